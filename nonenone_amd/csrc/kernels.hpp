// kernels.hpp — HIP kernels of the library and the Launch<NC> definitions.
// Included once per component count by nn_nc<NC>.hip.
#pragma once

#include "assembly.hpp"
#include "common.hpp"
#include "eos.hpp"
#include "flash.hpp"
#include "launch.hpp"
#include "phase_update.hpp"

namespace nn
{

// ------------------------------------------------------------ E1-E5 kernel
template <int NC>
__global__ void __launch_bounds__(kBlock, kWavesPerEu) eosPhaseKernel(
    const FluidParams *__restrict__ fp, int64_t n, const double *__restrict__ P,
    const double *__restrict__ T, const double *__restrict__ x, const int32_t *__restrict__ role,
    double *__restrict__ Z, double *__restrict__ lnphi, double *__restrict__ fug,
    int32_t *__restrict__ status)
{
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    const FluidParams &f = *fp;
    Err err;
    double xs[NC];
    for (int c = 0; c < NC; ++c) xs[c] = x[s * NC + c];
    const int r = role[s];
    double z = 1.0;
    double lp[NC];
    if (f.model == NN_MODEL_CPA)
    {
        CpaTemp<NC> ct;
        cpaTemperature<NC>(f, T[s], ct);
        cpaPhase<NC, double>(f, ct, P[s], T[s], xs, r != kGas, z, lp, err);
    }
    else
    {
        MixParams<NC, double> mp;
        mixingParameters<NC, double>(f, P[s], T[s], r, mp);
        PhaseMix<NC, double> mix;
        phaseMixing<NC, double>(xs, mp, mix);
        z = selectRoot(f, mix.A, mix.B, r != kGas, err);
        if (!err.bad())
            lnPhiAtRoot<NC, double>(f, mp, mix, z, lp, err);
    }
    Z[s] = z;
    for (int c = 0; c < NC; ++c)
    {
        const double phi = err.bad() ? 0.0 : ::exp(lp[c]);
        lnphi[s * NC + c] = err.bad() ? 0.0 : ::log(phi);
        fug[s * NC + c] = phi * P[s] * xs[c];
    }
    status[s] = err.code;
}

// ---------------------------------------------------------- flash kernels
template <int NC>
__device__ void storeFlash(const FlashResult<NC> &r, int64_t s, const nn_flash_out &out)
{
    out.presence[s] = r.presence;
    out.converged[s] = r.converged ? 1 : 0;
    out.iterations[s] = r.iterations;
    for (int p = 0; p < 3; ++p)
    {
        out.beta[s * 3 + p] = r.beta[p];
        out.saturation[s * 3 + p] = r.sat[p];
        out.Z[s * 3 + p] = r.Z[p];
        out.rho_m[s * 3 + p] = r.rho[p];
        for (int c = 0; c < NC; ++c)
            out.x[(s * 3 + p) * NC + c] = r.x[p][c];
    }
    for (int c = 0; c < NC; ++c)
    {
        out.k_go[s * NC + c] = r.kg[c];
        out.k_wo[s * NC + c] = r.kw[c];
    }
}

// One state of a flash batch.  Returns false when a kModeDefer run handed the
// state to the cooperative pass (nothing stored); in kModeCoop only lane 0 stores.
template <int NC, int Mode>
__device__ bool flashOne(const FluidParams &f, const FlashOptions &o, int64_t s,
                         const double *P, const double *T, const double *z,
                         const uint8_t *allowed, const double *seed, FlashHint h,
                         const nn_flash_out &out)
{
    Err err;
    Flash<NC, Mode> fl(f, o, err);
    double zs[NC];
    for (int c = 0; c < NC; ++c) zs[c] = z[s * NC + c];
    FlashResult<NC> r;
    if (allowed == nullptr)
    {
        if (h.hint)
        {
            fl.hint = h.hint[s] & 3;  // bits 2-3 only order the batch
            fl.hintLogK = h.logK ? h.logK + s * NC : nullptr;
        }
        r = fl.flash(P[s], T[s], zs);
    }
    else
    {
        const uint8_t a = allowed[s];
        if (a == kAllBits)
        {
            // flashRestricted with every phase allowed: the full flash followed
            // by the restricted post-processing (three_phase_flash.hpp:235-259).
            double zn[NC];
            for (int c = 0; c < NC; ++c) zn[c] = zs[c];
            fl.normalize(zn);
            if (!err.bad())
                r = fl.flash(P[s], T[s], zn);
            if (!err.bad() && r.converged)
            {
                fl.canonicalizeFinal(P[s], T[s], r, true);
                if (!err.bad() && r.converged && !fl.validatePhysical(P[s], T[s], zn, r))
                    r.converged = false;
            }
        }
        else if (seed == nullptr)
        {
            r = fl.flashRestricted(P[s], T[s], zs, a, nullptr);
        }
        else
        {
            double sd[3][NC];
            for (int p = 0; p < 3; ++p)
                for (int c = 0; c < NC; ++c) sd[p][c] = seed[(s * 3 + p) * NC + c];
            r = fl.flashRestricted(P[s], T[s], zs, a, sd);
        }
    }
    if (err.code == kErrDefer)
        return false;
    if (Mode == kModeCoop && waveLane() != 0)
        return true;
    storeFlash<NC>(r, s, out);
    int st = err.code;
    if (st == kErrNone)
        st = r.converged ? NN_STATUS_OK : NN_STATUS_NOT_CONVERGED;
    if (err.bad())
        out.converged[s] = 0;
    out.status[s] = st;
    return true;
}

// Defers item s of a lane-per-item pass to the cooperative pass.
__device__ inline void deferItem(WorkQueue *q, unsigned int *list, int64_t s)
{
    const unsigned int slot = atomicAdd(&q->deferred, 1u);
    list[slot] = static_cast<unsigned int>(s);
}

// Next deferred item for this wavefront (uniform across its lanes), or -1.
__device__ inline int64_t nextDeferred(WorkQueue *q, const unsigned int *list)
{
    unsigned int idx = 0;
    if (waveLane() == 0)
        idx = atomicAdd(&q->coopHead, 1u);
    idx = static_cast<unsigned int>(waveShfl(static_cast<int>(idx), 0));
    const unsigned int count = __atomic_load_n(&q->deferred, __ATOMIC_RELAXED);
    if (idx >= count)
        return -1;
    return static_cast<int64_t>(list[idx]);
}

// Drains the lane-per-item queue of a persistent pass, calling body(i) for
// each claimed position i < n.  q->claim == 0: every lane claims one item at a
// time, so a lane that finishes takes new work at once; q->claim != 0: lane 0
// claims a tile of 64 consecutive positions for its wavefront, so the lanes of a
// wavefront stay on neighbouring items of the processing order (less divergence
// when the order is sorted by condition).
template <class F>
__device__ inline void drainCounter(WorkQueue *q, unsigned int *head, int64_t n, F &&body)
{
    if (q->claim)
    {
        for (;;)
        {
            unsigned int base = 0;
            if (waveLane() == 0)
                base = atomicAdd(head, unsigned(kBlock));
            base = static_cast<unsigned int>(waveShfl(static_cast<int>(base), 0));
            if (int64_t(base) >= n)
                break;
            const int64_t i = int64_t(base) + waveLane();
            if (i < n)
                body(i);
        }
    }
    else
    {
        for (;;)
        {
            const int64_t i = int64_t(atomicAdd(head, 1u));
            if (i >= n)
                break;
            body(i);
        }
    }
}

template <class F>
__device__ inline void drainQueue(WorkQueue *q, int64_t n, F &&body)
{
    drainCounter(q, &q->head, n, static_cast<F &&>(body));
}

template <int NC>
__global__ void __launch_bounds__(kBlock, kWavesPerEu) flashKernel(
    const FluidParams *__restrict__ fp, const FlashOptions *__restrict__ op, int64_t n,
    const double *__restrict__ P, const double *__restrict__ T, const double *__restrict__ z,
    const uint8_t *__restrict__ allowed, const double *__restrict__ seed, FlashHint hint,
    nn_flash_out out, WorkQueue *__restrict__ q, unsigned int *__restrict__ list,
    const unsigned int *__restrict__ order)
{
    // Persistent work queue: work-items claim states until none are left, in
    // `order` (states of similar conditions next to each other) when given.
    drainQueue(q, n, [&](int64_t i) {
        const int64_t s = order ? int64_t(order[i]) : i;
        if (q->deferAll ||
            !flashOne<NC, kModeDefer>(*fp, *op, s, P, T, z, allowed, seed, hint, out))
            deferItem(q, list, s);
    });
}

// The baseline form's first kernel: one work-item per state in index order (no
// persistent queue, no sort); a state past the lane budget goes to the same
// cooperative pass as in the scheduled form, so both forms compute every state
// by the same numerics.
template <int NC>
__global__ void __launch_bounds__(kBlock, kWavesPerEu) flashPlainKernel(
    const FluidParams *__restrict__ fp, const FlashOptions *__restrict__ op, int64_t n,
    const double *__restrict__ P, const double *__restrict__ T, const double *__restrict__ z,
    const uint8_t *__restrict__ allowed, const double *__restrict__ seed, FlashHint hint,
    nn_flash_out out, WorkQueue *__restrict__ q, unsigned int *__restrict__ list)
{
    const int64_t s = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (s >= n)
        return;
    if (q->deferAll || !flashOne<NC, kModeDefer>(*fp, *op, s, P, T, z, allowed, seed, hint, out))
        deferItem(q, list, s);
}

// Sort keys that put states of similar (T, P, z) next to each other, so that
// the 64 states of a wavefront tend to follow the same iteration paths.
__device__ inline unsigned int quantize(double v, double lo, double hi, int bits)
{
    const double top = double((1u << bits) - 1u);
    double t = (v - lo) / (hi - lo);
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    return static_cast<unsigned int>(t * top);
}

template <int NC>
__global__ void flashKeyKernel(int64_t n, const double *__restrict__ P,
                               const double *__restrict__ T, const double *__restrict__ z,
                               const uint8_t *__restrict__ hint, unsigned int *__restrict__ keys,
                               unsigned int *__restrict__ idx)
{
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    const double lp = ::log(P[s] > 1.0 ? P[s] : 1.0);
    unsigned int k = quantize(T[s], 250.0, 750.0, 10) << 22;
    k |= quantize(lp, 11.5, 18.5, 10) << 12;
    k |= quantize(z[s * NC], 0.0, 1.0, 6) << 6;
    k |= quantize(z[s * NC + 1], 0.0, 1.0, 6);
    if (hint)  // learned hint first: states on the same solver path / of the same
               // predicted phase count share wavefronts
        k = (unsigned int)(hint[s] & 15u) << 28 | (k >> 4);
    keys[s] = k;
    idx[s] = static_cast<unsigned int>(s);
}

// One wavefront per deferred state (blocks of 64 = one wave).
template <int NC>
__global__ void __launch_bounds__(kBlock, kWavesPerEu) flashCoopKernel(
    const FluidParams *__restrict__ fp, const FlashOptions *__restrict__ op,
    const double *__restrict__ P, const double *__restrict__ T, const double *__restrict__ z,
    const uint8_t *__restrict__ allowed, const double *__restrict__ seed, FlashHint hint,
    nn_flash_out out, WorkQueue *__restrict__ q, const unsigned int *__restrict__ list)
{
    for (;;)
    {
        const int64_t s = nextDeferred(q, list);
        if (s < 0)
            break;
        flashOne<NC, kModeCoop>(*fp, *op, s, P, T, z, allowed, seed, hint, out);
    }
}

template <int NC>
__global__ void __launch_bounds__(kBlock, kWavesPerEu) stabilityKernel(
    const FluidParams *__restrict__ fp, const FlashOptions *__restrict__ op, int64_t n,
    const double *__restrict__ P, const double *__restrict__ T, const double *__restrict__ z,
    const uint8_t *__restrict__ active, const double *__restrict__ xin, uint8_t *stable,
    uint8_t *valid, uint8_t *mask, double *trial, double *incipient, int32_t *status)
{
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    Err err;
    Flash<NC, kModePlain> fl(*fp, *op, err);
    double zs[NC], xs[3][NC];
    for (int c = 0; c < NC; ++c) zs[c] = z[s * NC + c];
    for (int p = 0; p < 3; ++p)
        for (int c = 0; c < NC; ++c) xs[p][c] = xin[(s * 3 + p) * NC + c];
    const StabilityResult<NC> r = fl.stabilityTest(P[s], T[s], zs, active[s], xs);
    stable[s] = r.stable;
    valid[s] = r.valid;
    uint8_t m = 0;
    for (int p = 0; p < 3; ++p)
    {
        if (r.unstable[p]) m |= uint8_t(1u << p);
        trial[s * 3 + p] = r.trialSum[p];
        for (int c = 0; c < NC; ++c) incipient[(s * 3 + p) * NC + c] = r.incipient[p][c];
    }
    mask[s] = m;
    status[s] = err.code;
}

// ---------------------------------------------------- phase-update kernel
template <int NC, bool Well>
__global__ void __launch_bounds__(kBlock, kWavesPerEu) phaseUpdateKernel(
    const FluidParams *__restrict__ fp, const FlashOptions *__restrict__ op, int64_t n,
    const double *__restrict__ T, double *__restrict__ primary, double *__restrict__ phaseState,
    int32_t *__restrict__ status, uint8_t *__restrict__ flags, WorkQueue *__restrict__ q,
    unsigned int *__restrict__ list, const unsigned int *__restrict__ order,
    unsigned int *__restrict__ pendingList, StabRecord<NC> *__restrict__ records)
{
    drainQueue(q, n, [&](int64_t i) {
        const int64_t s = order ? int64_t(order[i]) : i;
        if (q->deferAll)
        {
            deferItem(q, list, s);
            return;
        }
        const int r = phaseUpdateCell<NC, Well, kModeDefer>(*fp, *op, T[s], primary, phaseState, s,
                                                            status, flags, kStabRecord, records);
        if (r == kItemDeferred)
            deferItem(q, list, s);
        else if (r == kItemPending)
            pendingList[atomicAdd(&q->pending, 1u)] = static_cast<unsigned int>(s);
    });
}

// Lockstep stability trials of the recorded cells: each lane runs every seed
// trial of its cell as a state machine (Flash::trialStep) whose one expensive
// step, the residual evaluation, all lanes of the wavefront take together; a
// lane whose cell is complete claims the next recorded cell.
template <int NC>
__global__ void __launch_bounds__(kBlock, kWavesPerEu) stabilityTrialKernel(
    const FluidParams *__restrict__ fp, const FlashOptions *__restrict__ op,
    WorkQueue *__restrict__ q, const unsigned int *__restrict__ pendingList,
    StabRecord<NC> *__restrict__ records)
{
    using F = Flash<NC, kModePlain>;
    const unsigned int count = q->pending;
    Err err;
    F fl(*fp, *op, err);
    typename F::TrialMachine m;
    StabRecord<NC> *rec = nullptr;
    for (;;)
    {
        while (rec == nullptr)
        {
            const unsigned int slot = atomicAdd(&q->trialHead, 1u);
            if (slot >= count)
                break;
            rec = records + pendingList[slot];
            err.code = kErrNone;
            if (fl.trialInit(*rec, m))
                rec = nullptr;  // nothing to evaluate (or an error): complete already
        }
        if (rec == nullptr)
            break;
        double x[NC];
        fl.trialPoint(m, x);
        typename F::StabResidual res;
        fl.evaluateStabilityResidual(m.cand, rec->P, rec->T, rec->logRef, x, res);
        typename F::Trial t;
        if (fl.trialStep(*rec, m, res, t) && fl.trialFold(*rec, m, t))
            rec = nullptr;
    }
}

// Finishing pass of the recorded cells: the phase update again, with the
// stability test replayed from the lockstep trials' results.
template <int NC, bool Well>
__global__ void __launch_bounds__(kBlock, kWavesPerEu) phaseUpdateReplayKernel(
    const FluidParams *__restrict__ fp, const FlashOptions *__restrict__ op, int64_t n,
    const double *__restrict__ T, double *__restrict__ primary, double *__restrict__ phaseState,
    int32_t *__restrict__ status, uint8_t *__restrict__ flags, WorkQueue *__restrict__ q,
    unsigned int *__restrict__ list, const unsigned int *__restrict__ pendingList,
    StabRecord<NC> *__restrict__ records)
{
    const int64_t count = int64_t(q->pending);
    unsigned int *reflashList = const_cast<unsigned int *>(pendingList) + n;
    drainCounter(q, &q->replayHead, count, [&](int64_t i) {
        const int64_t s = int64_t(pendingList[i]);
        const int r = phaseUpdateCell<NC, Well, kModeDefer>(
            *fp, *op, T[s], primary, phaseState, s, status, flags, kStabReplayHoldReflash, records);
        if (r == kItemDeferred)
            deferItem(q, list, s);
        else if (r == kItemReflash)
            reflashList[atomicAdd(&q->reflash, 1u)] = static_cast<unsigned int>(s);
    });
}

// Re-flash pass: the replayed cells whose missing phase turned out unstable,
// together, so that the wavefronts of this costly unrestricted flash are full.
template <int NC, bool Well>
__global__ void __launch_bounds__(kBlock, kWavesPerEu) phaseUpdateReflashKernel(
    const FluidParams *__restrict__ fp, const FlashOptions *__restrict__ op, int64_t n,
    const double *__restrict__ T, double *__restrict__ primary, double *__restrict__ phaseState,
    int32_t *__restrict__ status, uint8_t *__restrict__ flags, WorkQueue *__restrict__ q,
    unsigned int *__restrict__ list, const unsigned int *__restrict__ pendingList,
    StabRecord<NC> *__restrict__ records, const unsigned int *__restrict__ order)
{
    const int64_t count = int64_t(q->reflash);
    const unsigned int *reflashList = order ? order : pendingList + n;
    drainCounter(q, &q->reflashHead, count, [&](int64_t i) {
        const int64_t s = int64_t(reflashList[i]);
        const int r = phaseUpdateCell<NC, Well, kModeDefer>(*fp, *op, T[s], primary, phaseState, s,
                                                            status, flags, kStabReplay, records);
        if (r == kItemDeferred)
            deferItem(q, list, s);
    });
}

// Phase-update sort key: presence, whether an active saturation sits at or
// below the removal probe, then (T, P, water fraction of the oil slot).
template <int NC, bool Well>
__device__ inline unsigned int phaseUpdateKey(int64_t s, const double *__restrict__ T,
                                              const double *__restrict__ primary,
                                              const double *__restrict__ phaseState)
{
    using L = Layout<NC, Well>;
    const double flag = phaseState[s * L::NPS + L::flag];
    const unsigned int presence = static_cast<unsigned int>(flag > 0.0 && flag < 7.5 ? flag + 0.5 : 0.0) & 7u;
    unsigned int removed = 0;
    for (int p = 0; p < 3; ++p)
        if ((presence & (1u << p)) && primary[s * L::NP + L::satBegin + p] <= kProbeSaturation)
            removed = 1;
    const double P = primary[s * L::NP + L::pressure];
    const double lp = ::log(P > 1.0 ? P : 1.0);
    unsigned int k = presence << 29;
    k |= removed << 28;
    k |= quantize(T[s], 250.0, 750.0, 10) << 18;
    k |= quantize(lp, 11.5, 18.5, 10) << 8;
    k |= quantize(primary[s * L::NP + L::compBegin(0)], 0.0, 1.0, 8);
    return k;
}

template <int NC, bool Well>
__global__ void phaseUpdateKeyKernel(int64_t n, const double *__restrict__ T,
                                     const double *__restrict__ primary,
                                     const double *__restrict__ phaseState,
                                     unsigned int *__restrict__ keys, unsigned int *__restrict__ idx)
{
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    keys[s] = phaseUpdateKey<NC, Well>(s, T, primary, phaseState);
    idx[s] = static_cast<unsigned int>(s);
}

// Keys of the re-flash pass's cells (the held count is read on the device):
// position i < count holds reflashList[i] under its phase-update key (at most
// 0xFFFFFFFE), every later position sorts last.
template <int NC, bool Well>
__global__ void reflashKeyKernel(int64_t n, const WorkQueue *__restrict__ q,
                                 const unsigned int *__restrict__ reflashList,
                                 const double *__restrict__ T, const double *__restrict__ primary,
                                 const double *__restrict__ phaseState,
                                 unsigned int *__restrict__ keys, unsigned int *__restrict__ idx)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    if (i >= int64_t(q->reflash))
    {
        keys[i] = 0xFFFFFFFFu;
        idx[i] = 0u;
        return;
    }
    const unsigned int s = reflashList[i];
    const unsigned int k = phaseUpdateKey<NC, Well>(int64_t(s), T, primary, phaseState);
    keys[i] = k < 0xFFFFFFFFu ? k : 0xFFFFFFFEu;
    idx[i] = s;
}

template <int NC, bool Well>
__global__ void __launch_bounds__(kBlock, kWavesPerEu) phaseUpdateCoopKernel(
    const FluidParams *__restrict__ fp, const FlashOptions *__restrict__ op,
    const double *__restrict__ T, double *__restrict__ primary, double *__restrict__ phaseState,
    int32_t *__restrict__ status, uint8_t *__restrict__ flags, WorkQueue *__restrict__ q,
    const unsigned int *__restrict__ list)
{
    for (;;)
    {
        const int64_t s = nextDeferred(q, list);
        if (s < 0)
            break;
        (void)phaseUpdateCell<NC, Well, kModeCoop>(*fp, *op, T[s], primary, phaseState, s, status,
                                             flags);
    }
}

// ------------------------------------------------------- assembly kernels


template <int NC, bool Well>
__device__ bool loadCell(const FluidParams &f, const AssemblyView &v, const double *primary,
                         const double *phaseState, int64_t c, uint8_t &presence)
{
    using L = Layout<NC, Well>;
    const double flag = phaseState[c * L::NPS + L::flag];
    const int raw = int(flag + (flag >= 0.0 ? 0.5 : -0.5));
    if (raw <= 0 || raw > 7)
        return false;
    presence = uint8_t(raw);
    return true;
}

// Value-only property pass: the record a neighbour's face flux reads.
template <int NC, bool Well>
__global__ void __launch_bounds__(256, kWavesPerEu) cellRecordKernel(
    const FluidParams *__restrict__ fp, AssemblyView v, const double *__restrict__ primary,
    const double *__restrict__ phaseState, double *__restrict__ accum, int32_t *__restrict__ status)
{
    using L = Layout<NC, Well>;
    using S = Dual<0>;
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= v.ncell)
        return;
    const FluidParams &f = *fp;
    uint8_t presence = 0;
    Err err;
    CellStateT<NC, S> st;
    CellPropsT<NC, S> pr;
    double roots[9] = {-1.0, 0.0, 0.0, -1.0, 0.0, 0.0, -1.0, 0.0, 0.0};
    if (!loadCell<NC, Well>(f, v, primary, phaseState, c, presence))
        err.raise(kErrThermo);
    else
    {
        decodeState<NC, Well, S>(primary + c * L::NP, presence, st);
        evaluateProperties<NC, S>(f, st, v.porosity[c], v.T[c], pr, err,
                                  v.roots ? roots : nullptr);
    }
    if (status)
        status[c] = err.code;
    const double nan = __builtin_nan("");
    const int64_t n = v.ncell;
    if (v.roots)
        for (int k = 0; k < 9; ++k) v.roots[k * n + c] = err.bad() ? -1.0 : roots[k];
    if (v.record)
    {
        double *r = v.record;
        const bool ok = !err.bad();
        r[0 * n + c] = ok ? st.p.v : nan;
        for (int q = 0; q < 3; ++q)
        {
            r[(1 + q) * n + c] = ok ? pr.rho[q].v : nan;
            r[(4 + q) * n + c] = ok ? pr.sat[q].v : nan;
            r[(7 + q) * n + c] = ok ? pr.mob[q].v : nan;
            for (int k = 0; k < NC; ++k)
                r[(10 + q * NC + k) * n + c] = ok ? pr.w[q][k].v : nan;
        }
    }
    if (accum)
        for (int k = 0; k < NC; ++k)
            accum[c * NC + k] = err.bad() ? nan : componentAccumulation<NC, S>(pr, k).v;
}

// Derivative chunks of the assembly: chunk k evaluates a cell with Dual<CH> seeded
// on primary variables [k CH, k CH + CH).
template <int NC, bool Well>
constexpr int assemblyChunkCount()
{
    return (Layout<NC, Well>::NP + kAssemblyChunk - 1) / kAssemblyChunk;
}

// Residual + Jacobian columns of one cell (residual.inc:59-120, jacobian.inc:20-237),
// one work-item per (cell, derivative chunk), the chunk index fastest: a
// derivative component is computed by the same operations on the same values
// whatever the other components hold, so every column is bit-identical to a
// Dual<N_p> evaluation, while a lane keeps CH + 1 doubles per AD quantity instead
// of N_p + 1.  Neighbouring lanes hold neighbouring column pairs of the same
// cell, so each block row is stored by adjacent lanes of one store instruction.
// Chunk 0 also computes the residual row, the upwind decisions, the well sources
// and the status; the PR-CPA density roots with their correction slopes come from
// the record pass and are reused by every chunk.  Without a Jacobian (bsr null)
// the launch has one work-item per cell.
template <int NC, bool Well>
__global__ void __launch_bounds__(64, kWavesPerEu) assembleKernel(
    const FluidParams *__restrict__ fp, AssemblyView v, double dt, double fugScale,
    const double *__restrict__ primary, const double *__restrict__ phaseState,
    const double *__restrict__ accumPrev, double *__restrict__ residual, double *__restrict__ bsr,
    uint8_t *__restrict__ decisions, int32_t *__restrict__ status)
{
    using L = Layout<NC, Well>;
    constexpr int NP = L::NP;
    constexpr int CH = kAssemblyChunk;
    constexpr int NCHUNK = (NP + CH - 1) / CH;
    using S = Dual<CH>;
    const int nchunk = bsr ? NCHUNK : 1;
    const int64_t item = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t c = item / nchunk;
    const int chunk = int(item - c * nchunk);
    if (c >= v.ncell)
        return;
    const FluidParams &f = *fp;
    uint8_t presence = 0;
    const bool loaded = loadCell<NC, Well>(f, v, primary, phaseState, c, presence);
    // row scale of each equation of this cell (natural_petsc_runtime_scaling.inc:26-89)
    double rowScale[NP];
    for (int r = 0; r < NP; ++r) rowScale[r] = v.rowScale[r];
    if constexpr (Well)
        if (v.wellRowScale)
            rowScale[NP - 1] = v.wellRowScale[c];
    double *self = bsr ? bsr + v.selfSlot[c] * NP * NP : nullptr;
    const int64_t n = v.ncell;
    const double *rec = v.record;
    const double vIn = v.volume[c];
    const int64_t kBegin = v.connOffsets[c], kEnd = v.connOffsets[c + 1];
    // density roots and correction slopes of the record pass (PR-CPA), which
    // evaluated the same cell values
    double roots[9] = {-1.0, 0.0, 0.0, -1.0, 0.0, 0.0, -1.0, 0.0, 0.0};
    if (v.roots)
        for (int k = 0; k < 9; ++k) roots[k] = v.roots[k * v.ncell + c];
    int32_t wellStatus = NN_STATUS_OK;   // a perforation source the reference rejects
    {
        const int col0 = chunk * CH;
        Err err;
        CellStateT<NC, S> st;
        CellPropsT<NC, S> pr;
        if (!loaded)
            err.raise(kErrThermo);
        else
        {
            decodeState<NC, Well, S>(primary + c * NP, presence, st, col0);
            evaluateProperties<NC, S>(f, st, v.porosity[c], v.T[c], pr, err, roots);
        }
        if (err.bad())
        {
            // the values are the same in every chunk, so every chunk of the cell
            // gets here: chunk 0 writes the status and the residual, each chunk
            // its columns of the cell's blocks
            const double nan = __builtin_nan("");
            if (chunk == 0)
            {
                if (status) status[c] = err.code;
                if (residual)
                    for (int r = 0; r < NP; ++r) residual[c * NP + r] = nan;
            }
            if (bsr)
                for (int q = 0; q < CH && col0 + q < NP; ++q)
                    for (int r = 0; r < NP; ++r)
                    {
                        self[r * NP + col0 + q] = nan;
                        for (int64_t k = kBegin; k < kEnd; ++k)
                            bsr[v.recipSlot[k] * NP * NP + r * NP + col0 + q] = nan;
                    }
            return;
        }
        S res[NP];
        localResidual<NC, Well, S>(st, pr, accumPrev + c * NC, v.volume[c], dt, fugScale, res);

        // jacobian.inc:262-268: (derivative * rowScale) * variableScale; the mass rows
        // gather the face-flux derivatives below in the order of the connections
        double selfMass[NC][CH];
        if (self)
        {
            for (int q = 0; q < CH; ++q)
            {
                const int col = col0 + q;
                if (col >= NP)
                    break;
                for (int r = 0; r < NC; ++r)
                    selfMass[r][q] = res[r].d[q] * rowScale[r] * v.varScale[col];
                for (int r = NC; r < NP; ++r)
                    self[r * NP + col] = res[r].d[q] * rowScale[r] * v.varScale[col];
            }
        }
        double outward[NC];
        for (int k = 0; k < NC; ++k) outward[k] = 0.0;
        for (int64_t k = kBegin; k < kEnd; ++k)
        {
            const int64_t nb = v.connNeighbor[k];
            const double trans = v.connTrans[k];
            const double gdz = v.connGdz[k];
            const double pEx = rec[nb];
            const double vEx = v.volume[nb];
            bool upIn[3];
            double faceFlux[NC];
            for (int i = 0; i < NC; ++i) faceFlux[i] = 0.0;
            // scalar face flux (face_flux.hpp:330-430): the upwind decisions, and in
            // the first chunk the residual's flux
            for (int q = 0; q < 3; ++q)
            {
                const double rhoIn = pr.rho[q].v, rhoEx = rec[(1 + q) * n + nb];
                const double satIn = pr.sat[q].v, satEx = rec[(4 + q) * n + nb];
                const double avg = 0.5 * (rhoIn * satIn + rhoEx * satEx) /
                                   maxd(0.5 * (satIn + satEx), 1.0e-8);
                const double dPhi = pEx + avg * gdz - st.p.v;
                upIn[q] = interiorUpwind(dPhi, vIn, vEx, c, nb);
                if (chunk != 0)
                    continue;
                if (decisions)
                    decisions[k * 3 + q] = upIn[q] ? 0 : 1;
                const double mob = upIn[q] ? pr.mob[q].v : rec[(7 + q) * n + nb];
                const double dens = upIn[q] ? rhoIn : rhoEx;
                const double total = -trans * dPhi * mob * dens;
                for (int i = 0; i < NC; ++i)
                {
                    const double wv = upIn[q] ? pr.w[q][i].v : rec[(10 + q * NC + i) * n + nb];
                    faceFlux[i] += total * wv;
                }
            }
            if (chunk == 0)
                for (int i = 0; i < NC; ++i) outward[i] += faceFlux[i];
            if (!self)
                continue;
            // AD face flux with the recorded decision (face_flux.hpp:174-321)
            S flux[NC];
            for (int i = 0; i < NC; ++i) flux[i] = S(0.0);
            for (int q = 0; q < 3; ++q)
            {
                const double rhoEx = rec[(1 + q) * n + nb];
                const double satEx = rec[(4 + q) * n + nb];
                const S satAvg = 0.5 * (pr.sat[q] + satEx);
                const S avg = 0.5 * (pr.rho[q] * pr.sat[q] + rhoEx * satEx) / maxd(satAvg.v, 1.0e-8);
                const S dPhi = pEx + avg * gdz - st.p;
                if (upIn[q])
                {
                    const S darcy = -trans * dPhi * pr.mob[q];
                    const S total = darcy * pr.rho[q];
                    for (int i = 0; i < NC; ++i) flux[i] += total * pr.w[q][i];
                }
                else
                {
                    const double mobEx = rec[(7 + q) * n + nb];
                    const S darcy = -trans * dPhi * mobEx;
                    const S total = darcy * rhoEx;
                    for (int i = 0; i < NC; ++i)
                        flux[i] += total * rec[(10 + q * NC + i) * n + nb];
                }
            }
            // jacobian.inc:196-213: mass rows carry one scale in every cell
            double *rb = bsr + v.recipSlot[k] * NP * NP;
            for (int q = 0; q < CH; ++q)
            {
                const int col = col0 + q;
                if (col >= NP)
                    break;
                for (int i = 0; i < NC; ++i)
                {
                    selfMass[i][q] += flux[i].d[q] * rowScale[i] * v.varScale[col];
                    rb[i * NP + col] = -flux[i].d[q] * v.rowScale[i] * v.varScale[col];
                }
            }
            for (int q = 0; q < CH && col0 + q < NP; ++q)
                for (int r = NC; r < NP; ++r) rb[r * NP + col0 + q] = 0.0;
        }
        if (self)
            for (int q = 0; q < CH; ++q)
            {
                const int col = col0 + q;
                if (col >= NP)
                    break;
                for (int r = 0; r < NC; ++r) self[r * NP + col] = selfMass[r][q];
            }
        if (chunk == 0 && residual)
        {
            // the perforations' sources enter as outward flux - source
            // (addCellConservationTransportAndWell, cell_residual.hpp:435-447), summed in
            // well then perforation order (natural_petsc_runtime_wells.inc:93-165)
            double wellSource[NC];
            bool perforated = false;
            if constexpr (Well)
                if (v.wellCellOffsets && v.wellCellOffsets[c + 1] > v.wellCellOffsets[c])
                {
                    perforated = true;
                    for (int i = 0; i < NC; ++i) wellSource[i] = 0.0;
                    for (int64_t k = v.wellCellOffsets[c]; k < v.wellCellOffsets[c + 1]; ++k)
                    {
                        const int32_t perf = v.wellCellPerfs[k];
                        const DeviceWell &wl = v.wells[v.perfWell[perf]];
                        const S bhp(primary[wl.bhpCell * NP + L::well]);
                        S src[NC];
                        perforationSource<NC, S>(wl, v.perfIndex[perf], v.surfaceDensity, st.p,
                                                 bhp, pr.rho, pr.mob, pr.w, src, err);
                        for (int i = 0; i < NC; ++i) wellSource[i] += src[i].v;
                    }
                    if (err.bad())
                        wellStatus = err.code;
                }
            // natural_petsc_runtime_common.inc:345-356
            for (int i = 0; i < NC; ++i)
                residual[c * NP + i] = perforated
                    ? (res[i].v + (outward[i] - wellSource[i])) * rowScale[i]
                    : (res[i].v + outward[i]) * rowScale[i];
            for (int r = NC; r < NP; ++r)
                residual[c * NP + r] = res[r].v * rowScale[r];
        }
    }
    if (status && chunk == 0)
        status[c] = wellStatus;
}

// insertWellJacobian_ (natural_petsc_runtime_jacobian.inc:301-457) for BHP-controlled
// wells: one thread per perforated cell adds, perforation by perforation, the
// derivatives of -source to its mass rows: the cell's own columns into its diagonal
// block, the well-pressure column into block (cell, bhp cell).  Derivative chunks
// as in assembleKernel.
template <int NC, bool Well>
__global__ void __launch_bounds__(64, kWavesPerEu) wellJacobianKernel(
    const FluidParams *__restrict__ fp, AssemblyView v, const int64_t *__restrict__ cells,
    int64_t ncells, const double *__restrict__ primary, const double *__restrict__ phaseState,
    double *__restrict__ bsr, int32_t *__restrict__ status)
{
    using L = Layout<NC, Well>;
    constexpr int NP = L::NP;
    constexpr int CH = kAssemblyChunk;
    constexpr int NCHUNK = (NP + CH - 1) / CH;
    using S = Dual<CH>;
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= ncells)
        return;
    const int64_t c = cells[t];
    const FluidParams &f = *fp;
    uint8_t presence = 0;
    if (!loadCell<NC, Well>(f, v, primary, phaseState, c, presence))
        return;   // the assembly pass has reported this cell
    double roots[9] = {-1.0, 0.0, 0.0, -1.0, 0.0, 0.0, -1.0, 0.0, 0.0};
    if (v.roots)
        for (int k = 0; k < 9; ++k) roots[k] = v.roots[k * v.ncell + c];
    double *self = bsr + v.selfSlot[c] * NP * NP;
    for (int chunk = 0; chunk < NCHUNK; ++chunk)
    {
        const int col0 = chunk * CH;
        Err err;
        CellStateT<NC, S> st;
        CellPropsT<NC, S> pr;
        decodeState<NC, Well, S>(primary + c * NP, presence, st, col0);
        evaluateProperties<NC, S>(f, st, v.porosity[c], v.T[c], pr, err, roots);
        if (err.bad())
            return;
        for (int64_t k = v.wellCellOffsets[c]; k < v.wellCellOffsets[c + 1]; ++k)
        {
            const int32_t perf = v.wellCellPerfs[k];
            const DeviceWell &wl = v.wells[v.perfWell[perf]];
            // the bottom-hole pressure: the bhp cell's well-pressure unknown
            S bhp(primary[wl.bhpCell * NP + L::well]);
            if (L::well >= col0 && L::well < col0 + CH)
                bhp.d[L::well - col0] = 1.0;
            S src[NC];
            perforationSource<NC, S>(wl, v.perfIndex[perf], v.surfaceDensity, st.p, bhp, pr.rho,
                                     pr.mob, pr.w, src, err);
            if (err.bad())
            {
                if (status) status[c] = err.code;
                return;
            }
            double *coupling = wl.bhpCell == c ? self : bsr + v.perfBhpSlot[perf] * NP * NP;
            for (int q = 0; q < CH; ++q)
            {
                const int col = col0 + q;
                if (col >= NP)
                    break;
                double *block = col == L::well ? coupling : self;
                for (int i = 0; i < NC; ++i)
                {
                    const S eq = -src[i];
                    block[i * NP + col] += eq.d[q] * v.rowScale[i] * v.varScale[col];
                }
            }
        }
    }
}

// overwriteWellControlResiduals_ (natural_petsc_runtime_wells.inc:167-243), BHP
// control: R_w = p_bhp - target on the bhp cell's well-control row, in well order.
template <int NC, bool Well>
__global__ void wellControlKernel(AssemblyView v, int nwells, const double *__restrict__ primary,
                                  double *__restrict__ residual)
{
    using L = Layout<NC, Well>;
    if (blockIdx.x != 0 || threadIdx.x != 0)
        return;
    for (int w = 0; w < nwells; ++w)
    {
        const DeviceWell &wl = v.wells[w];
        const int64_t r = wl.bhpCell;
        const double scale = v.wellRowScale ? v.wellRowScale[r] : v.rowScale[L::NP - 1];
        residual[r * L::NP + L::well] = (primary[r * L::NP + L::well] - wl.target) * scale;
    }
}
// ------------------------------------------------------------ launchers
template <int NC>
hipError_t Launch<NC>::eos(unsigned grid, hipStream_t st, const FluidParams *fp, int64_t n,
                           const double *P, const double *T, const double *x,
                           const int32_t *role, double *Z, double *lnphi, double *fug,
                           int32_t *status)
{
    eosPhaseKernel<NC><<<grid, kBlock, 0, st>>>(fp, n, P, T, x, role, Z, lnphi, fug, status);
    return hipGetLastError();
}

template <int NC>
hipError_t Launch<NC>::flash(unsigned grid, unsigned coopGrid, hipStream_t st,
                             const FluidParams *fp, const FlashOptions *op, int64_t n,
                             const double *P, const double *T, const double *z,
                             const uint8_t *allowed, const double *seed, FlashHint hint,
                             nn_flash_out out, WorkQueue *q, unsigned int *list,
                             const unsigned int *order, bool plain, hipEvent_t splitA,
                             hipEvent_t splitB)
{
    if (plain)
        flashPlainKernel<NC><<<grid, kBlock, 0, st>>>(fp, op, n, P, T, z, allowed, seed, hint,
                                                      out, q, list);
    else
        flashKernel<NC><<<grid, kBlock, 0, st>>>(fp, op, n, P, T, z, allowed, seed, hint, out, q,
                                                 list, order);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    if (splitA && (e = hipEventRecord(splitA, st)) != hipSuccess)
        return e;
    if (splitB && (e = hipEventRecord(splitB, st)) != hipSuccess)
        return e;
    flashCoopKernel<NC><<<coopGrid, kBlock, 0, st>>>(fp, op, P, T, z, allowed, seed, hint, out,
                                                     q, list);
    return hipGetLastError();
}

template <int NC>
hipError_t Launch<NC>::flashKeys(hipStream_t st, int64_t n, const double *P, const double *T,
                                 const double *z, const uint8_t *hint, unsigned int *keys,
                                 unsigned int *idx)
{
    const unsigned grid = static_cast<unsigned>((n + 255) / 256);
    flashKeyKernel<NC><<<grid, 256, 0, st>>>(n, P, T, z, hint, keys, idx);
    return hipGetLastError();
}

template <int NC>
hipError_t Launch<NC>::phaseUpdateKeys(bool well, hipStream_t st, int64_t n, const double *T,
                                       const double *primary, const double *phaseState,
                                       unsigned int *keys, unsigned int *idx)
{
    const unsigned grid = static_cast<unsigned>((n + 255) / 256);
    if (well)
        phaseUpdateKeyKernel<NC, true><<<grid, 256, 0, st>>>(n, T, primary, phaseState, keys, idx);
    else
        phaseUpdateKeyKernel<NC, false><<<grid, 256, 0, st>>>(n, T, primary, phaseState, keys, idx);
    return hipGetLastError();
}

template <int NC>
hipError_t Launch<NC>::stability(unsigned grid, hipStream_t st, const FluidParams *fp,
                                 const FlashOptions *op, int64_t n, const double *P,
                                 const double *T, const double *z, const uint8_t *active,
                                 const double *x, uint8_t *stable, uint8_t *valid,
                                 uint8_t *mask, double *trial, double *incipient,
                                 int32_t *status)
{
    stabilityKernel<NC><<<grid, kBlock, 0, st>>>(fp, op, n, P, T, z, active, x, stable, valid,
                                                 mask, trial, incipient, status);
    return hipGetLastError();
}

template <int NC>
hipError_t Launch<NC>::phaseUpdate(bool well, unsigned grid, unsigned coopGrid, hipStream_t st,
                                   const FluidParams *fp, const FlashOptions *op, int64_t n,
                                   const double *T, double *primary, double *phaseState,
                                   int32_t *status, uint8_t *flags, WorkQueue *q,
                                   unsigned int *list, const unsigned int *order,
                                   unsigned int *pendingList, void *records,
                                   const ReflashOrder *reorder, hipEvent_t splitA,
                                   hipEvent_t splitB)
{
    StabRecord<NC> *recs = static_cast<StabRecord<NC> *>(records);
    if (!pendingList)
        recs = nullptr;
    if (well)
        phaseUpdateKernel<NC, true><<<grid, kBlock, 0, st>>>(fp, op, n, T, primary, phaseState,
                                                             status, flags, q, list, order,
                                                             pendingList, recs);
    else
        phaseUpdateKernel<NC, false><<<grid, kBlock, 0, st>>>(fp, op, n, T, primary, phaseState,
                                                              status, flags, q, list, order,
                                                              pendingList, recs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    if (recs)
    {
        stabilityTrialKernel<NC><<<grid, kBlock, 0, st>>>(fp, op, q, pendingList, recs);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
        if (well)
            phaseUpdateReplayKernel<NC, true><<<grid, kBlock, 0, st>>>(
                fp, op, n, T, primary, phaseState, status, flags, q, list, pendingList, recs);
        else
            phaseUpdateReplayKernel<NC, false><<<grid, kBlock, 0, st>>>(
                fp, op, n, T, primary, phaseState, status, flags, q, list, pendingList, recs);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
        const unsigned int *reflashOrder = nullptr;
        if (reorder)
        {
            const unsigned keyGrid = static_cast<unsigned>((n + 255) / 256);
            if (well)
                reflashKeyKernel<NC, true><<<keyGrid, 256, 0, st>>>(
                    n, q, pendingList + n, T, primary, phaseState, reorder->keys, reorder->idx);
            else
                reflashKeyKernel<NC, false><<<keyGrid, 256, 0, st>>>(
                    n, q, pendingList + n, T, primary, phaseState, reorder->keys, reorder->idx);
            if ((e = hipGetLastError()) != hipSuccess)
                return e;
            if ((e = reorder->sort(reorder->arg, st, &reflashOrder)) != hipSuccess)
                return e;
        }
        if (well)
            phaseUpdateReflashKernel<NC, true><<<grid, kBlock, 0, st>>>(
                fp, op, n, T, primary, phaseState, status, flags, q, list, pendingList, recs,
                reflashOrder);
        else
            phaseUpdateReflashKernel<NC, false><<<grid, kBlock, 0, st>>>(
                fp, op, n, T, primary, phaseState, status, flags, q, list, pendingList, recs,
                reflashOrder);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    if (splitA && (e = hipEventRecord(splitA, st)) != hipSuccess)
        return e;
    if (splitB && (e = hipEventRecord(splitB, st)) != hipSuccess)
        return e;
    if (well)
        phaseUpdateCoopKernel<NC, true><<<coopGrid, kBlock, 0, st>>>(fp, op, T, primary,
                                                                     phaseState, status, flags,
                                                                     q, list);
    else
        phaseUpdateCoopKernel<NC, false><<<coopGrid, kBlock, 0, st>>>(fp, op, T, primary,
                                                                      phaseState, status, flags,
                                                                      q, list);
    return hipGetLastError();
}

template <int NC>
hipError_t Launch<NC>::record(bool well, hipStream_t st, const FluidParams *fp, AssemblyView v,
                              const double *primary, const double *phaseState, double *accum,
                              int32_t *status)
{
    const unsigned grid = static_cast<unsigned>((v.ncell + 255) / 256);
    if (well)
        cellRecordKernel<NC, true><<<grid, 256, 0, st>>>(fp, v, primary, phaseState, accum, status);
    else
        cellRecordKernel<NC, false><<<grid, 256, 0, st>>>(fp, v, primary, phaseState, accum, status);
    return hipGetLastError();
}

template <int NC>
hipError_t Launch<NC>::assemble(bool well, hipStream_t st, const FluidParams *fp, AssemblyView v,
                                double dt, double fugScale, const double *primary,
                                const double *phaseState, const double *accumPrev,
                                double *residual, double *bsr, uint8_t *decisions,
                                int32_t *status)
{
    if constexpr (NC > kAssemblyMaxNc)
    {
        return hipErrorNotSupported;
    }
    else
    {
    const int chunks = well ? assemblyChunkCount<NC, true>() : assemblyChunkCount<NC, false>();
    const int64_t items = v.ncell * (bsr ? chunks : 1);
    const unsigned grid = static_cast<unsigned>((items + 63) / 64);
    if (well)
        assembleKernel<NC, true><<<grid, 64, 0, st>>>(fp, v, dt, fugScale, primary, phaseState,
                                                      accumPrev, residual, bsr, decisions, status);
    else
        assembleKernel<NC, false><<<grid, 64, 0, st>>>(fp, v, dt, fugScale, primary, phaseState,
                                                       accumPrev, residual, bsr, decisions, status);
    return hipGetLastError();
    }
}

template <int NC>
hipError_t Launch<NC>::wells(hipStream_t st, const FluidParams *fp, AssemblyView v, int nwells,
                             const int64_t *cells, int64_t ncells, const double *primary,
                             const double *phaseState, double *residual, double *bsr,
                             int32_t *status)
{
    if constexpr (NC > kAssemblyMaxNc)
    {
        return hipErrorNotSupported;
    }
    else
    {
        if (bsr && ncells > 0)
        {
            const unsigned grid = static_cast<unsigned>((ncells + 63) / 64);
            wellJacobianKernel<NC, true><<<grid, 64, 0, st>>>(fp, v, cells, ncells, primary,
                                                              phaseState, bsr, status);
            if (const hipError_t e = hipGetLastError(); e != hipSuccess)
                return e;
        }
        if (residual && nwells > 0)
        {
            wellControlKernel<NC, true><<<1, 1, 0, st>>>(v, nwells, primary, residual);
            if (const hipError_t e = hipGetLastError(); e != hipSuccess)
                return e;
        }
        return hipSuccess;
    }
}

} // namespace nn
