// phase_update.hpp — hysteretic O/G/W active-set update of one cell.
//
// Follows models/include/natural/state/three_phase_equilibrium.hpp:
//   sanitizePrimaryBeforeFlash           :505-543 (useVariableBounds = false, the
//                                        runtime default, natural_runtime_options.hpp:40)
//   updatePhaseState                     :282-496
//   updateSecondary                      :164-270
//   assignFlashResult                    :128-156
//   canonicalizeSwNonAqueousRoles_       :555-622
//   assignCanonicalSingleNonaqueousReducedSet_ :651-695
//   hasStrongMissingPhaseInstability_    :698-722
//   transitionOverallComposition_        :763-788
// with the primary / phase-state layouts of indices/include/indices/indices.hpp:46-292
// and the Vec codec of models/include/natural/petsc/phase_state_codec.hpp:27-150.
#pragma once

#include "common.hpp"
#include "flash.hpp"

namespace nn
{

// Primary / phase-state layout of the fully compositional mode.
template <int NC, bool Well>
struct Layout
{
    static constexpr int NP = 3 * NC + 1 + (Well ? 1 : 0);
    static constexpr int NPS = 3 * NC + 8;
    static constexpr int pressure = 0;
    __host__ __device__ static constexpr int compBegin(int phase) { return 1 + phase * (NC - 1); }
    static constexpr int satBegin = 3 * NC - 2;  // S_o, S_g, S_w
    static constexpr int well = 3 * NC + 1;
    // phase-state Vec entries
    static constexpr int flag = 0;
    static constexpr int kgBegin = 1;
    static constexpr int kwBegin = 1 + NC;
    static constexpr int zBegin = 1 + 2 * NC;
    static constexpr int betaBegin = 1 + 3 * NC;
    static constexpr int zfBegin = 4 + 3 * NC;
    static constexpr int suppression = 7 + 3 * NC;
};

constexpr double kProbeSaturation = 1.0e-4;        // numerics.hpp:30
constexpr double kAppearanceMargin = 1.0e-6;       // numerics.hpp:43
constexpr double kHysteresisMargin = 1.0e-4;       // numerics.hpp:54
constexpr double kMinNormDenominator = 1.0e-14;    // numerics.hpp:55

template <int NC>
struct PhaseStateData
{
    uint8_t presence;
    uint8_t suppression;
    double kg[NC], kw[NC], z[NC];
    double beta[3];
    double Zf[3];
};

enum : uint8_t
{
    kFlagRemoved = 1,
    kFlagStabilityInvalid = 2,
    kFlagMissingUnstable = 4,
    kFlagRestrictedFailed = 8,
    kFlagUnrestrictedFailed = 16,
    kFlagStableReduced = 32,
};

template <int NC, bool Well, int Mode = kModePlain>
struct PhaseUpdater
{
    using L = Layout<NC, Well>;
    using Result = FlashResult<NC>;
    static constexpr int NP = L::NP;

    const FluidParams &f;
    const FlashOptions &o;
    Err &err;
    Flash<NC, Mode> flash;
    double T;
    // Record/replay of the stability test this update runs directly on its
    // active set (the common path: no phase removed, fewer than three phases).
    int stabMode = kStabNone;
    StabRecord<NC> *stabRec = nullptr;

    __host__ __device__ PhaseUpdater(const FluidParams &fp, const FlashOptions &op, Err &e, double temp)
        : f(fp), o(op), err(e), flash(fp, op, e), T(temp)
    {
    }

    __host__ __device__ bool sw() const { return f.model == NN_MODEL_SW; }

    __host__ __device__ void normalizeFloor(double *x)
    {
        double sum = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            x[c] = maxd(x[c], o.floor);
            sum += x[c];
        }
        if (!(sum > 0.0) || !finite(sum))
        {
            err.raise(kErrThermo);
            return;
        }
        for (int c = 0; c < NC; ++c) x[c] /= sum;
    }

    __host__ __device__ static void compositionFromPrimary(const double *prim, int phase, double *x)
    {
        x[NC - 1] = 1.0;
        for (int i = 0; i < NC - 1; ++i)
        {
            x[i] = prim[L::compBegin(phase) + i];
            x[NC - 1] -= x[i];
        }
    }

    __host__ __device__ static void writeComposition(double *prim, const double *x, int phase)
    {
        for (int i = 0; i < NC - 1; ++i)
            prim[L::compBegin(phase) + i] = x[i];
    }

    __host__ __device__ void normalizedCompositions(const double *prim, double (*x)[NC])
    {
        for (int p = 0; p < 3; ++p)
        {
            compositionFromPrimary(prim, p, x[p]);
            normalizeFloor(x[p]);
        }
    }

    // :505-543
    __host__ __device__ void sanitize(double *prim) const
    {
        for (int p = 0; p < 3; ++p)
        {
            double sum = 0.0;
            for (int i = 0; i < NC - 1; ++i)
            {
                double &v = prim[L::compBegin(p) + i];
                v = clampd(v, o.floor, 1.0 - o.floor);
                sum += v;
            }
            const double maxSum = 1.0 - o.floor;
            if (sum > maxSum)
            {
                const double scale = maxSum / sum;
                for (int i = 0; i < NC - 1; ++i) prim[L::compBegin(p) + i] *= scale;
            }
        }
    }

    __host__ __device__ double wilsonLnK(double P, int c) const
    {
        return ::log(f.pc[c] / P) + 5.373 * (1.0 + f.omega[c]) * (1.0 - f.tc[c] / T);
    }

    // :555-622
    __host__ __device__ void canonicalizeSwNonAqueousRoles(double *prim, PhaseStateData<NC> &ps)
    {
        if (!sw() || !(ps.presence & kOilBit) || !(ps.presence & kGasBit))
            return;
        double comp[3][NC];
        normalizedCompositions(prim, comp);
        if (err.bad())
            return;
        const double P = prim[L::pressure];
        double orientation = 0.0, magnitude = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            const double term = (comp[1][c] - comp[0][c]) * wilsonLnK(P, c);
            orientation += term;
            magnitude += ::fabs(term);
        }
        if (orientation >= -1.0e-10 * maxd(1.0, magnitude))
            return;
        double t = prim[L::satBegin];
        prim[L::satBegin] = prim[L::satBegin + 1];
        prim[L::satBegin + 1] = t;
        for (int i = 0; i < NC - 1; ++i)
        {
            t = prim[L::compBegin(0) + i];
            prim[L::compBegin(0) + i] = prim[L::compBegin(1) + i];
            prim[L::compBegin(1) + i] = t;
        }
        t = ps.beta[0]; ps.beta[0] = ps.beta[1]; ps.beta[1] = t;
        t = ps.Zf[0]; ps.Zf[0] = ps.Zf[1]; ps.Zf[1] = t;
        double canon[3][NC];
        normalizedCompositions(prim, canon);
        for (int c = 0; c < NC; ++c)
        {
            ps.kg[c] = canon[1][c] / maxd(canon[0][c], o.floor);
            if (ps.presence & kWaterBit)
                ps.kw[c] = canon[2][c] / maxd(canon[0][c], o.floor);
        }
    }

    // :763-788
    __host__ __device__ void transitionOverall(const double *prim, const PhaseStateData<NC> &ps, double *z)
    {
        double comp[3][NC];
        normalizedCompositions(prim, comp);
        for (int c = 0; c < NC; ++c) z[c] = 0.0;
        double betaSum = 0.0;
        for (int p = 0; p < 3; ++p)
        {
            if (!(ps.presence & (1u << p)))
                continue;
            const double beta = maxd(ps.beta[p], 0.0);
            betaSum += beta;
            for (int c = 0; c < NC; ++c) z[c] += beta * comp[p][c];
        }
        if (!(betaSum > kMinNormDenominator))
        {
            for (int c = 0; c < NC; ++c) z[c] = ps.z[c];
            return;
        }
        for (int c = 0; c < NC; ++c) z[c] /= betaSum;
        normalizeFloor(z);
    }

    // :128-156 (+ overallFromFlash_ :749-761)
    __host__ __device__ void assignFlashResult(double *prim, PhaseStateData<NC> &ps, const Result &r)
    {
        if (!r.converged)
        {
            err.raise(kErrThermo);
            return;
        }
        prim[L::satBegin] = r.sat[0];
        prim[L::satBegin + 1] = r.sat[1];
        prim[L::satBegin + 2] = r.sat[2];
        for (int p = 0; p < 3; ++p) writeComposition(prim, r.x[p], p);
        ps.presence = r.presence;
        ps.suppression &= uint8_t(~r.presence) & kAllBits;
        double z[NC];
        for (int c = 0; c < NC; ++c) z[c] = 0.0;
        for (int p = 0; p < 3; ++p)
            for (int c = 0; c < NC; ++c) z[c] += r.beta[p] * r.x[p][c];
        normalizeFloor(z);
        for (int c = 0; c < NC; ++c)
        {
            ps.z[c] = z[c];
            ps.kg[c] = r.kg[c];
            ps.kw[c] = r.kw[c];
        }
        for (int p = 0; p < 3; ++p)
        {
            ps.beta[p] = r.beta[p];
            ps.Zf[p] = r.Z[p];
        }
    }

    // :164-270
    __host__ __device__ void updateSecondary(double *prim, PhaseStateData<NC> &ps)
    {
        double comp[3][NC];
        normalizedCompositions(prim, comp);
        if (err.bad())
            return;
        const double sat[3] = {prim[L::satBegin], prim[L::satBegin + 1], prim[L::satBegin + 2]};
        double moles[3] = {0.0, 0.0, 0.0};
        const double P = prim[L::pressure];
        double total = 0.0;
        const bool cpa = f.model == NN_MODEL_CPA;
        MixParams<NC, double> nonAq, aq;
        CpaTemp<NC> cpaT;
        if (cpa)
        {
            cpaTemperature<NC>(f, T, cpaT);
        }
        else
        {
            mixingParameters<NC, double>(f, P, T, kOil, nonAq);
            if (sw())
                mixingParameters<NC, double>(f, P, T, kWater, aq);
        }
        for (int p = 0; p < 3; ++p)
        {
            if (!(ps.presence & (1u << p)))
                continue;
            double Z;
            if (cpa)
            {
                // :188-200: eos.compressibility (CPA density root)
                Z = cpaCompressibility<NC, Mode == kModeCoop>(f, cpaT, P, T, comp[p], p != kGas, err);
            }
            else
            {
                const MixParams<NC, double> &mp = (p == kWater && sw()) ? aq : nonAq;
                PhaseMix<NC, double> mix;
                phaseMixing<NC, double>(comp[p], mp, mix);
                Z = selectRoot(f, mix.A, mix.B, p != kGas, err);
            }
            if (err.bad())
                return;
            ps.Zf[p] = Z;
            const double rhoM = phaseMolarDensity<NC, double>(f, P, T, comp[p], Z, p == kWater, err);
            if (err.bad())
                return;
            moles[p] = rhoM * maxd(sat[p], 0.0);
            total += moles[p];
        }
        if (!(total > kMinNormDenominator))
        {
            err.raise(kErrThermo);
            return;
        }
        for (int c = 0; c < NC; ++c) ps.z[c] = 0.0;
        for (int p = 0; p < 3; ++p)
        {
            ps.beta[p] = moles[p] / total;
            for (int c = 0; c < NC; ++c) ps.z[c] += ps.beta[p] * comp[p][c];
        }
        normalizeFloor(ps.z);
        if ((ps.presence & kOilBit) && (ps.presence & kGasBit))
            for (int c = 0; c < NC; ++c)
                ps.kg[c] = comp[1][c] / maxd(comp[0][c], o.floor);
        if ((ps.presence & kOilBit) && (ps.presence & kWaterBit))
            for (int c = 0; c < NC; ++c)
                ps.kw[c] = comp[2][c] / maxd(comp[0][c], o.floor);
    }

    // :698-722
    __host__ __device__ static bool strongInstability(const StabilityResult<NC> &st, uint8_t present,
                                                      uint8_t suppression)
    {
        if (!st.valid)
            return false;
        for (int p = 0; p < 3; ++p)
        {
            if (present & (1u << p))
                continue;
            const double margin = (suppression & (1u << p)) ? kHysteresisMargin : kAppearanceMargin;
            if (st.unstable[p] && st.trialSum[p] > 1.0 + margin)
                return true;
        }
        return false;
    }

    // :651-695
    __host__ __device__ bool assignCanonicalSingleNonaqueous(double *prim, PhaseStateData<NC> &ps,
                                                             const double *overall, uint8_t active,
                                                             double (*seed)[NC])
    {
        if (sw())
            return false;
        const bool hasOil = active & kOilBit;
        const bool hasGas = active & kGasBit;
        if (hasOil == hasGas)
            return false;
        const int act = hasOil ? kOil : kGas;
        const double P = prim[L::pressure];
        double vol = 0.0, mag = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            const double term = seed[act][c] * wilsonLnK(P, c);
            vol += term;
            mag += ::fabs(term);
        }
        const int expected = vol >= -1.0e-10 * maxd(1.0, mag) ? kGas : kOil;
        if (expected == act)
            return false;
        uint8_t corrected = uint8_t((active & ~(1u << act)) | (1u << expected));
        for (int c = 0; c < NC; ++c) seed[expected][c] = seed[act][c];
        const Result canon = flash.flashRestricted(P, T, overall, corrected, seed);
        if (err.bad())
            return false;
        if (!canon.converged || canon.presence != corrected)
            return false;
        assignFlashResult(prim, ps, canon);
        return true;
    }

    // :282-496; returns (status << 8) | update flags.
    __host__ __device__ int update(double *prim, PhaseStateData<NC> &ps)
    {
        int flags = 0;
        canonicalizeSwNonAqueousRoles(prim, ps);
        if (err.bad())
            return flags;
        double original[NP];
        for (int k = 0; k < NP; ++k) original[k] = prim[k];
        const PhaseStateData<NC> originalPs = ps;

        double overall[NC];
        transitionOverall(prim, ps, overall);
        if (err.bad())
            return flags;

        uint8_t active = ps.presence;
        ps.suppression &= uint8_t(~active) & kAllBits;
        bool removed = false;
        const double newtonSat[3] = {prim[L::satBegin], prim[L::satBegin + 1], prim[L::satBegin + 2]};
        for (int p = 0; p < 3; ++p)
        {
            if ((active & (1u << p)) && newtonSat[p] <= kProbeSaturation)
            {
                active &= uint8_t(~(1u << p));
                ps.suppression |= uint8_t(1u << p);
                prim[L::satBegin + p] = 0.0;
                removed = true;
            }
        }
        if (active == 0)
        {
            int survivor = 0;
            for (int p = 1; p < 3; ++p)
                if (newtonSat[survivor] < newtonSat[p])
                    survivor = p;
            active |= uint8_t(1u << survivor);
            ps.suppression &= uint8_t(~(1u << survivor));
            prim[L::satBegin + survivor] = 1.0;
            removed = true;
        }
        if (removed)
            flags |= kFlagRemoved;
        ps.presence = active;
        for (int c = 0; c < NC; ++c) ps.z[c] = overall[c];

        double comps[3][NC];
        normalizedCompositions(prim, comps);
        if (err.bad())
            return flags;
        const double P = prim[L::pressure];

        if (removed)
        {
            const Result reduced = flash.flashRestricted(P, T, overall, active, comps);
            NN_TRACE("update: restricted flash active=%d conv=%d pres=%d err=%d\n", (int)active,
                     (int)reduced.converged, (int)reduced.presence, err.code);
            if (err.bad())
                return flags;
            if (reduced.converged)
            {
                bool unstableMissing = false;
                if (popcount3(reduced.presence) < 3)
                {
                    const StabilityResult<NC> st =
                        flash.stabilityTest(P, T, overall, reduced.presence, reduced.x);
                    NN_TRACE("update: stability of reduced set valid=%d err=%d\n", (int)st.valid,
                             err.code);
                    if (err.bad())
                        return flags;
                    if (!st.valid)
                    {
                        flags |= kFlagStabilityInvalid;
                        unstableMissing = true;
                    }
                    else
                    {
                        unstableMissing = strongInstability(st, reduced.presence, ps.suppression);
                        if (unstableMissing)
                            flags |= kFlagMissingUnstable;
                    }
                }
                if (!unstableMissing)
                {
                    double seed[3][NC];
                    for (int p = 0; p < 3; ++p)
                        for (int c = 0; c < NC; ++c) seed[p][c] = reduced.x[p][c];
                    if (assignCanonicalSingleNonaqueous(prim, ps, overall, reduced.presence, seed))
                        return flags | kFlagStableReduced;
                    if (err.bad())
                        return flags;
                    assignFlashResult(prim, ps, reduced);
                    return flags | kFlagStableReduced;
                }
            }
            if constexpr (Mode == kModeDefer && NC != 3)
            {
                // The lane-per-cell device build returned wrong statuses, flags and
                // states for cells on this path (restricted flash, stability test,
                // then the unrestricted re-flash in one lane) at six components,
                // while the host build of the same code and the wavefront-
                // cooperative device build were right: such cells go to the
                // cooperative pass, which redoes the update from the inputs.  At
                // three components the lane build of the path matches the
                // reference on every parity grid (PR, SW, PR-CPA), and stays.
                err.raise(kErrDefer);
                return flags;
            }
            const Result reflashed = flash.flash(P, T, overall);
            NN_TRACE("update: reflash conv=%d pres=%d err=%d\n", (int)reflashed.converged,
                     (int)reflashed.presence, err.code);
            if (err.bad())
                return flags;
            if (reflashed.converged)
            {
                assignFlashResult(prim, ps, reflashed);
                return flags;
            }
            if (!reduced.converged)
                flags |= kFlagRestrictedFailed;
            flags |= kFlagUnrestrictedFailed;
            for (int k = 0; k < NP; ++k) prim[k] = original[k];
            ps = originalPs;
            return flags | (NN_STATUS_RECOVERABLE_FAILURE << 8);
        }
        else if (popcount3(active) < 3)
        {
            flash.stabMode = stabMode == kStabReplayHoldReflash ? kStabReplay : stabMode;
            flash.stabRec = stabRec;
            const StabilityResult<NC> st = flash.stabilityTest(P, T, overall, active, comps);
            flash.stabMode = kStabNone;
            if (err.bad())
                return flags;
            const bool invalid = !st.valid;
            const bool missing = !invalid && strongInstability(st, active, ps.suppression);
            if (invalid)
                flags |= kFlagStabilityInvalid;
            if (missing)
                flags |= kFlagMissingUnstable;
            if (invalid || missing)
            {
                if (stabMode == kStabReplayHoldReflash)
                {
                    err.raise(kErrReflash);
                    return flags;
                }
                const Result reflashed = flash.flash(P, T, overall);
                if (err.bad())
                    return flags;
                if (reflashed.converged)
                {
                    assignFlashResult(prim, ps, reflashed);
                    return flags;
                }
                flags |= kFlagUnrestrictedFailed;
                for (int k = 0; k < NP; ++k) prim[k] = original[k];
                ps = originalPs;
                return flags | (NN_STATUS_RECOVERABLE_FAILURE << 8);
            }
        }

        if (popcount3(active) < 3 && assignCanonicalSingleNonaqueous(prim, ps, overall, active, comps))
            return flags | kFlagStableReduced;
        if (err.bad())
            return flags;

        for (int p = 0; p < 3; ++p)
        {
            if (ps.presence & (1u << p))
                continue;
            prim[L::satBegin + p] = 0.0;
            writeComposition(prim, overall, p);
        }
        updateSecondary(prim, ps);
        if (popcount3(ps.presence) < 3)
            flags |= kFlagStableReduced;
        return flags;
    }
};

template <int NC, bool Well>
__host__ __device__ bool decodePhaseState(const double *v, PhaseStateData<NC> &ps)
{
    using L = Layout<NC, Well>;
    const double flag = v[L::flag];
    const int raw = int(flag + (flag >= 0.0 ? 0.5 : -0.5));
    if (raw <= 0 || raw > 7)
        return false;
    ps.presence = uint8_t(raw);
    ps.suppression = uint8_t(int(v[L::suppression] + 0.5)) & kAllBits;
    for (int c = 0; c < NC; ++c)
    {
        ps.kg[c] = v[L::kgBegin + c];
        ps.kw[c] = v[L::kwBegin + c];
        ps.z[c] = v[L::zBegin + c];
    }
    for (int p = 0; p < 3; ++p)
    {
        ps.beta[p] = v[L::betaBegin + p];
        ps.Zf[p] = v[L::zfBegin + p];
    }
    return true;
}

template <int NC, bool Well>
__host__ __device__ void encodePhaseState(const PhaseStateData<NC> &ps, double *v)
{
    using L = Layout<NC, Well>;
    v[L::flag] = double(ps.presence);
    for (int c = 0; c < NC; ++c)
    {
        v[L::kgBegin + c] = ps.kg[c];
        v[L::kwBegin + c] = ps.kw[c];
        v[L::zBegin + c] = ps.z[c];
    }
    for (int p = 0; p < 3; ++p)
    {
        v[L::betaBegin + p] = ps.beta[p];
        v[L::zfBegin + p] = ps.Zf[p];
    }
    v[L::suppression] = double(ps.suppression);
}

// One cell of nn_phase_update_batch (natural_petsc_runtime_state.inc:215-221):
// kItemDone, kItemDeferred (a kModeDefer run handed the cell to the cooperative
// pass), kItemPending (kStabRecord: its stability-test inputs are in rec; the
// lockstep trial pass and a replay finish it) or kItemReflash
// (kStabReplayHoldReflash: its re-flash is left to a pass over all such cells).
// Deferred, pending and held cells leave primary/phaseState untouched; in
// kModeCoop only lane 0 writes.
template <int NC, bool Well, int Mode = kModePlain>
__host__ __device__ int phaseUpdateCell(const FluidParams &f, const FlashOptions &o, double T,
                                        double *primary, double *phaseState, int64_t s,
                                        int32_t *status, uint8_t *flags,
                                        int stabMode = kStabNone, StabRecord<NC> *rec = nullptr)
{
    using L = Layout<NC, Well>;
    const bool writer = Mode != kModeCoop || waveLane() == 0;
    double prim[L::NP];
    double psv[L::NPS];
    for (int k = 0; k < L::NP; ++k) prim[k] = primary[s * L::NP + k];
    for (int k = 0; k < L::NPS; ++k) psv[k] = phaseState[s * L::NPS + k];
    PhaseStateData<NC> ps;
    if (!decodePhaseState<NC, Well>(psv, ps))
    {
        // decodePhasePresence throws std::runtime_error (phase_behavior.hpp:80-86).
        if (writer)
        {
            status[s] = NN_STATUS_THERMO_ERROR;
            if (flags) flags[s] = 0;
        }
        return kItemDone;
    }
    Err err;
    PhaseUpdater<NC, Well, Mode> up(f, o, err, T);
    up.stabMode = rec ? stabMode : kStabNone;
    up.stabRec = rec ? rec + s : nullptr;
    up.sanitize(prim);
    const int packed = up.update(prim, ps);
    const int st = packed >> 8;
    const uint8_t fl = uint8_t(packed & 0xff);
    if (err.code == kErrDefer)
        return kItemDeferred;
    if (err.code == kErrPending)
        return kItemPending;
    if (err.code == kErrReflash)
        return kItemReflash;
    if (!writer)
        return kItemDone;
    if (err.bad())
    {
        // The reference throws out of updatePhaseState: inputs stay untouched and
        // there is no PhaseUpdateReport, so no flags either.
        status[s] = err.code;
        if (flags) flags[s] = 0;
        return kItemDone;
    }
    encodePhaseState<NC, Well>(ps, psv);
    for (int k = 0; k < L::NP; ++k) primary[s * L::NP + k] = prim[k];
    for (int k = 0; k < L::NPS; ++k) phaseState[s * L::NPS + k] = psv[k];
    status[s] = st;
    if (flags) flags[s] = fl;
    return kItemDone;
}

} // namespace nn
