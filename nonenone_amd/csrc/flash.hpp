// flash.hpp — O/G/W three-phase P-T-z flash and tangent-plane stability for
// one state per work-item, templated on the component count.
//
// The control flow follows the reference's CubicThreePhaseFlash so that phase
// labels, iteration counts and converged values agree:
//   flash                     thermo/three_phase_flash.hpp:83-227
//   flashRestricted           three_phase_flash.hpp:235-292
//   stabilityTest             three_phase_flash.hpp:300-489
//   updateSaturationsFromMoles three_phase_flash.hpp:527-575
//   core helpers              thermo/detail/three_phase_flash_core.inc:66-188
//   Rachford-Rice, flashThree thermo/detail/three_phase_flash_nonlinear.inc:8-513
//   pairInitialK              three_phase_flash_nonlinear.inc:515-622
//   flashTwo, single phase,
//   reduced-set fallbacks     thermo/detail/three_phase_flash_restricted.inc:8-779
//   equilibrateActiveSet      thermo/detail/three_phase_flash_allocation.inc:707-858
//   canonicalization          thermo/detail/three_phase_flash_canonicalization.inc:8-839
//   stability trials          thermo/detail/three_phase_flash_stability.inc:32-355
// Where the reference throws, Err is raised and the public call ends with a
// THERMO_ERROR / INVALID_INPUT status.
//
// The nested solver instances of the reference (SW role permutations,
// pressure-continuation anchors) run without further recovery, so they are
// expressed here as a second Flash object whose flash() stops before those two
// stages; no device recursion is needed.
#pragma once

#include "common.hpp"
#include "aqueous.hpp"
#include "cpa.hpp"
#include "eos.hpp"

// Optional control-flow trace (printf-style), compiled out unless defined.
#ifndef NN_TRACE
#define NN_TRACE(...) ((void)0)
#endif

namespace nn
{

template <int NC>
struct FlashResult
{
    bool converged;
    int iterations;
    uint8_t presence;
    double beta[3];
    double sat[3];
    double x[3][NC];
    double Z[3];
    double rho[3];
    double kg[NC];
    double kw[NC];

    __host__ __device__ FlashResult()
        : converged(false), iterations(0), presence(kAllBits)
    {
        beta[0] = 1.0; beta[1] = 0.0; beta[2] = 0.0;
        sat[0] = 1.0; sat[1] = 0.0; sat[2] = 0.0;
        for (int p = 0; p < 3; ++p)
        {
            Z[p] = 1.0;
            rho[p] = 0.0;
            for (int c = 0; c < NC; ++c) x[p][c] = 0.0;
        }
        for (int c = 0; c < NC; ++c) { kg[c] = 0.0; kw[c] = 0.0; }
    }
};

template <int NC>
struct StabilityResult
{
    bool stable;
    bool valid;
    uint8_t tested;
    bool unstable[3];
    double incipient[3][NC];
    double trialSum[3];

    __host__ __device__ StabilityResult() : stable(true), valid(true), tested(0)
    {
        for (int p = 0; p < 3; ++p)
        {
            unstable[p] = false;
            trialSum[p] = 1.0;
            for (int c = 0; c < NC; ++c) incipient[p][c] = 0.0;
        }
    }
};

// Mutable per-(P,T) cache of the mixing parameters (three_phase_flash_core.inc:66-99).
template <int NC>
struct ParamCache
{
    bool valid = false;
    double P = 0.0, T = 0.0;
    MixParams<NC, double> nonAqueous;
    MixParams<NC, double> aqueous;
    CpaTemp<NC> cpa;  // CPA: temperature part of the mixing state (cubic_eos.hpp:1356-1414)
    // CPA: the last two phase evaluations (inputs and results).  A CPA phase
    // evaluation is a pure function of (P, T, x, liquid) once the temperature
    // part is cached, and the flash's loops often repeat the previous one or
    // the one before (e.g. the converged phases again for the canonical
    // thermodynamics or a stability test's reference phase), so a repeated call
    // takes the stored result: the same bits, without the density root.
    struct CpaMemo
    {
        double P = -1.0, T = -1.0, x[NC] = {};
        bool liquid = false;
        double Z = 0.0, phi[NC] = {};
    } memo[2];
    int memoNext = 0;
};

// Gaussian elimination with partial pivoting (three_phase_flash_stability.inc:98-146).
template <int M>
__host__ __device__ bool solveDense(double (&a)[M][M], double (&rhs)[M], double (&sol)[M])
{
    for (int i = 0; i < M; ++i) sol[i] = 0.0;
    for (int col = 0; col < M; ++col)
    {
        int pivot = col;
        for (int row = col + 1; row < M; ++row)
            if (::fabs(a[row][col]) > ::fabs(a[pivot][col]))
                pivot = row;
        if (!(::fabs(a[pivot][col]) > 1.0e-13))
            return false;
        if (pivot != col)
        {
            for (int j = 0; j < M; ++j)
            {
                const double t = a[pivot][j];
                a[pivot][j] = a[col][j];
                a[col][j] = t;
            }
            const double t = rhs[pivot];
            rhs[pivot] = rhs[col];
            rhs[col] = t;
        }
        for (int row = col + 1; row < M; ++row)
        {
            const double factor = a[row][col] / a[col][col];
            if (!finite(factor))
                return false;
            for (int j = col; j < M; ++j)
                a[row][j] -= factor * a[col][j];
            rhs[row] -= factor * rhs[col];
        }
    }
    for (int rev = 0; rev < M; ++rev)
    {
        const int row = M - 1 - rev;
        double value = rhs[row];
        for (int j = row + 1; j < M; ++j)
            value -= a[row][j] * sol[j];
        if (!(::fabs(a[row][row]) > 1.0e-13))
            return false;
        sol[row] = value / a[row][row];
        if (!finite(sol[row]))
            return false;
    }
    return true;
}

template <int NC, int Mode = kModePlain>
class Flash
{
public:
    using Result = FlashResult<NC>;
    using Stab = StabilityResult<NC>;
    static constexpr int K2 = 2 * NC;

    const FluidParams &f;
    const FlashOptions &o;
    Err &err;
    int roles[3];
    bool enablePermutation;
    bool enableContinuation;
    ParamCache<NC> cache;
    // Set when a state reaches the Soreide-Whitson allocation basin search,
    // which this build does not provide (the state then reports not converged).
    bool allocationFallbackReached = false;
    // Set when a lowest-Gibbs selection compared two candidates whose reduced
    // Gibbs energies agree to 1e-12 (equivalent solutions): which one is kept
    // then follows the last bits of the arithmetic.  Diagnostic only.
    bool nearTie = false;
    // RR3 residual evaluations so far (kModeDefer's budget; per lane in kModeCoop).
    mutable int work = 0;
    // Learned-initialiser hint for this solver's first gas-oil start (no
    // counterpart in the reference): NN_HINT_SINGLE_PHASE skips that start and
    // goes to the single-phase candidate, NN_HINT_LOG_K replaces its Wilson K by
    // exp(hintLogK).  The active-set equilibration's stability tests judge the
    // outcome either way; solvers made for recovery paths carry no hint.
    int hint = 0;
    const double *hintLogK = nullptr;
    // Record/replay of the next stabilityTest's seed trials (kStabRecord /
    // kStabReplay, see StabRecord); reset to kStabNone by that call.
    int stabMode = kStabNone;
    StabRecord<NC> *stabRec = nullptr;

    __host__ __device__ Flash(const FluidParams &fp, const FlashOptions &op, Err &e)
        : f(fp), o(op), err(e), enablePermutation(true), enableContinuation(true)
    {
        roles[0] = kOil; roles[1] = kGas; roles[2] = kWater;
    }

    __host__ __device__ Flash(const FluidParams &fp, const FlashOptions &op, Err &e,
                              const int r[3], bool perm, bool cont)
        : f(fp), o(op), err(e), enablePermutation(perm), enableContinuation(cont)
    {
        roles[0] = r[0]; roles[1] = r[1]; roles[2] = r[2];
    }

    __host__ __device__ bool sw() const { return f.model == NN_MODEL_SW; }
    __host__ __device__ int roleOf(int slot) const { return sw() ? roles[slot] : slot; }

    // ------------------------------------------------------------ EOS access
    __host__ __device__ void ensureCache(double P, double T)
    {
        if (cache.valid && cache.P == P && cache.T == T)
            return;
        mixingParameters<NC, double>(f, P, T, kOil, cache.nonAqueous);
        if (sw())
            mixingParameters<NC, double>(f, P, T, kWater, cache.aqueous);
        cache.P = P;
        cache.T = T;
        cache.valid = true;
    }

    __host__ __device__ void thermoForRole(int role, double P, double T, const double *x,
                                           PhaseThermo<NC> &out)
    {
        if (f.model == NN_MODEL_CPA)
        {
            // three_phase_flash_core.inc:108-112: the EOS's own role-aware CPA result
            if (cache.cpa.T != T)
                cpaTemperature<NC>(f, T, cache.cpa);
            const bool liquid = role != kGas;
            for (int k = 0; k < 2; ++k)
            {
                const auto &m = cache.memo[k];
                bool same = m.P == P && m.T == T && m.liquid == liquid;
                for (int i = 0; i < NC; ++i) same = same && m.x[i] == x[i];
                if (same)
                {
                    out.Z = m.Z;
                    for (int i = 0; i < NC; ++i) out.phi[i] = m.phi[i];
                    return;
                }
            }
            double lnPhi[NC];
            cpaPhase<NC, double, Mode == kModeCoop>(f, cache.cpa, P, T, x, liquid, out.Z, lnPhi, err);
            for (int i = 0; i < NC; ++i) out.phi[i] = err.bad() ? 1.0 : ::exp(lnPhi[i]);
            if (!err.bad())
            {
                auto &m = cache.memo[cache.memoNext];
                cache.memoNext ^= 1;
                m.P = P;
                m.T = T;
                m.liquid = liquid;
                m.Z = out.Z;
                for (int i = 0; i < NC; ++i)
                {
                    m.x[i] = x[i];
                    m.phi[i] = out.phi[i];
                }
            }
            return;
        }
        ensureCache(P, T);
        const MixParams<NC, double> &mp =
            (role == kWater && sw()) ? cache.aqueous : cache.nonAqueous;
        phaseThermoFromParams<NC>(f, mp, role, x, out, err);
    }

    __host__ __device__ void thermo(int slot, double P, double T, const double *x,
                                    PhaseThermo<NC> &out)
    {
        thermoForRole(roleOf(slot), P, T, x, out);
    }

    // ------------------------------------------------------------- helpers
    __host__ __device__ bool validatePT(double P, double T)
    {
        if (!(P > 0.0) || !finite(P) || !(T > 0.0) || !finite(T))
        {
            NN_TRACE("invalid P/T %.17g %.17g\n", P, T);
            err.raise(kErrInvalid);
            return false;
        }
        return true;
    }

    __host__ __device__ void normalize(double *x)
    {
        double sum = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            if (!finite(x[c]))
            {
                NN_TRACE("invalid composition entry %d = %.17g\n", c, x[c]);
                err.raise(kErrInvalid);
                return;
            }
            x[c] = maxd(x[c], o.floor);
            sum += x[c];
        }
        if (!(sum > 0.0))
        {
            NN_TRACE("invalid composition sum %.17g\n", sum);
            err.raise(kErrInvalid);
            return;
        }
        for (int c = 0; c < NC; ++c) x[c] /= sum;
    }

    __host__ __device__ void wilsonK(double P, double T, double *K) const
    {
        for (int i = 0; i < NC; ++i)
        {
            const double lnK = ::log(f.pc[i] / P) +
                               5.373 * (1.0 + f.omega[i]) * (1.0 - f.tc[i] / T);
            K[i] = clampd(::exp(lnK), 1.0e-30, 1.0e30);
        }
    }

    __host__ __device__ void initialWaterOilK(double *K) const
    {
        for (int c = 0; c < NC; ++c) K[c] = 0.05;
        if (f.water >= 0)
            K[f.water] = 50.0;
        else
            K[NC - 1] = 50.0;
    }

    __host__ __device__ uint8_t presenceFromFractions(const double *beta) const
    {
        uint8_t b = 0;
        for (int p = 0; p < 3; ++p)
            if (beta[p] > o.betaTol)
                b |= uint8_t(1u << p);
        return b;
    }

    __host__ __device__ static int firstActive(uint8_t bits)
    {
        for (int p = 0; p < 3; ++p)
            if (bits & (1u << p))
                return p;
        return 0;
    }

    __host__ __device__ static void activePair(uint8_t bits, int &a, int &b)
    {
        int n = 0;
        int pair[2] = {0, 0};
        for (int p = 0; p < 3; ++p)
            if ((bits & (1u << p)) && n < 2)
                pair[n++] = p;
        a = pair[0];
        b = pair[1];
    }

    __host__ __device__ void deriveReferenceK(Result &r) const
    {
        for (int c = 0; c < NC; ++c)
        {
            r.kg[c] = r.x[1][c] / maxd(r.x[0][c], o.floor);
            r.kw[c] = r.x[2][c] / maxd(r.x[0][c], o.floor);
        }
    }

    __host__ __device__ static void swapSlots(Result &r, int a, int b)
    {
        if (a == b)
            return;
        double t = r.beta[a]; r.beta[a] = r.beta[b]; r.beta[b] = t;
        t = r.sat[a]; r.sat[a] = r.sat[b]; r.sat[b] = t;
        for (int c = 0; c < NC; ++c) { t = r.x[a][c]; r.x[a][c] = r.x[b][c]; r.x[b][c] = t; }
        t = r.Z[a]; r.Z[a] = r.Z[b]; r.Z[b] = t;
        t = r.rho[a]; r.rho[a] = r.rho[b]; r.rho[b] = t;
    }

    // The Water-role closures' composition domain (cubic_eos.hpp:449-468): every
    // composition when none is configured.
    __host__ __device__ bool aqueousSupported(const double *x) const
    {
        return aqueousDomain<NC>(f, x);
    }

    __host__ __device__ bool waterIsDominant(const double *x) const
    {
        if (f.water < 0)
            return false;
        if (!aqueousSupported(x))
            return false;
        double largest = 0.0;
        for (int c = 0; c < NC; ++c)
            if (c != f.water)
                largest = maxd(largest, x[c]);
        return x[f.water] >= largest;
    }

    // three_phase_flash.hpp:527-575
    __host__ __device__ void updateSaturationsFromMoles(double P, double T, Result &r)
    {
        double volumeSum = 0.0;
        double volume[3] = {0.0, 0.0, 0.0};
        for (int p = 0; p < 3; ++p)
        {
            if (!(r.presence & (1u << p)))
            {
                r.beta[p] = 0.0;
                r.sat[p] = 0.0;
                r.rho[p] = 0.0;
                continue;
            }
            const double Z = r.Z[p];
            if (!(Z > 0.0))
            {
                err.raise(kErrThermo);
                return;
            }
            const bool waterRole = roleOf(p) == kWater;
            if (waterRole && !aqueousSupported(r.x[p]))
            {
                // outside the Water-role closures' domain: not a converged state
                r.converged = false;
                for (int q = 0; q < 3; ++q)
                {
                    r.sat[q] = 0.0;
                    r.rho[q] = 0.0;
                }
                return;
            }
            const double rhoM = phaseMolarDensity<NC, double>(f, P, T, r.x[p], Z, waterRole, err);
            if (err.bad())
                return;
            r.rho[p] = rhoM;
            volume[p] = maxd(r.beta[p], 0.0) / rhoM;
            volumeSum += volume[p];
        }
        if (!(volumeSum > 0.0))
        {
            err.raise(kErrThermo);
            return;
        }
        for (int p = 0; p < 3; ++p)
            r.sat[p] = volume[p] / volumeSum;
    }

    // restricted.inc:425-447
    __host__ __device__ Result singlePhaseResult(double P, double T, const double *z, uint8_t active)
    {
        Result r;
        r.converged = true;
        r.presence = active;
        r.beta[0] = r.beta[1] = r.beta[2] = 0.0;
        const int p = firstActive(active);
        r.beta[p] = 1.0;
        for (int c = 0; c < NC; ++c) r.x[p][c] = z[c];
        for (int q = 0; q < 3; ++q)
            if (!(active & (1u << q)))
                for (int c = 0; c < NC; ++c) r.x[q][c] = z[c];
        PhaseThermo<NC> t;
        thermo(p, P, T, z, t);
        r.Z[p] = t.Z;
        if (err.bad())
            return r;
        deriveReferenceK(r);
        updateSaturationsFromMoles(P, T, r);
        return r;
    }

    // restricted.inc:449-478
    __host__ __device__ uint8_t chooseSinglePhase(double P, double T, const double *z)
    {
        PhaseThermo<NC> to, tg, tw;
        thermo(kOil, P, T, z, to);
        thermo(kGas, P, T, z, tg);
        thermo(kWater, P, T, z, tw);
        if (err.bad())
            return kOilBit;
        double go = 0.0, gg = 0.0, gw = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            go += z[c] * ::log(maxd(to.phi[c], 1.0e-300));
            gg += z[c] * ::log(maxd(tg.phi[c], 1.0e-300));
            gw += z[c] * ::log(maxd(tw.phi[c], 1.0e-300));
        }
        const double tie = 1.0e-10;
        if (gg + tie < go && gg + tie < gw)
            return kGasBit;
        if (gw + tie < go && gw + tie < gg)
            return kWaterBit;
        if (waterIsDominant(z) && ::fabs(gw - go) <= tie)
            return kWaterBit;
        if (gg + tie < go)
            return kGasBit;
        return kOilBit;
    }

    // restricted.inc:480-506 (slot index used as the thermodynamic role)
    __host__ __device__ double dimensionlessGibbs(double P, double T, const Result &r)
    {
        double g = 0.0;
        for (int p = 0; p < 3; ++p)
        {
            const double beta = r.beta[p];
            if (!(beta > o.betaTol))
                continue;
            PhaseThermo<NC> t;
            thermoForRole(p, P, T, r.x[p], t);
            if (err.bad())
                return 0.0;
            for (int c = 0; c < NC; ++c)
            {
                const double x = maxd(r.x[p][c], o.floor);
                const double phi = maxd(t.phi[c], 1.0e-300);
                g += beta * r.x[p][c] * (::log(x) + ::log(phi));
            }
        }
        return g;
    }

    // ----------------------------------------------------- Rachford-Rice
    struct Rr3 { bool converged; double bg; double bw; };
    struct Rr2 { bool converged; double beta; };

    __host__ __device__ double rr3Norm(const double *z, const double *kg, const double *kw,
                                       double cg, double cw, double *fgO, double *fwO,
                                       double *jggO, double *jwwO, double *jgwO) const
    {
        const double inf = __builtin_inf();
        if (!(cg > 0.0) || !(cw > 0.0) || !(cg + cw < 1.0))
            return inf;
        double fg = 0.0, fw = 0.0, jgg = 0.0, jww = 0.0, jgw = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            const double ag = kg[c] - 1.0;
            const double aw = kw[c] - 1.0;
            const double co = 1.0 - cg - cw;
            const double D = co + cg * kg[c] + cw * kw[c];
            if (!(D > 1.0e-12) || !finite(D))
                return inf;
            fg += z[c] * ag / D;
            fw += z[c] * aw / D;
            const double invD2 = z[c] / (D * D);
            jgg -= invD2 * ag * ag;
            jww -= invD2 * aw * aw;
            jgw -= invD2 * ag * aw;
        }
        if (fgO) { *fgO = fg; *fwO = fw; *jggO = jgg; *jwwO = jww; *jgwO = jgw; }
        return maxd(::fabs(fg), ::fabs(fw));
    }

    // nonlinear.inc:8-93
    // Start of one RR3 Newton solve from (ibg, ibw): clamped into the interior.
    __host__ __device__ static void rr3Start(double ibg, double ibw, double &bg, double &bw)
    {
        bg = clampd(ibg, 1.0e-4, 0.8);
        bw = clampd(ibw, 1.0e-4, 0.8);
        if (bg + bw >= 0.95)
        {
            const double scale = 0.8 / (bg + bw);
            bg *= scale;
            bw *= scale;
        }
    }

    // One iteration of the RR3 Newton solve with its backtracking line search:
    // kRr3Running (moved to a point of smaller residual norm), kRr3Converged
    // (the norm at (bg, bw) is below the tolerance) or kRr3Failed.
    enum : int
    {
        kRr3Running = 0,
        kRr3Converged = 1,
        kRr3Failed = 2,
    };
    __host__ __device__ int rr3Step(const double *z, const double *kg, const double *kw,
                                    double &bg, double &bw) const
    {
        double fg, fw, jgg, jww, jgw;
        ++work;
        const double norm = rr3Norm(z, kg, kw, bg, bw, &fg, &fw, &jgg, &jww, &jgw);
        if (!finite(norm))
            return kRr3Failed;
        if (norm < o.rrTol)
            return kRr3Converged;
        const double det = jgg * jww - jgw * jgw;
        if (::fabs(det) < 1.0e-24 || !finite(det))
            return kRr3Failed;
        const double dbg = (-fg * jww + fw * jgw) / det;
        const double dbw = (-fw * jgg + fg * jgw) / det;
        double lambda = 1.0;
        for (int ls = 0; ls < 40; ++ls)
        {
            const double cg = bg + lambda * dbg;
            const double cw = bw + lambda * dbw;
            ++work;
            const double cn = rr3Norm(z, kg, kw, cg, cw, nullptr, nullptr, nullptr, nullptr, nullptr);
            if (finite(cn) && cn < norm)
            {
                bg = cg;
                bw = cw;
                return kRr3Running;
            }
            lambda *= 0.5;
        }
        return kRr3Failed;
    }

    __host__ __device__ Rr3 solveRr3(const double *z, const double *kg, const double *kw,
                                     double ibg, double ibw) const
    {
        double bg, bw;
        rr3Start(ibg, ibw, bg, bw);
        for (int it = 0; it < o.maxRrIterations; ++it)
        {
            const int r = rr3Step(z, kg, kw, bg, bw);
            if (r == kRr3Converged)
                return {true, bg, bw};
            if (r == kRr3Failed)
                return {false, 0.0, 0.0};
        }
        return {false, 0.0, 0.0};
    }

    // nonlinear.inc:95-127: the first of 17 starting guesses that converges.
    __host__ __device__ Rr3 solveRr3Robust(const double *z, const double *kg, const double *kw,
                                           double pg, double pw)
    {
        const double guesses[17][2] = {
            {pg, pw}, {0.333, 0.333}, {0.50, 1.0e-4}, {1.0e-4, 0.50}, {0.80, 1.0e-4},
            {1.0e-4, 0.80}, {0.20, 0.20}, {0.60, 0.20}, {0.20, 0.60}, {0.05, 0.05},
            {0.90, 0.01}, {0.95, 1.0e-4}, {0.98, 1.0e-4}, {1.0e-4, 0.95}, {1.0e-4, 0.98},
            {0.01, 0.90}, {0.45, 0.45}};
        if constexpr (Mode == kModeCoop)
        {
            // Guess g on lane g, all lanes one Newton iteration at a time; the
            // lowest converged lane is the serial loop's answer, and the wave stops
            // as soon as every lane below it has failed (the serial loop would not
            // have tried the guesses above it).
            const int lane = waveLane();
            double bg = 0.0, bw = 0.0;
            int state = kRr3Failed;
            if (lane < 17)
            {
                rr3Start(guesses[lane][0], guesses[lane][1], bg, bw);
                state = kRr3Running;
            }
            for (int it = 0; it < o.maxRrIterations; ++it)
            {
                if (state == kRr3Running)
                    state = rr3Step(z, kg, kw, bg, bw);
                const unsigned long long running = waveBallot(state == kRr3Running);
                const int first = firstSetLane(waveBallot(state == kRr3Converged));
                if (running == 0ull ||
                    (first >= 0 && (running & ((1ull << first) - 1ull)) == 0ull))
                    break;
            }
            // a lane still running here has used the serial solve's iterations: failed
            const int first = firstSetLane(waveBallot(state == kRr3Converged));
            if (first < 0)
                return {false, 0.0, 0.0};
            return {true, waveShfl(bg, first), waveShfl(bw, first)};
        }
        else
        {
            if constexpr (Mode == kModeDefer)
            {
                if (work > kDeferBudget)
                {
                    err.raise(kErrDefer);
                    return {false, 0.0, 0.0};
                }
            }
            for (int g = 0; g < 17; ++g)
            {
                const Rr3 r = solveRr3(z, kg, kw, guesses[g][0], guesses[g][1]);
                if (r.converged)
                    return r;
            }
            return {false, 0.0, 0.0};
        }
    }

    __host__ __device__ double rr2Eval(const double *z, const double *K, double beta, double *df) const
    {
        double fv = 0.0, d = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            const double a = K[c] - 1.0;
            const double D = (1.0 - beta) + beta * K[c];
            if (!(D > 1.0e-300) || !finite(D))
                return __builtin_nan("");
            fv += z[c] * a / D;
            d -= z[c] * a * a / (D * D);
        }
        if (df) *df = d;
        return fv;
    }

    // nonlinear.inc:129-187
    __host__ __device__ Rr2 solveRr2(const double *z, const double *K, double initialBeta) const
    {
        const double f0 = rr2Eval(z, K, 0.0, nullptr);
        const double f1 = rr2Eval(z, K, 1.0, nullptr);
        if (!finite(f0) || !finite(f1))
            return {false, 0.0};
        if (f0 <= 0.0)
            return {true, 0.0};
        if (f1 >= 0.0)
            return {true, 1.0};
        double lower = 0.0, upper = 1.0;
        double beta = clampd(initialBeta, lower, upper);
        for (int it = 0; it < o.maxRrIterations; ++it)
        {
            double df = 0.0;
            const double fv = rr2Eval(z, K, beta, &df);
            if (!finite(fv) || !finite(df))
                return {false, 0.0};
            if (::fabs(fv) < o.rrTol)
                return {true, beta};
            if (fv > 0.0) lower = beta; else upper = beta;
            double cand = beta;
            if (::fabs(df) > 1.0e-24)
                cand = beta - fv / df;
            if (!finite(cand) || cand <= lower || cand >= upper)
                cand = 0.5 * (lower + upper);
            beta = cand;
        }
        return {false, 0.0};
    }

    // -------------------------------------------------------- two-phase flash
    struct TwoState
    {
        bool valid;
        double norm;
        double residual[NC];
        Result result;
    };

    // restricted.inc:71-134
    __host__ __device__ void evaluateTwoPhaseLogK(double P, double T, const double *z, int first,
                                                  int second, const double *logK, double pref,
                                                  TwoState &s)
    {
        s.valid = false;
        s.norm = __builtin_inf();
        s.result = Result();
        double K[NC];
        for (int c = 0; c < NC; ++c) K[c] = ::exp(clampd(logK[c], -230.0, 230.0));
        const Rr2 rr = solveRr2(z, K, pref);
        if (!rr.converged)
            return;
        const double bs = rr.beta;
        const double bf = 1.0 - bs;
        double a[NC], b[NC];
        for (int c = 0; c < NC; ++c)
        {
            const double D = (1.0 - bs) + bs * K[c];
            if (!(D > 0.0) || !finite(D))
                return;
            a[c] = z[c] / D;
            b[c] = K[c] * a[c];
        }
        normalize(a);
        normalize(b);
        PhaseThermo<NC> ta, tb;
        thermo(first, P, T, a, ta);
        thermo(second, P, T, b, tb);
        if (err.bad())
            return;
        s.norm = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            const double pa = maxd(ta.phi[c], 1.0e-300);
            const double pb = maxd(tb.phi[c], 1.0e-300);
            s.residual[c] = logK[c] - ::log(pa / pb);
            s.norm = maxd(s.norm, ::fabs(s.residual[c]));
        }
        Result &r = s.result;
        r.presence = uint8_t((1u << first) | (1u << second));
        r.beta[0] = r.beta[1] = r.beta[2] = 0.0;
        r.beta[first] = bf;
        r.beta[second] = bs;
        for (int c = 0; c < NC; ++c) { r.x[first][c] = a[c]; r.x[second][c] = b[c]; }
        r.Z[first] = ta.Z;
        r.Z[second] = tb.Z;
        for (int p = 0; p < 3; ++p)
            if (!(r.presence & (1u << p)))
                for (int c = 0; c < NC; ++c) r.x[p][c] = z[c];
        s.valid = finite(s.norm);
    }

    // restricted.inc:136-327
    __host__ __device__ Result flashTwo(double P, double T, const double *z, int first, int second,
                                        const double *Kin)
    {
        Result result;
        result.presence = uint8_t((1u << first) | (1u << second));
        double K[NC];
        for (int c = 0; c < NC; ++c) K[c] = clampd(Kin[c], 1.0e-100, 1.0e100);
        double betaSecond = 0.5;
        int total = 0;
        const int ssiSteps = (o.maxIterations < (o.twoPhaseSsi > 0 ? o.twoPhaseSsi : 0))
                                 ? o.maxIterations
                                 : (o.twoPhaseSsi > 0 ? o.twoPhaseSsi : 0);

        // returns false when the SSI step fails (reference's oneSsiStep)
        auto oneSsiStep = [&]() -> bool {
            const Rr2 rr = solveRr2(z, K, betaSecond);
            if (!rr.converged)
                return false;
            betaSecond = rr.beta;
            const double betaFirst = 1.0 - betaSecond;
            double a[NC], b[NC];
            for (int c = 0; c < NC; ++c)
            {
                const double D = (1.0 - betaSecond) + betaSecond * K[c];
                if (!(D > 0.0) || !finite(D))
                    return false;
                a[c] = z[c] / D;
                b[c] = K[c] * a[c];
            }
            normalize(a);
            normalize(b);
            PhaseThermo<NC> ta, tb;
            thermo(first, P, T, a, ta);
            thermo(second, P, T, b, tb);
            if (err.bad())
                return false;
            double maxResidual = 0.0;
            for (int c = 0; c < NC; ++c)
            {
                const double target = maxd(ta.phi[c], 1.0e-300) / maxd(tb.phi[c], 1.0e-300);
                const double d = clampd(::log(target / K[c]), -o.logKStep, o.logKStep);
                maxResidual = maxd(maxResidual, ::fabs(d));
                K[c] *= ::exp(0.6 * d);
                K[c] = clampd(K[c], 1.0e-100, 1.0e100);
            }
            ++total;
            NN_TRACE("two ssi pair=%d%d it=%d res=%.17g beta=%.17g\n", first, second, total, maxResidual, betaSecond);
            result.iterations = total;
            result.beta[0] = result.beta[1] = result.beta[2] = 0.0;
            result.beta[first] = betaFirst;
            result.beta[second] = betaSecond;
            for (int c = 0; c < NC; ++c) { result.x[first][c] = a[c]; result.x[second][c] = b[c]; }
            result.Z[first] = ta.Z;
            result.Z[second] = tb.Z;
            for (int p = 0; p < 3; ++p)
                if (!(result.presence & (1u << p)))
                    for (int c = 0; c < NC; ++c) result.x[p][c] = z[c];
            if (maxResidual < o.fugTol)
            {
                result.converged = true;
                result.presence = presenceFromFractions(result.beta);
                deriveReferenceK(result);
                updateSaturationsFromMoles(P, T, result);
            }
            return true;
        };

        for (int it = 0; it < ssiSteps; ++it)
        {
            if (!oneSsiStep() || err.bad())
                return result;
            if (result.converged)
                return result;
        }

        double logK[NC];
        for (int c = 0; c < NC; ++c) logK[c] = ::log(K[c]);

        TwoState base, shifted;
        for (int it = 0; it < o.maxTwoPhaseNewton; ++it)
        {
            evaluateTwoPhaseLogK(P, T, z, first, second, logK, betaSecond, base);
            if (err.bad())
                return result;
            if (!base.valid)
                break;
            result = base.result;
            result.iterations = ++total;
            betaSecond = result.beta[second];
            NN_TRACE("two newton pair=%d%d it=%d norm=%.17g\n", first, second, total, base.norm);
            if (base.norm < o.fugTol)
            {
                result.converged = true;
                result.presence = presenceFromFractions(result.beta);
                deriveReferenceK(result);
                updateSaturationsFromMoles(P, T, result);
                return result;
            }
            double jac[NC][NC];
            bool jacValid = true;
            for (int col = 0; col < NC && jacValid; ++col)
            {
                double pert[NC];
                for (int c = 0; c < NC; ++c) pert[c] = logK[c];
                const double step = o.fdStep * maxd(1.0, ::fabs(logK[col]));
                pert[col] += step;
                evaluateTwoPhaseLogK(P, T, z, first, second, pert, betaSecond, shifted);
                if (err.bad())
                    return result;
                if (!shifted.valid)
                {
                    jacValid = false;
                    break;
                }
                for (int row = 0; row < NC; ++row)
                    jac[row][col] = (shifted.residual[row] - base.residual[row]) / step;
            }
            if (!jacValid)
                break;
            double rhs[NC], delta[NC];
            for (int c = 0; c < NC; ++c) rhs[c] = -base.residual[c];
            if (!solveDense<NC>(jac, rhs, delta))
                break;
            bool accepted = false;
            double damping = 1.0;
            for (int ls = 0; ls < 16; ++ls)
            {
                double trialLogK[NC];
                for (int c = 0; c < NC; ++c) trialLogK[c] = logK[c] + damping * delta[c];
                evaluateTwoPhaseLogK(P, T, z, first, second, trialLogK, betaSecond, shifted);
                if (err.bad())
                    return result;
                if (shifted.valid && shifted.norm < base.norm)
                {
                    for (int c = 0; c < NC; ++c) logK[c] = trialLogK[c];
                    accepted = true;
                    break;
                }
                damping *= 0.5;
            }
            if (!accepted)
                break;
        }

        for (int it = ssiSteps; it < o.maxIterations; ++it)
        {
            for (int c = 0; c < NC; ++c) K[c] = ::exp(clampd(logK[c], -230.0, 230.0));
            if (!oneSsiStep() || err.bad())
                return result;
            if (result.converged)
                return result;
            for (int c = 0; c < NC; ++c) logK[c] = ::log(K[c]);
        }
        return result;
    }

    // ------------------------------------------------------ three-phase flash
    struct ThreeState
    {
        bool valid;
        double norm;
        double residual[K2];
        Result result;
    };

    // nonlinear.inc:200-277
    __host__ __device__ void evaluateThreePhaseLogK(double P, double T, const double *z,
                                                    const double *logK, double pg, double pw,
                                                    ThreeState &s)
    {
        s.valid = false;
        s.norm = __builtin_inf();
        s.result = Result();
        double kg[NC], kw[NC];
        for (int c = 0; c < NC; ++c)
        {
            kg[c] = ::exp(clampd(logK[c], -230.0, 230.0));
            kw[c] = ::exp(clampd(logK[NC + c], -230.0, 230.0));
        }
        const Rr3 rr = solveRr3Robust(z, kg, kw, pg, pw);
        if (!rr.converged)
            return;
        const double bg = rr.bg, bw = rr.bw, bo = 1.0 - bg - bw;
        if (!(bo >= -o.betaTol) || !(bg >= -o.betaTol) || !(bw >= -o.betaTol))
            return;
        double xo[NC], yg[NC], xw[NC];
        for (int c = 0; c < NC; ++c)
        {
            const double D = bo + bg * kg[c] + bw * kw[c];
            if (!(D > 0.0) || !finite(D))
                return;
            xo[c] = z[c] / D;
            yg[c] = kg[c] * xo[c];
            xw[c] = kw[c] * xo[c];
        }
        normalize(xo);
        normalize(yg);
        normalize(xw);
        PhaseThermo<NC> to, tg, tw;
        thermo(kOil, P, T, xo, to);
        thermo(kGas, P, T, yg, tg);
        thermo(kWater, P, T, xw, tw);
        if (err.bad())
            return;
        s.norm = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            const double po = maxd(to.phi[c], 1.0e-300);
            const double pgp = maxd(tg.phi[c], 1.0e-300);
            const double pwp = maxd(tw.phi[c], 1.0e-300);
            s.residual[c] = logK[c] - ::log(po / pgp);
            s.residual[NC + c] = logK[NC + c] - ::log(po / pwp);
            s.norm = maxd(maxd(s.norm, ::fabs(s.residual[c])), ::fabs(s.residual[NC + c]));
        }
        Result &r = s.result;
        r.presence = kAllBits;
        r.beta[0] = bo; r.beta[1] = bg; r.beta[2] = bw;
        for (int c = 0; c < NC; ++c)
        {
            r.x[0][c] = xo[c]; r.x[1][c] = yg[c]; r.x[2][c] = xw[c];
            r.kg[c] = kg[c]; r.kw[c] = kw[c];
        }
        r.Z[0] = to.Z; r.Z[1] = tg.Z; r.Z[2] = tw.Z;
        s.valid = finite(s.norm);
    }

    // nonlinear.inc:279-513
    __host__ __device__ Result flashThree(double P, double T, const double *z, const double *kgIn,
                                          const double *kwIn)
    {
        Result result;
        result.presence = kAllBits;
        double kg[NC], kw[NC];
        for (int c = 0; c < NC; ++c)
        {
            kg[c] = clampd(kgIn[c], 1.0e-100, 1.0e100);
            kw[c] = clampd(kwIn[c], 1.0e-100, 1.0e100);
        }
        double bg = 0.15, bw = 0.15;
        const int ssiSteps = (o.maxIterations < (o.threePhaseSsi > 0 ? o.threePhaseSsi : 0))
                                 ? o.maxIterations
                                 : (o.threePhaseSsi > 0 ? o.threePhaseSsi : 0);
        int total = 0;

        auto oneSsiStep = [&]() -> bool {
            const Rr3 rr = solveRr3Robust(z, kg, kw, bg, bw);
            if (!rr.converged)
                return false;
            bg = rr.bg;
            bw = rr.bw;
            const double bo = 1.0 - bg - bw;
            double xo[NC], yg[NC], xw[NC];
            for (int c = 0; c < NC; ++c)
            {
                const double D = bo + bg * kg[c] + bw * kw[c];
                if (!(D > 0.0) || !finite(D))
                    return false;
                xo[c] = z[c] / D;
                yg[c] = kg[c] * xo[c];
                xw[c] = kw[c] * xo[c];
            }
            normalize(xo);
            normalize(yg);
            normalize(xw);
            PhaseThermo<NC> to, tg, tw;
            thermo(kOil, P, T, xo, to);
            thermo(kGas, P, T, yg, tg);
            thermo(kWater, P, T, xw, tw);
            if (err.bad())
                return false;
            double maxResidual = 0.0;
            for (int c = 0; c < NC; ++c)
            {
                const double tkg = maxd(to.phi[c], 1.0e-300) / maxd(tg.phi[c], 1.0e-300);
                const double tkw = maxd(to.phi[c], 1.0e-300) / maxd(tw.phi[c], 1.0e-300);
                const double dg = clampd(::log(tkg / kg[c]), -o.logKStep, o.logKStep);
                const double dw = clampd(::log(tkw / kw[c]), -o.logKStep, o.logKStep);
                maxResidual = maxd(maxd(maxResidual, ::fabs(dg)), ::fabs(dw));
                kg[c] *= ::exp(0.6 * dg);
                kw[c] *= ::exp(0.6 * dw);
                kg[c] = clampd(kg[c], 1.0e-100, 1.0e100);
                kw[c] = clampd(kw[c], 1.0e-100, 1.0e100);
            }
            ++total;
            result.iterations = total;
            result.beta[0] = bo; result.beta[1] = bg; result.beta[2] = bw;
            for (int c = 0; c < NC; ++c)
            {
                result.x[0][c] = xo[c]; result.x[1][c] = yg[c]; result.x[2][c] = xw[c];
                result.kg[c] = kg[c]; result.kw[c] = kw[c];
            }
            result.Z[0] = to.Z; result.Z[1] = tg.Z; result.Z[2] = tw.Z;
            if (maxResidual < o.fugTol)
            {
                result.converged = true;
                result.presence = presenceFromFractions(result.beta);
                updateSaturationsFromMoles(P, T, result);
            }
            return true;
        };

        for (int it = 0; it < ssiSteps; ++it)
        {
            if (!oneSsiStep() || err.bad())
                return result;
            if (result.converged)
                return result;
        }

        double logK[K2];
        for (int c = 0; c < NC; ++c)
        {
            logK[c] = ::log(kg[c]);
            logK[NC + c] = ::log(kw[c]);
        }

        ThreeState base, side;
        for (int it = 0; it < o.maxThreePhaseNewton; ++it)
        {
            evaluateThreePhaseLogK(P, T, z, logK, bg, bw, base);
            if (err.bad())
                return result;
            if (!base.valid)
                break;
            result = base.result;
            result.iterations = ++total;
            bg = result.beta[1];
            bw = result.beta[2];
            if (base.norm < o.fugTol)
            {
                result.converged = true;
                result.presence = presenceFromFractions(result.beta);
                updateSaturationsFromMoles(P, T, result);
                return result;
            }
            double jac[K2][K2];
            bool jacValid = true;
            for (int col = 0; col < K2; ++col)
            {
                const double step = o.fdStep * maxd(1.0, ::fabs(logK[col]));
                double pert[K2];
                for (int k = 0; k < K2; ++k) pert[k] = logK[k];
                pert[col] += step;
                evaluateThreePhaseLogK(P, T, z, pert, bg, bw, side);
                if (err.bad())
                    return result;
                if (side.valid)
                {
                    for (int row = 0; row < K2; ++row)
                        jac[row][col] = (side.residual[row] - base.residual[row]) / step;
                    continue;
                }
                for (int k = 0; k < K2; ++k) pert[k] = logK[k];
                pert[col] -= step;
                evaluateThreePhaseLogK(P, T, z, pert, bg, bw, side);
                if (err.bad())
                    return result;
                if (!side.valid)
                {
                    jacValid = false;
                    break;
                }
                for (int row = 0; row < K2; ++row)
                    jac[row][col] = (base.residual[row] - side.residual[row]) / step;
            }
            if (!jacValid)
                break;
            double rhs[K2], delta[K2];
            for (int k = 0; k < K2; ++k) rhs[k] = -base.residual[k];
            if (!solveDense<K2>(jac, rhs, delta))
                break;
            bool accepted = false;
            double damping = 1.0;
            for (int ls = 0; ls < 16; ++ls)
            {
                double trial[K2];
                for (int k = 0; k < K2; ++k) trial[k] = logK[k] + damping * delta[k];
                evaluateThreePhaseLogK(P, T, z, trial, bg, bw, side);
                if (err.bad())
                    return result;
                if (side.valid && side.norm < base.norm)
                {
                    for (int k = 0; k < K2; ++k) logK[k] = trial[k];
                    accepted = true;
                    break;
                }
                damping *= 0.5;
            }
            if (!accepted)
                break;
        }

        for (int it = ssiSteps; it < o.maxIterations; ++it)
        {
            for (int c = 0; c < NC; ++c)
            {
                kg[c] = ::exp(clampd(logK[c], -230.0, 230.0));
                kw[c] = ::exp(clampd(logK[NC + c], -230.0, 230.0));
            }
            if (!oneSsiStep() || err.bad())
                return result;
            if (result.converged)
                return result;
            for (int c = 0; c < NC; ++c)
            {
                logK[c] = ::log(kg[c]);
                logK[NC + c] = ::log(kw[c]);
            }
        }
        return result;
    }

    // nonlinear.inc:515-622
    __host__ __device__ void pairInitialK(int first, int second, double P, double T,
                                          const double *z, double *K)
    {
        if (sw() && (first == kWater || second == kWater))
        {
            const int water = f.water >= 0 ? f.water : NC - 1;
            double seeds[2][NC];
            const int which[2] = {first, second};
            for (int s = 0; s < 2; ++s)
            {
                double *seed = seeds[s];
                for (int c = 0; c < NC; ++c) seed[c] = z[c];
                if (which[s] == kWater)
                {
                    double nonWater = 0.0;
                    for (int i = 0; i < NC; ++i)
                    {
                        if (i == water) continue;
                        seed[i] = maxd(z[i], o.floor);
                        nonWater += seed[i];
                    }
                    const double dissolved = 0.02;
                    if (nonWater > 0.0)
                        for (int i = 0; i < NC; ++i)
                            if (i != water)
                                seed[i] *= dissolved / nonWater;
                    seed[water] = 1.0 - dissolved;
                }
                else
                {
                    const double waterSeed = 0.01;
                    double nonWater = 0.0;
                    for (int i = 0; i < NC; ++i)
                        if (i != water)
                            nonWater += maxd(z[i], o.floor);
                    for (int i = 0; i < NC; ++i)
                    {
                        if (i == water)
                            seed[i] = waterSeed;
                        else
                            seed[i] = nonWater > 0.0
                                          ? (1.0 - waterSeed) * maxd(z[i], o.floor) / nonWater
                                          : (1.0 - waterSeed) / double(NC - 1 > 1 ? NC - 1 : 1);
                    }
                }
                normalize(seed);
            }
            PhaseThermo<NC> ft, st;
            thermo(first, P, T, seeds[0], ft);
            thermo(second, P, T, seeds[1], st);
            for (int c = 0; c < NC; ++c)
                K[c] = clampd(maxd(ft.phi[c], 1.0e-300) / maxd(st.phi[c], 1.0e-300),
                              1.0e-100, 1.0e100);
            return;
        }
        double kg[NC], kw[NC];
        wilsonK(P, T, kg);
        initialWaterOilK(kw);
        for (int c = 0; c < NC; ++c)
        {
            auto rel = [&](int p) {
                if (p == kOil) return 1.0;
                if (p == kGas) return kg[c];
                return kw[c];
            };
            K[c] = rel(second) / rel(first);
        }
    }

    // -------------------------------------------------------- canonicalization
    // canonicalization.inc:100-134
    __host__ __device__ void canonicalizeLiquidRoles(Result &r) const
    {
        if (f.water < 0)
            return;
        const bool hasOil = r.beta[0] > o.betaTol;
        const bool hasWater = r.beta[2] > o.betaTol;
        bool swap = false;
        if (hasOil && hasWater)
            swap = r.x[0][f.water] > r.x[2][f.water];
        else if (hasOil != hasWater)
        {
            const int act = hasOil ? 0 : 2;
            const bool shouldBeWater = waterIsDominant(r.x[act]);
            swap = (hasOil && shouldBeWater) || (hasWater && !shouldBeWater);
        }
        if (swap)
            swapSlots(r, 0, 2);
    }

    // canonicalization.inc:47-65
    __host__ __device__ void refreshCanonicalThermodynamics(double P, double T, Result &r)
    {
        for (int p = 0; p < 3; ++p)
        {
            if (!(r.presence & (1u << p)))
                continue;
            normalize(r.x[p]);
            PhaseThermo<NC> t;
            thermo(p, P, T, r.x[p], t);
            if (err.bad())
                return;
            r.Z[p] = t.Z;
        }
        deriveReferenceK(r);
        updateSaturationsFromMoles(P, T, r);
    }

    // canonicalization.inc:8-45
    __host__ __device__ void canonicalize(double P, double T, Result &r)
    {
        uint8_t p = presenceFromFractions(r.beta);
        if (p == 0)
            p = r.presence == 0 ? kOilBit : r.presence;
        r.presence = p;
        double sum = 0.0;
        for (int i = 0; i < 3; ++i)
        {
            double &beta = r.beta[i];
            if (!(p & (1u << i)) || beta <= o.betaTol || !finite(beta))
            {
                beta = 0.0;
                r.presence &= uint8_t(~(1u << i));
            }
            else
                sum += beta;
        }
        if (!(sum > 0.0))
        {
            r.presence = chooseSinglePhase(P, T, r.x[0]);
            r.beta[0] = r.beta[1] = r.beta[2] = 0.0;
            r.beta[firstActive(r.presence)] = 1.0;
        }
        else
        {
            for (int i = 0; i < 3; ++i) r.beta[i] /= sum;
        }
        if (!sw())
            canonicalizeLiquidRoles(r);
        r.presence = presenceFromFractions(r.beta);
        refreshCanonicalThermodynamics(P, T, r);
    }

    // canonicalization.inc:176-222
    __host__ __device__ int canonicalizeGlobalWaterRole(Result &r) const
    {
        if (f.water < 0)
            return -1;
        const int wc = f.water;
        int richest = 2;
        double richestWater = -1.0;
        int activeCount = 0;
        bool hasWaterSlot = false;
        for (int s = 0; s < 3; ++s)
        {
            if (!(r.beta[s] > o.betaTol))
                continue;
            ++activeCount;
            if (s == 2)
                hasWaterSlot = true;
            const double xw = r.x[s][wc];
            if (xw > richestWater)
            {
                richestWater = xw;
                richest = s;
            }
        }
        if (activeCount == 0 || richest == 2)
            return -1;
        if (!hasWaterSlot && !waterIsDominant(r.x[richest]))
            return -1;
        const double currentWater = hasWaterSlot ? r.x[2][wc] : -1.0;
        const double relTol = 1.0e-10;
        if (hasWaterSlot && richestWater <= currentWater + relTol * maxd(1.0, ::fabs(richestWater)))
            return -1;
        swapSlots(r, richest, 2);
        return richest;
    }

    // canonicalization.inc:225-321
    __host__ __device__ bool validatePhysical(double P, double T, const double *z, const Result &r)
    {
        if (!r.converged)
            return false;
        const double closureTol = maxd(maxd(10.0 * o.fugTol, 10.0 * o.rrTol), 1.0e-10);
        double betaSum = 0.0;
        uint8_t active = 0;
        for (int p = 0; p < 3; ++p)
        {
            const double beta = r.beta[p];
            if (!finite(beta) || beta < -closureTol)
                return false;
            if (beta > o.betaTol)
            {
                active |= uint8_t(1u << p);
                betaSum += beta;
                double xsum = 0.0;
                for (int c = 0; c < NC; ++c)
                {
                    const double xv = r.x[p][c];
                    if (!finite(xv) || xv < -closureTol)
                        return false;
                    xsum += xv;
                }
                if (::fabs(xsum - 1.0) > closureTol)
                    return false;
            }
        }
        if (active == 0 || ::fabs(betaSum - 1.0) > closureTol)
            return false;
        for (int c = 0; c < NC; ++c)
        {
            double rec = 0.0;
            for (int p = 0; p < 3; ++p)
                if (active & (1u << p))
                    rec += r.beta[p] * r.x[p][c];
            if (::fabs(rec - z[c]) > closureTol)
                return false;
        }
        if (popcount3(active) <= 1)
            return true;
        const int ref = firstActive(active);
        PhaseThermo<NC> rt;
        thermoForRole(ref, P, T, r.x[ref], rt);
        if (err.bad())
            return false;
        for (int p = 0; p < 3; ++p)
        {
            if (!(active & (1u << p)) || p == ref)
                continue;
            PhaseThermo<NC> t;
            thermoForRole(p, P, T, r.x[p], t);
            if (err.bad())
                return false;
            for (int c = 0; c < NC; ++c)
            {
                if (r.x[ref][c] <= 10.0 * o.floor || r.x[p][c] <= 10.0 * o.floor)
                    continue;
                const double ra = maxd(r.x[ref][c] * rt.phi[c], 1.0e-300);
                const double a = maxd(r.x[p][c] * t.phi[c], 1.0e-300);
                if (::fabs(::log(ra / a)) > 10.0 * o.fugTol)
                    return false;
            }
        }
        return true;
    }

    // canonicalization.inc:323-337
    __host__ __device__ void overallComposition(const Result &r, double *z)
    {
        for (int c = 0; c < NC; ++c) z[c] = 0.0;
        for (int s = 0; s < 3; ++s)
        {
            const double beta = r.beta[s];
            if (!(beta > o.betaTol))
                continue;
            for (int c = 0; c < NC; ++c) z[c] += beta * r.x[s][c];
        }
        normalize(z);
    }

    // canonicalization.inc:340-379
    __host__ __device__ Result reequilibrateCanonicalRoles(double P, double T, const double *z, Result seed)
    {
        seed.presence = presenceFromFractions(seed.beta);
        const int count = popcount3(seed.presence);
        if (count <= 1)
            return singlePhaseResult(P, T, z, seed.presence);
        if (count == 2)
        {
            int a, b;
            activePair(seed.presence, a, b);
            double K[NC];
            for (int c = 0; c < NC; ++c)
                K[c] = clampd(seed.x[b][c] / maxd(seed.x[a][c], o.floor), 1.0e-100, 1.0e100);
            return flashTwo(P, T, z, a, b, K);
        }
        deriveReferenceK(seed);
        Result r = flashThree(P, T, z, seed.kg, seed.kw);
        if (r.converged || err.bad())
            return r;
        double kg[NC], kw[NC];
        wilsonK(P, T, kg);
        initialWaterOilK(kw);
        return flashThree(P, T, z, kg, kw);
    }

    __host__ __device__ double wilsonLogK(double P, double T, int c) const
    {
        const double lnK = ::log(f.pc[c] / P) + 5.373 * (1.0 + f.omega[c]) * (1.0 - f.tc[c] / T);
        const double K = clampd(::exp(lnK), 1.0e-30, 1.0e30);
        return ::log(clampd(K, 1.0e-30, 1.0e30));
    }

    // canonicalization.inc:381-414
    __host__ __device__ bool canonicalizeNonAqueousRoles(double P, double T, Result &r) const
    {
        const bool hasOil = r.beta[0] > o.betaTol;
        const bool hasGas = r.beta[1] > o.betaTol;
        if (!(hasOil && hasGas))
            return false;
        double orientation = 0.0, magnitude = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            const double lnK = wilsonLogK(P, T, c);
            const double term = (r.x[1][c] - r.x[0][c]) * lnK;
            orientation += term;
            magnitude += ::fabs(term);
        }
        if (orientation >= -1.0e-10 * maxd(1.0, magnitude))
            return false;
        swapSlots(r, 0, 1);
        return true;
    }

    // canonicalization.inc:416-438
    __host__ __device__ int expectedSingleNonAqueousRole(double P, double T, const double *x) const
    {
        double vol = 0.0, mag = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            const double term = x[c] * wilsonLogK(P, T, c);
            vol += term;
            mag += ::fabs(term);
        }
        return vol >= -1.0e-10 * maxd(1.0, mag) ? kGas : kOil;
    }

    // canonicalization.inc:440-464
    __host__ __device__ bool canonicalizeSingleNonAqueousRole(double P, double T, Result &r) const
    {
        const bool hasOil = r.beta[0] > o.betaTol;
        const bool hasGas = r.beta[1] > o.betaTol;
        if (hasOil == hasGas)
            return false;
        const int act = hasOil ? 0 : 1;
        const int expected = expectedSingleNonAqueousRole(P, T, r.x[act]);
        if (act == expected)
            return false;
        swapSlots(r, act, expected);
        r.presence = presenceFromFractions(r.beta);
        return true;
    }

    __host__ __device__ bool rootsCoalesced(double P, double T, const double *x)
    {
        PhaseThermo<NC> ot, gt;
        thermoForRole(kOil, P, T, x, ot);
        thermoForRole(kGas, P, T, x, gt);
        const double scale = maxd(maxd(1.0, ::fabs(ot.Z)), ::fabs(gt.Z));
        return ::fabs(ot.Z - gt.Z) <= 1.0e-8 * scale;
    }

    // canonicalization.inc:467-503
    __host__ __device__ void canonicalizeCoalescedSwSingleNonAqueousRole(double P, double T, Result &r)
    {
        if (!sw())
            return;
        const bool hasOil = r.beta[0] > o.betaTol;
        const bool hasGas = r.beta[1] > o.betaTol;
        if (hasOil == hasGas)
            return;
        const int act = hasOil ? 0 : 1;
        if (!rootsCoalesced(P, T, r.x[act]))
            return;
        (void)canonicalizeSingleNonAqueousRole(P, T, r);
    }

    // canonicalization.inc:513-600
    __host__ __device__ void collapseCoincidentSwPhases(double P, double T, Result &r)
    {
        if (!sw() || !(o.coincidentTol > 0.0))
            return;
        const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        bool changed = false;
        for (int pass = 0; pass < 2; ++pass)
        {
            bool merged = false;
            for (int k = 0; k < 3; ++k)
            {
                const int a = pairs[k][0], b = pairs[k][1];
                const double ba = r.beta[a], bb = r.beta[b];
                if (!(ba > o.betaTol) || !(bb > o.betaTol))
                    continue;
                double maxDiff = 0.0;
                for (int c = 0; c < NC; ++c)
                    maxDiff = maxd(maxDiff, ::fabs(r.x[a][c] - r.x[b][c]));
                if (maxDiff > o.coincidentTol)
                    continue;
                const double combined = ba + bb;
                double comb[NC];
                for (int c = 0; c < NC; ++c)
                    comb[c] = (ba * r.x[a][c] + bb * r.x[b][c]) / combined;
                int dest = a;
                if (a == 2 || b == 2)
                    dest = 2;
                else
                    dest = expectedSingleNonAqueousRole(P, T, comb);
                const int src = dest == a ? b : a;
                r.beta[dest] = combined;
                r.beta[src] = 0.0;
                for (int c = 0; c < NC; ++c) { r.x[dest][c] = comb[c]; r.x[src][c] = comb[c]; }
                r.sat[src] = 0.0;
                r.rho[src] = 0.0;
                r.Z[src] = 1.0;
                merged = true;
                changed = true;
                break;
            }
            if (!merged)
                break;
        }
        if (!changed)
            return;
        r.presence = presenceFromFractions(r.beta);
        refreshCanonicalThermodynamics(P, T, r);
    }

    // canonicalization.inc:602-704
    __host__ __device__ bool swRolesConsistent(double P, double T, const Result &r, bool preserve)
    {
        if (!sw() || f.water < 0)
            return true;
        const int wc = f.water;
        int activeCount = 0, aqueousSlot = -1, oilSlot = -1, gasSlot = -1, richestSlot = -1;
        double richestWater = -1.0;
        for (int s = 0; s < 3; ++s)
        {
            if (!(r.beta[s] > o.betaTol))
                continue;
            ++activeCount;
            const int role = roles[s];
            if (role == kWater) aqueousSlot = s;
            if (role == kOil) oilSlot = s;
            if (role == kGas) gasSlot = s;
            const double xw = r.x[s][wc];
            if (xw > richestWater)
            {
                richestWater = xw;
                richestSlot = s;
            }
        }
        if (activeCount == 0)
            return false;
        if (aqueousSlot >= 0)
        {
            if (richestSlot != aqueousSlot &&
                richestWater > r.x[aqueousSlot][wc] + 1.0e-10 * maxd(1.0, ::fabs(richestWater)))
                return false;
        }
        else if (richestSlot >= 0 && waterIsDominant(r.x[richestSlot]))
        {
            return false;
        }
        if ((oilSlot >= 0) != (gasSlot >= 0))
        {
            const int slot = oilSlot >= 0 ? oilSlot : gasSlot;
            const double *x = r.x[slot];
            bool coalesced = false;
            if (preserve)
                coalesced = rootsCoalesced(P, T, x);
            if (err.bad())
                return false;
            if (!coalesced)
            {
                double pseudoTc = 0.0;
                for (int c = 0; c < NC; ++c) pseudoTc += x[c] * f.tc[c];
                const double scale = maxd(maxd(1.0, ::fabs(T)), ::fabs(pseudoTc));
                if (::fabs(T - pseudoTc) > 1.0e-8 * scale)
                {
                    const int expected = T > pseudoTc ? kGas : kOil;
                    if (roles[slot] != expected)
                        return false;
                }
            }
        }
        return true;
    }

    // canonicalization.inc:706-730
    __host__ __device__ void canonicalizeSwThermodynamicRoles(Result &r) const
    {
        Result c = r;
        c.beta[0] = c.beta[1] = c.beta[2] = 0.0;
        c.sat[0] = c.sat[1] = c.sat[2] = 0.0;
        c.Z[0] = c.Z[1] = c.Z[2] = 1.0;
        c.rho[0] = c.rho[1] = c.rho[2] = 0.0;
        for (int p = 0; p < 3; ++p)
            for (int k = 0; k < NC; ++k) c.x[p][k] = 0.0;
        for (int s = 0; s < 3; ++s)
        {
            const int phys = roles[s];
            c.beta[phys] = r.beta[s];
            c.sat[phys] = r.sat[s];
            for (int k = 0; k < NC; ++k) c.x[phys][k] = r.x[s][k];
            c.Z[phys] = r.Z[s];
            c.rho[phys] = r.rho[s];
        }
        c.presence = presenceFromFractions(c.beta);
        deriveReferenceK(c);
        r = c;
    }

    // canonicalization.inc:732-810
    __host__ __device__ void canonicalizeFinal(double P, double T, Result &r, bool preserve = false)
    {
        NN_TRACE("final in conv=%d pres=%d beta=%.17g %.17g %.17g\n", (int)r.converged, (int)r.presence, r.beta[0], r.beta[1], r.beta[2]);
        if (!r.converged)
            return;
        if (sw())
        {
            if (!swRolesConsistent(P, T, r, preserve))
            {
                r.converged = false;
                return;
            }
            canonicalizeSwThermodynamicRoles(r);
            collapseCoincidentSwPhases(P, T, r);
            if (!preserve)
                canonicalizeCoalescedSwSingleNonAqueousRole(P, T, r);
            (void)canonicalizeNonAqueousRoles(P, T, r);
            r.presence = presenceFromFractions(r.beta);
            return;
        }
        for (int pass = 0; pass < 4; ++pass)
        {
            if (err.bad())
                return;
            r.presence = presenceFromFractions(r.beta);
            double z[NC];
            overallComposition(r, z);
            const int waterSource = canonicalizeGlobalWaterRole(r);
            const bool crossedGasWater = waterSource == kGas;
            const bool crossedSingle = !preserve && canonicalizeSingleNonAqueousRole(P, T, r);
            if (crossedGasWater || crossedSingle)
            {
                r.presence = presenceFromFractions(r.beta);
                Result corrected = reequilibrateCanonicalRoles(P, T, z, r);
                if (!corrected.converged)
                {
                    r.converged = false;
                    return;
                }
                r = corrected;
                continue;
            }
            (void)canonicalizeNonAqueousRoles(P, T, r);
            r.presence = presenceFromFractions(r.beta);
            refreshCanonicalThermodynamics(P, T, r);
            return;
        }
        r.converged = false;
    }

    // ---------------------------------------------------------- stability
    struct Trial
    {
        bool valid = false;
        bool trivial = false;
        double sum = 0.0;
        double comp[NC];
    };

    struct StabResidual
    {
        bool valid;
        double norm;
        double sum;
        double comp[NC];
        double residual[NC];
        double phi[NC];
    };

    // stability.inc:157-204 (phi retained so the SSI update reuses it)
    __host__ __device__ void evaluateStabilityResidual(int candidate, double P, double T,
                                                       const double *logRef, const double *logW,
                                                       StabResidual &s)
    {
        s.valid = false;
        s.norm = __builtin_inf();
        s.sum = 0.0;
        double maxLogW = logW[0];
        for (int c = 1; c < NC; ++c)
            if (maxLogW < logW[c])
                maxLogW = logW[c];
        if (!finite(maxLogW))
            return;
        double scaledSum = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            const double v = ::exp(logW[c] - maxLogW);
            if (!(v > 0.0) || !finite(v))
                return;
            s.comp[c] = v;
            scaledSum += v;
        }
        if (!(scaledSum > 0.0) || !finite(scaledSum))
            return;
        for (int c = 0; c < NC; ++c) s.comp[c] /= scaledSum;
        const double logSum = maxLogW + ::log(scaledSum);
        if (logSum > kLogDblMax)
            s.sum = __builtin_inf();
        else
            s.sum = ::exp(logSum);
        PhaseThermo<NC> t;
        thermo(candidate, P, T, s.comp, t);
        if (err.bad())
            return;
        double norm = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            const double phi = t.phi[c];
            if (!(phi > 0.0) || !finite(phi))
            {
                s.norm = __builtin_inf();
                s.sum = 0.0;
                return;
            }
            s.phi[c] = phi;
            s.residual[c] = logW[c] + ::log(phi) - logRef[c];
            norm = maxd(norm, ::fabs(s.residual[c]));
        }
        s.norm = norm;
        s.valid = finite(norm);
    }

    __host__ __device__ Trial finalizeTrial(const StabResidual &s, const double *ref) const
    {
        Trial t;
        if (!s.valid)
            return t;
        double dist = 0.0;
        for (int c = 0; c < NC; ++c) dist += ::fabs(s.comp[c] - ref[c]);
        t.valid = true;
        t.trivial = dist < 1.0e-7 && ::fabs(s.sum - 1.0) < 1.0e-7;
        t.sum = s.sum;
        for (int c = 0; c < NC; ++c) t.comp[c] = s.comp[c];
        return t;
    }

    // stability.inc:225-355
    __host__ __device__ Trial stabilityTrial(int candidate, double P, double T, const double *logRef,
                                             const double *refComp, double *seed)
    {
        normalize(seed);
        double logW[NC];
        for (int c = 0; c < NC; ++c) logW[c] = ::log(maxd(seed[c], o.floor));
        const int ssi = o.maxStabilityIterations < (o.stabilitySsi > 0 ? o.stabilitySsi : 0)
                            ? o.maxStabilityIterations
                            : (o.stabilitySsi > 0 ? o.stabilitySsi : 0);
        StabResidual cur, shifted;
        for (int it = 0; it < ssi; ++it)
        {
            evaluateStabilityResidual(candidate, P, T, logRef, logW, cur);
            if (err.bad() || !cur.valid)
                return Trial();
            if (cur.norm < o.stabTol)
                return finalizeTrial(cur, refComp);
            for (int c = 0; c < NC; ++c)
                logW[c] = logRef[c] - ::log(maxd(cur.phi[c], 1.0e-300));
        }
        for (int it = 0; it < o.maxStabilityNewton; ++it)
        {
            evaluateStabilityResidual(candidate, P, T, logRef, logW, cur);
            if (err.bad() || !cur.valid)
                return Trial();
            if (cur.norm < o.stabTol)
                return finalizeTrial(cur, refComp);
            double jac[NC][NC];
            for (int col = 0; col < NC; ++col)
            {
                double pert[NC];
                for (int c = 0; c < NC; ++c) pert[c] = logW[c];
                const double step = o.fdStep * maxd(1.0, ::fabs(logW[col]));
                pert[col] += step;
                evaluateStabilityResidual(candidate, P, T, logRef, pert, shifted);
                if (err.bad() || !shifted.valid)
                    return Trial();
                for (int row = 0; row < NC; ++row)
                    jac[row][col] = (shifted.residual[row] - cur.residual[row]) / step;
            }
            double rhs[NC], delta[NC];
            for (int c = 0; c < NC; ++c) rhs[c] = -cur.residual[c];
            if (!solveDense<NC>(jac, rhs, delta))
                break;
            bool accepted = false;
            double damping = 1.0;
            for (int ls = 0; ls < 16; ++ls)
            {
                double cand[NC];
                for (int c = 0; c < NC; ++c) cand[c] = logW[c] + damping * delta[c];
                evaluateStabilityResidual(candidate, P, T, logRef, cand, shifted);
                if (err.bad())
                    return Trial();
                if (shifted.valid && shifted.norm < cur.norm)
                {
                    for (int c = 0; c < NC; ++c) logW[c] = cand[c];
                    accepted = true;
                    break;
                }
                damping *= 0.5;
            }
            if (!accepted)
                break;
        }
        const int remaining = o.maxStabilityIterations - ssi > 0 ? o.maxStabilityIterations - ssi : 0;
        for (int it = 0; it < remaining; ++it)
        {
            evaluateStabilityResidual(candidate, P, T, logRef, logW, cur);
            if (err.bad() || !cur.valid)
                return Trial();
            if (cur.norm < o.stabTol)
                return finalizeTrial(cur, refComp);
            for (int c = 0; c < NC; ++c)
                logW[c] = logRef[c] - ::log(maxd(cur.phi[c], 1.0e-300));
        }
        return Trial();
    }

    // stability.inc:32-96: seed k of the NC+7 multistart set; returns false past the end.
    __host__ __device__ bool stabilitySeed(int k, double P, double T, const double *z,
                                           const double *refComp, double *seed)
    {
        if (k < 4)
        {
            const double exps[4] = {1.0, -1.0, 1.0 / 3.0, -1.0 / 3.0};
            double w[NC];
            wilsonK(P, T, w);
            for (int c = 0; c < NC; ++c)
            {
                const double kk = clampd(w[c], 1.0e-100, 1.0e100);
                seed[c] = z[c] * ::pow(kk, exps[k]);
            }
            normalize(seed);
            return true;
        }
        k -= 4;
        if (k < NC)
        {
            if (NC == 1)
            {
                seed[0] = 1.0;
                return true;
            }
            const double rem = 0.1 / double(NC - 1);
            for (int c = 0; c < NC; ++c) seed[c] = rem;
            seed[k] = 0.9;
            return true;
        }
        k -= NC;
        if (k == 0)
        {
            for (int c = 0; c < NC; ++c) seed[c] = z[c];
            normalize(seed);
            return true;
        }
        if (k == 1)
        {
            for (int c = 0; c < NC; ++c) seed[c] = refComp[c];
            normalize(seed);
            return true;
        }
        if (k == 2 && f.water >= 0 && NC > 1)
        {
            const double trace = 1.0e-6 / double(NC - 1);
            for (int c = 0; c < NC; ++c) seed[c] = trace;
            seed[f.water] = 1.0 - 1.0e-6;
            return true;
        }
        return false;
    }

    // ------------------------------------------- lockstep seed trials
    // Every seed trial of a recorded stability test (all missing candidates,
    // all seeds, in the serial order) as a per-lane state machine.  One step of
    // the machine is one residual evaluation at the point trialPoint() names,
    // so a wavefront's lanes evaluate the EOS together whatever trial, seed or
    // stage each lane is in; trialAdvance() then moves the lane on exactly as
    // stabilityTrial's loops would (same operations, same order, hence the same
    // Trial per seed and the same best per candidate).
    enum : int
    {
        kTrialSsi = 0,
        kTrialNewtonBase = 1,
        kTrialNewtonFd = 2,
        kTrialLineSearch = 3,
        kTrialSsiTail = 4,
    };

    struct TrialMachine
    {
        int cand, k, stage, it, col, ls;
        double damping;
        double logW[NC], delta[NC];
        double jac[NC][NC];
        StabResidual cur;
        Trial best;
    };

    __host__ __device__ int trialSsiSteps() const
    {
        return o.maxStabilityIterations < (o.stabilitySsi > 0 ? o.stabilitySsi : 0)
                   ? o.maxStabilityIterations
                   : (o.stabilitySsi > 0 ? o.stabilitySsi : 0);
    }

    // Next stage after the SSI stage (or after Newton stops): returns false when
    // the trial ends there without another evaluation (an invalid Trial).
    __host__ __device__ bool trialEnterNewton(TrialMachine &m) const
    {
        if (o.maxStabilityNewton > 0)
        {
            m.stage = kTrialNewtonBase;
            m.it = 0;
            return true;
        }
        return trialEnterTail(m);
    }

    __host__ __device__ bool trialEnterTail(TrialMachine &m) const
    {
        const int remaining = o.maxStabilityIterations - trialSsiSteps() > 0
                                  ? o.maxStabilityIterations - trialSsiSteps()
                                  : 0;
        m.stage = kTrialSsiTail;
        m.it = 0;
        return remaining > 0;
    }

    // Starts seed m.k of candidate m.cand; false when no such seed (or the trial
    // ends before its first evaluation, recorded through `ended`).
    __host__ __device__ bool trialBeginSeed(const StabRecord<NC> &r, TrialMachine &m, bool &ended)
    {
        ended = false;
        double seed[NC];
        if (!stabilitySeed(m.k, r.P, r.T, r.z, r.refComp, seed))
            return false;
        if (err.bad())
            return true;
        normalize(seed);
        for (int c = 0; c < NC; ++c) m.logW[c] = ::log(maxd(seed[c], o.floor));
        bool live;
        if (trialSsiSteps() > 0)
        {
            m.stage = kTrialSsi;
            m.it = 0;
            live = true;
        }
        else
        {
            live = trialEnterNewton(m);
        }
        ended = !live;
        return true;
    }

    __host__ __device__ void trialPoint(const TrialMachine &m, double *x) const
    {
        for (int c = 0; c < NC; ++c) x[c] = m.logW[c];
        if (m.stage == kTrialNewtonFd)
            x[m.col] += o.fdStep * maxd(1.0, ::fabs(m.logW[m.col]));
        else if (m.stage == kTrialLineSearch)
            for (int c = 0; c < NC; ++c) x[c] = m.logW[c] + m.damping * m.delta[c];
    }

    // Applies the residual at trialPoint(m); returns true when the current
    // trial ended, with its outcome in `t`.
    __host__ __device__ bool trialStep(const StabRecord<NC> &r, TrialMachine &m,
                                       const StabResidual &s, Trial &t)
    {
        t = Trial();
        switch (m.stage)
        {
        case kTrialSsi:
        case kTrialSsiTail:
        {
            if (err.bad() || !s.valid)
                return true;
            if (s.norm < o.stabTol)
            {
                t = finalizeTrial(s, r.refComp);
                return true;
            }
            for (int c = 0; c < NC; ++c)
                m.logW[c] = r.logRef[c] - ::log(maxd(s.phi[c], 1.0e-300));
            ++m.it;
            if (m.stage == kTrialSsi)
            {
                if (m.it >= trialSsiSteps())
                    return !trialEnterNewton(m);
                return false;
            }
            const int remaining = o.maxStabilityIterations - trialSsiSteps();
            return m.it >= remaining;
        }
        case kTrialNewtonBase:
            if (err.bad() || !s.valid)
                return true;
            if (s.norm < o.stabTol)
            {
                t = finalizeTrial(s, r.refComp);
                return true;
            }
            m.cur = s;
            m.col = 0;
            m.stage = kTrialNewtonFd;
            return false;
        case kTrialNewtonFd:
        {
            if (err.bad() || !s.valid)
                return true;
            const double step = o.fdStep * maxd(1.0, ::fabs(m.logW[m.col]));
            for (int row = 0; row < NC; ++row)
                m.jac[row][m.col] = (s.residual[row] - m.cur.residual[row]) / step;
            if (++m.col < NC)
                return false;
            double rhs[NC];
            for (int c = 0; c < NC; ++c) rhs[c] = -m.cur.residual[c];
            if (!solveDense<NC>(m.jac, rhs, m.delta))
                return !trialEnterTail(m);
            m.ls = 0;
            m.damping = 1.0;
            m.stage = kTrialLineSearch;
            return false;
        }
        default:  // kTrialLineSearch
            if (err.bad())
                return true;
            if (s.valid && s.norm < m.cur.norm)
            {
                for (int c = 0; c < NC; ++c) m.logW[c] = m.logW[c] + m.damping * m.delta[c];
                if (++m.it >= o.maxStabilityNewton)
                    return !trialEnterTail(m);
                m.stage = kTrialNewtonBase;
                return false;
            }
            m.damping *= 0.5;
            if (++m.ls >= 16)
                return !trialEnterTail(m);
            return false;
        }
    }

    // First missing candidate at or after `from` (3 when none).
    __host__ __device__ static int nextMissing(uint8_t active, int from)
    {
        int c = from;
        while (c < 3 && (active & (1u << c)))
            ++c;
        return c;
    }

    // Moves the machine to its next seed (or candidate) after a trial ended;
    // returns true when the record is complete (or an error ended it).
    __host__ __device__ bool trialNext(StabRecord<NC> &r, TrialMachine &m)
    {
        for (;;)
        {
            if (err.bad())
            {
                r.errCode = err.code;
                return true;
            }
            ++m.k;
            bool ended = false;
            if (trialBeginSeed(r, m, ended))
            {
                if (err.bad())
                    continue;
                if (!ended)
                    return false;
                // a trial that ends before any evaluation is invalid: keep best
                continue;
            }
            // candidate finished
            r.valid[m.cand] = m.best.valid ? 1 : 0;
            r.trivial[m.cand] = m.best.trivial ? 1 : 0;
            r.sum[m.cand] = m.best.sum;
            for (int c = 0; c < NC; ++c) r.comp[m.cand][c] = m.best.comp[c];
            m.cand = nextMissing(uint8_t(r.active), m.cand + 1);
            if (m.cand >= 3)
                return true;
            m.best = Trial();
            m.best.sum = -__builtin_inf();
            m.k = -1;
        }
    }

    // Initialises the machine for a recorded cell; true when it is already complete.
    __host__ __device__ bool trialInit(StabRecord<NC> &r, TrialMachine &m)
    {
        r.errCode = kErrNone;
        for (int p = 0; p < 3; ++p)
        {
            r.valid[p] = 0;
            r.trivial[p] = 0;
            r.sum[p] = -__builtin_inf();
        }
        m.cand = nextMissing(uint8_t(r.active), 0);
        if (m.cand >= 3)
            return true;
        m.best = Trial();
        m.best.sum = -__builtin_inf();
        m.k = -1;
        return trialNext(r, m);
    }

    // After trialStep reported the end of a trial: fold it into the candidate's best.
    __host__ __device__ bool trialFold(StabRecord<NC> &r, TrialMachine &m, const Trial &t)
    {
        if (err.bad())
        {
            r.errCode = err.code;
            return true;
        }
        if (t.valid && (!m.best.valid || t.sum > m.best.sum))
            m.best = t;
        return trialNext(r, m);
    }

    __host__ __device__ Result flashRestrictedImpl(double P, double T, const double *zin,
                                                   uint8_t allowed, const double (*seed)[NC]);

    // three_phase_flash.hpp:300-489
    __host__ __device__ Stab stabilityTest(double P, double T, const double *zin, uint8_t active,
                                           const double (*activeComp)[NC])
    {
        Stab result;
        if (!validatePT(P, T))
            return result;
        if (active == 0)
        {
            NN_TRACE("stability with no active phase\n");
            err.raise(kErrInvalid);
            return result;
        }
        double z[NC];
        for (int c = 0; c < NC; ++c) z[c] = zin[c];
        normalize(z);
        result.tested = active;
        const int ref = firstActive(active);
        double comps[3][NC];
        for (int p = 0; p < 3; ++p)
            for (int c = 0; c < NC; ++c) comps[p][c] = activeComp[p][c];
        if (sw() && popcount3(active) == 1)
            for (int c = 0; c < NC; ++c) comps[ref][c] = z[c];
        for (int p = 0; p < 3; ++p)
            if (active & (1u << p))
                normalize(comps[p]);
        if (err.bad())
            return result;

        const int stabUse = stabRec ? stabMode : kStabNone;
        stabMode = kStabNone;
        bool certifiedRun = false;
        if (sw() && popcount3(active) > 1)
        {
            PhaseThermo<NC> rawRef;
            thermo(ref, P, T, comps[ref], rawRef);
            double mismatch = 0.0;
            for (int p = 0; p < 3; ++p)
            {
                if (!(active & (1u << p)) || p == ref)
                    continue;
                PhaseThermo<NC> t;
                thermo(p, P, T, comps[p], t);
                for (int c = 0; c < NC; ++c)
                {
                    const double xr = comps[ref][c];
                    const double xv = comps[p][c];
                    if (xr <= 10.0 * o.floor || xv <= 10.0 * o.floor)
                        continue;
                    const double ra = maxd(xr * rawRef.phi[c], 1.0e-300);
                    const double a = maxd(xv * t.phi[c], 1.0e-300);
                    mismatch = maxd(mismatch, ::fabs(::log(ra / a)));
                }
            }
            if (err.bad())
                return result;
            const double certTol = maxd(10.0 * o.fugTol, 10.0 * o.stabTol);
            if (mismatch > certTol)
            {
                certifiedRun = true;
                Result certified = flashRestrictedImpl(P, T, z, active, comps);
                if (err.bad())
                    return result;
                if (!certified.converged || certified.presence != active)
                {
                    result.valid = false;
                    result.stable = false;
                    return result;
                }
                for (int p = 0; p < 3; ++p)
                    for (int c = 0; c < NC; ++c) comps[p][c] = certified.x[p][c];
            }
        }

        double refComp[NC];
        for (int c = 0; c < NC; ++c) refComp[c] = comps[ref][c];
        normalize(refComp);
        PhaseThermo<NC> refThermo;
        thermo(ref, P, T, refComp, refThermo);
        if (err.bad())
            return result;
        double logRef[NC];
        for (int c = 0; c < NC; ++c)
        {
            const double xi = maxd(refComp[c], o.floor);
            const double phi = maxd(refThermo.phi[c], 1.0e-300);
            logRef[c] = ::log(xi) + ::log(phi);
            NN_TRACE("stab ref=%d c=%d x=%.17g phi=%.17g logRef=%.17g\n", ref, c, refComp[c], refThermo.phi[c], logRef[c]);
        }

        if (stabUse == kStabRecord && !certifiedRun)
        {
            // Hand the seed trials to the lockstep pass (a state whose SW
            // certification ran a restricted flash finishes here instead).
            StabRecord<NC> &rec = *stabRec;
            rec.P = P;
            rec.T = T;
            rec.active = active;
            rec.errCode = kErrNone;
            for (int c = 0; c < NC; ++c)
            {
                rec.z[c] = z[c];
                rec.refComp[c] = refComp[c];
                rec.logRef[c] = logRef[c];
            }
            err.raise(kErrPending);
            return result;
        }
        const bool replay = stabUse == kStabReplay;
        if (replay && stabRec->errCode != kErrNone)
        {
            err.raise(stabRec->errCode);
            return result;
        }

        for (int cand = 0; cand < 3; ++cand)
        {
            if (active & (1u << cand))
                continue;
            Trial best;
            best.sum = -__builtin_inf();
            if (replay)
            {
                best.valid = stabRec->valid[cand] != 0;
                best.trivial = stabRec->trivial[cand] != 0;
                best.sum = stabRec->sum[cand];
                for (int c = 0; c < NC; ++c) best.comp[c] = stabRec->comp[cand][c];
            }
            double seed[NC];
            for (int k = 0; !replay && stabilitySeed(k, P, T, z, refComp, seed); ++k)
            {
                if (err.bad())
                    return result;
                const Trial t = stabilityTrial(cand, P, T, logRef, refComp, seed);
                NN_TRACE("trial cand=%d k=%d valid=%d trivial=%d sum=%.17g\n", cand, k, (int)t.valid, (int)t.trivial, t.sum);
                if (err.bad())
                    return result;
                if (t.valid && (!best.valid || t.sum > best.sum))
                    best = t;
            }
            if (!best.valid)
            {
                result.valid = false;
                result.stable = false;
                continue;
            }
            result.trialSum[cand] = best.sum;
            for (int c = 0; c < NC; ++c) result.incipient[cand][c] = best.comp[c];
            bool coincides = false;
            const double ctol = maxd(10.0 * o.coincidentTol, 1.0e-7);
            for (int p = 0; p < 3; ++p)
            {
                if (!(active & (1u << p)))
                    continue;
                double md = 0.0;
                for (int c = 0; c < NC; ++c) md = maxd(md, ::fabs(best.comp[c] - comps[p][c]));
                if (md <= ctol)
                {
                    coincides = true;
                    break;
                }
            }
            bool unstable = !best.trivial && !coincides && best.sum > 1.0 + o.stabTol;
            if (cand == kWater && unstable && !aqueousSupported(best.comp))
                unstable = false;
            result.unstable[cand] = unstable;
            result.stable = result.stable && !unstable;
        }
        return result;
    }

    // ------------------------------------------------- active-set equilibrium
    // allocation.inc:707-858
    __host__ __device__ Result equilibrateActiveSet(double P, double T, const double *z,
                                                    uint8_t active, Result seed, bool reuse = true)
    {
        if (active == 0)
            active = chooseSinglePhase(P, T, z);
        canonicalize(P, T, seed);
        if (err.bad())
            return seed;
        if (seed.presence != active)
            active = seed.presence;
        Result lastPhysical = seed;
        bool useSeed = reuse && seed.converged && seed.presence == active && !sw();

        for (int outer = 0; outer < 8; ++outer)
        {
            if (err.bad())
                return seed;
            Result current;
            const int count = popcount3(active);
            if (useSeed)
            {
                current = seed;
                useSeed = false;
            }
            else if (count == 1)
            {
                current = singlePhaseResult(P, T, z, active);
            }
            else if (count == 2)
            {
                int a, b;
                activePair(active, a, b);
                double K[NC];
                pairInitialK(a, b, P, T, z, K);
                if (seed.converged && seed.presence == active)
                    for (int c = 0; c < NC; ++c)
                        K[c] = clampd(seed.x[b][c] / maxd(seed.x[a][c], o.floor), 1.0e-100, 1.0e100);
                current = flashTwo(P, T, z, a, b, K);
                if (err.bad())
                    return current;
                if (!current.converged)
                    current = singlePhaseResult(P, T, z, chooseSinglePhase(P, T, z));
            }
            else
            {
                double kg[NC], kw[NC];
                bool kgValid = true, kwValid = true;
                for (int c = 0; c < NC; ++c)
                {
                    kg[c] = seed.kg[c];
                    kw[c] = seed.kw[c];
                    kgValid = kgValid && kg[c] > 0.0;
                    kwValid = kwValid && kw[c] > 0.0;
                }
                if (!kgValid) wilsonK(P, T, kg);
                if (!kwValid) initialWaterOilK(kw);
                current = flashThree(P, T, z, kg, kw);
                if (err.bad())
                    return current;
                if (!current.converged)
                {
                    lastPhysical.converged = false;
                    canonicalize(P, T, lastPhysical);
                    return lastPhysical;
                }
            }
            if (err.bad())
                return current;

            uint8_t fractionPresence = presenceFromFractions(current.beta);
            NN_TRACE("eq outer=%d active=%d conv=%d it=%d beta=%.17g %.17g %.17g\n", outer, (int)active, (int)current.converged, current.iterations, current.beta[0], current.beta[1], current.beta[2]);
            if (fractionPresence == 0)
                fractionPresence = chooseSinglePhase(P, T, z);
            if (fractionPresence != active)
            {
                active = fractionPresence;
                seed = current;
                canonicalize(P, T, seed);
                lastPhysical = seed;
                useSeed = seed.converged && seed.presence == active && !sw();
                continue;
            }
            canonicalize(P, T, current);
            lastPhysical = current;
            if (popcount3(active) == 3)
            {
                canonicalize(P, T, current);
                return current;
            }
            const Stab st = stabilityTest(P, T, z, active, current.x);
            NN_TRACE("stab active=%d valid=%d stable=%d unst=%d%d%d sums=%.17g %.17g %.17g\n", (int)active, (int)st.valid, (int)st.stable, (int)st.unstable[0], (int)st.unstable[1], (int)st.unstable[2], st.trialSum[0], st.trialSum[1], st.trialSum[2]);
            if (err.bad())
                return current;
            if (!st.valid)
            {
                current.converged = false;
                canonicalize(P, T, current);
                return current;
            }
            if (st.stable)
            {
                canonicalize(P, T, current);
                return current;
            }
            bool added = false;
            for (int p = 0; p < 3; ++p)
            {
                if (!(active & (1u << p)) && st.unstable[p])
                {
                    active |= uint8_t(1u << p);
                    for (int c = 0; c < NC; ++c) seed.x[p][c] = st.incipient[p][c];
                    added = true;
                }
            }
            if (!added)
            {
                current.converged = false;
                canonicalize(P, T, current);
                return current;
            }
            deriveReferenceK(seed);
            seed.presence = active;
            useSeed = false;
        }
        canonicalize(P, T, seed);
        return seed;
    }

    // --------------------------------------------------- recovery helpers
    // restricted.inc:336-412
    __host__ __device__ void finalizePhysicallyConvergedTwoPhase(double P, double T, const double *z,
                                                                 int first, int second, Result &r)
    {
        if (r.converged)
            return;
        const double ba = r.beta[first], bb = r.beta[second];
        if (!(ba > o.betaTol) || !(bb > o.betaTol) || !finite(ba) || !finite(bb))
            return;
        const double closure = ::fabs(ba + bb - 1.0);
        const double tol = maxd(o.fugTol, 10.0 * o.rrTol);
        if (closure > tol)
            return;
        double xa[NC], xb[NC];
        for (int c = 0; c < NC; ++c) { xa[c] = r.x[first][c]; xb[c] = r.x[second][c]; }
        normalize(xa);
        normalize(xb);
        double mat = 0.0;
        for (int c = 0; c < NC; ++c)
            mat = maxd(mat, ::fabs(ba * xa[c] + bb * xb[c] - z[c]));
        if (mat > tol)
            return;
        PhaseThermo<NC> ta, tb;
        thermo(first, P, T, xa, ta);
        thermo(second, P, T, xb, tb);
        if (err.bad())
            return;
        double fres = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            const double aa = maxd(xa[c] * ta.phi[c], 1.0e-300);
            const double ab = maxd(xb[c] * tb.phi[c], 1.0e-300);
            fres = maxd(fres, ::fabs(::log(aa / ab)));
        }
        if (fres > o.fugTol)
            return;
        for (int c = 0; c < NC; ++c) { r.x[first][c] = xa[c]; r.x[second][c] = xb[c]; }
        r.presence = uint8_t((1u << first) | (1u << second));
        r.Z[first] = ta.Z;
        r.Z[second] = tb.Z;
        r.converged = true;
        deriveReferenceK(r);
        updateSaturationsFromMoles(P, T, r);
    }

    // three_phase_flash.hpp:235-292 (seed may be null)
    __host__ __device__ Result flashRestricted(double P, double T, const double *zin, uint8_t allowed,
                                               const double (*seed)[NC])
    {
        double z[NC];
        for (int c = 0; c < NC; ++c) z[c] = zin[c];
        normalize(z);
        if (err.bad())
            return Result();
        Result r = flashRestrictedImpl(P, T, z, allowed, seed);
        if (err.bad())
            return r;
        if (!r.converged && popcount3(allowed) == 2 && sw())
        {
            int a, b;
            activePair(allowed, a, b);
            finalizePhysicallyConvergedTwoPhase(P, T, z, a, b, r);
        }
        if (r.converged)
        {
            canonicalizeFinal(P, T, r, true);
            if (r.converged && !validatePhysical(P, T, z, r))
                r.converged = false;
        }
        return r;
    }

    // restricted.inc:631-648
    __host__ __device__ Result flashRestrictedRecovery(double P, double T, const double *zin, uint8_t allowed)
    {
        double z[NC];
        for (int c = 0; c < NC; ++c) z[c] = zin[c];
        normalize(z);
        Result r = flashRestrictedImpl(P, T, z, allowed, nullptr);
        if (err.bad())
            return r;
        if (r.converged)
        {
            canonicalizeFinal(P, T, r, false);
            if (r.converged && !validatePhysical(P, T, z, r))
                r.converged = false;
        }
        return r;
    }

    // restricted.inc:657-696
    __host__ __device__ Result stableReducedSetFallback(double P, double T, const double *z)
    {
        Result best;
        double bestG = __builtin_inf();
        const uint8_t pairs[3] = {kOilBit | kGasBit, kOilBit | kWaterBit, kGasBit | kWaterBit};
        for (int k = 0; k < 3; ++k)
        {
            Result cand = flashRestrictedRecovery(P, T, z, pairs[k]);
            if (err.bad())
                return best;
            if (!cand.converged)
                continue;
            cand.presence = presenceFromFractions(cand.beta);
            if (cand.presence == 0)
                continue;
            const Stab st = stabilityTest(P, T, z, cand.presence, cand.x);
            if (err.bad())
                return best;
            if (!st.valid || !st.stable)
                continue;
            const double g = dimensionlessGibbs(P, T, cand);
            noteGibbsTie(g, bestG);
            if (finite(g) && g < bestG)
            {
                bestG = g;
                best = cand;
            }
        }
        return best;
    }

    // restricted.inc:706-779
    __host__ __device__ Result preferStableSwReducedOverSingle(double P, double T, const double *z,
                                                               Result cand)
    {
        if (!sw() || popcount3(presenceFromFractions(cand.beta)) != 1)
            return cand;
        cand.presence = presenceFromFractions(cand.beta);
        const Stab single = stabilityTest(P, T, z, cand.presence, cand.x);
        if (err.bad())
            return cand;
        const bool singleKnownUnstable = single.valid && !single.stable;
        Result reduced;
        double reducedG = __builtin_inf();
        const uint8_t pairs[3] = {kOilBit | kGasBit, kOilBit | kWaterBit, kGasBit | kWaterBit};
        for (int k = 0; k < 3; ++k)
        {
            Result trial = flashRestricted(P, T, z, pairs[k], nullptr);
            if (err.bad())
                return cand;
            if (!trial.converged)
                continue;
            trial.presence = presenceFromFractions(trial.beta);
            if (trial.presence == 0)
                continue;
            const Stab st = stabilityTest(P, T, z, trial.presence, trial.x);
            if (err.bad())
                return cand;
            if (!st.valid || !st.stable)
                continue;
            const double g = dimensionlessGibbs(P, T, trial);
            if (finite(g) && g < reducedG)
            {
                reducedG = g;
                reduced = trial;
            }
        }
        if (!reduced.converged)
        {
            if (singleKnownUnstable)
                cand.converged = false;
            return cand;
        }
        const double candG = dimensionlessGibbs(P, T, cand);
        const double scale = maxd(maxd(1.0, ::fabs(candG)), ::fabs(reducedG));
        if (finite(reducedG) && (!finite(candG) || reducedG < candG - 1.0e-12 * scale))
            return reduced;
        if (singleKnownUnstable)
            cand.converged = false;
        return cand;
    }

    // restricted.inc:514-543
    __host__ __device__ Result numericalSwRoleResult(const Result &canonical) const
    {
        if (!sw())
            return canonical;
        Result n = canonical;
        for (int p = 0; p < 3; ++p)
        {
            n.beta[p] = 0.0; n.sat[p] = 0.0; n.Z[p] = 1.0; n.rho[p] = 0.0;
            for (int c = 0; c < NC; ++c) n.x[p][c] = 0.0;
        }
        for (int s = 0; s < 3; ++s)
        {
            const int phys = roles[s];
            n.beta[s] = canonical.beta[phys];
            n.sat[s] = canonical.sat[phys];
            for (int c = 0; c < NC; ++c) n.x[s][c] = canonical.x[phys][c];
            n.Z[s] = canonical.Z[phys];
            n.rho[s] = canonical.rho[phys];
        }
        n.presence = presenceFromFractions(n.beta);
        deriveReferenceK(n);
        return n;
    }

    __host__ __device__ void noteGibbsTie(double g, double bestG)
    {
        if (finite(g) && finite(bestG) && ::fabs(g - bestG) <= 1.0e-12 * maxd(1.0, ::fabs(bestG)))
            nearTie = true;
    }

    // ------------------------------------------- SW exact-mass allocation
    // allocation.inc:8-700.  Unknowns u = (u_g[NC], u_w[NC]) are per-component
    // softmax logits of the split of z_i over (oil, gas, water), so the
    // material balance holds exactly for every u; the search minimises Gibbs
    // energy (BFGS) and then drives the chemical-potential differences to the
    // fugacity tolerance (Levenberg-Marquardt on a central-difference Jacobian).
    static constexpr int kAllocDim = 2 * NC;

    struct Allocation
    {
        bool valid;
        double gibbs;
        double closure;
        double norm2;
        double residual[kAllocDim];
        double gradient[kAllocDim];
        double x[3][NC];
        double beta[3];
        double Z[3];
    };

    // allocation.inc:32-128
    __host__ __device__ void evaluateAllocation(double P, double T, const double *z,
                                                const double *u, Allocation &s)
    {
        s.valid = false;
        s.gibbs = 0.0;
        s.closure = __builtin_inf();
        s.norm2 = __builtin_inf();
        for (int k = 0; k < kAllocDim; ++k)
        {
            s.residual[k] = 0.0;
            s.gradient[k] = 0.0;
        }
        double amount[3][NC];
        for (int p = 0; p < 3; ++p)
        {
            s.beta[p] = 0.0;
            s.Z[p] = 0.0;
            for (int c = 0; c < NC; ++c) s.x[p][c] = 0.0;
        }
        for (int c = 0; c < NC; ++c)
        {
            const double ug = u[c];
            const double uw = u[NC + c];
            double shift = 0.0;
            if (shift < ug) shift = ug;
            if (shift < uw) shift = uw;
            const double qo0 = ::exp(-shift);
            const double qg0 = ::exp(ug - shift);
            const double qw0 = ::exp(uw - shift);
            const double qsum = qo0 + qg0 + qw0;
            amount[0][c] = z[c] * (qo0 / qsum);
            amount[1][c] = z[c] * (qg0 / qsum);
            amount[2][c] = z[c] * (qw0 / qsum);
            s.beta[0] += amount[0][c];
            s.beta[1] += amount[1][c];
            s.beta[2] += amount[2][c];
        }
        for (int p = 0; p < 3; ++p)
        {
            if (!(s.beta[p] > 1.0e-14))
                return;
            for (int c = 0; c < NC; ++c) s.x[p][c] = amount[p][c] / s.beta[p];
        }
        double mu[3][NC];
        for (int p = 0; p < 3; ++p)
        {
            PhaseThermo<NC> t;
            thermo(p, P, T, s.x[p], t);
            if (err.bad())
                return;
            s.Z[p] = t.Z;
            for (int c = 0; c < NC; ++c)
            {
                const double xc = maxd(s.x[p][c], o.floor);
                const double phi = maxd(t.phi[c], 1.0e-300);
                mu[p][c] = ::log(xc) + ::log(phi);
                s.gibbs += amount[p][c] * mu[p][c];
            }
        }
        s.closure = 0.0;
        s.norm2 = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            const int w = NC + c;
            if (!(z[c] > o.floor))
                continue;  // absent component: zero residual and gradient rows
            s.residual[c] = mu[1][c] - mu[0][c];
            s.residual[w] = mu[2][c] - mu[0][c];
            s.norm2 += s.residual[c] * s.residual[c] + s.residual[w] * s.residual[w];
            s.closure = maxd(s.closure, ::fabs(mu[0][c] - mu[1][c]));
            s.closure = maxd(s.closure, ::fabs(mu[0][c] - mu[2][c]));
            s.closure = maxd(s.closure, ::fabs(mu[1][c] - mu[2][c]));
            const double qo = amount[0][c] / z[c];
            const double qg = amount[1][c] / z[c];
            const double qw = amount[2][c] / z[c];
            const double average = qo * mu[0][c] + qg * mu[1][c] + qw * mu[2][c];
            s.gradient[c] = z[c] * qg * (mu[1][c] - average);
            s.gradient[w] = z[c] * qw * (mu[2][c] - average);
        }
        s.valid = finite(s.gibbs) && finite(s.norm2) && finite(s.closure);
    }

    // allocation.inc:135-238: BFGS on the Gibbs energy (inverse-Hessian form),
    // Armijo backtracking, steps capped at 4 in the max norm.
    __host__ __device__ void gibbsPreconditionAllocation(double P, double T, const double *z,
                                                         double *u, Allocation &state)
    {
        double H[kAllocDim][kAllocDim];
        for (int i = 0; i < kAllocDim; ++i)
            for (int j = 0; j < kAllocDim; ++j) H[i][j] = (i == j) ? 1.0 : 0.0;
        evaluateAllocation(P, T, z, u, state);
        if (err.bad() || !state.valid)
            return;
        Allocation trial;
        for (int iteration = 0; iteration < 160; ++iteration)
        {
            double dir[kAllocDim];
            for (int i = 0; i < kAllocDim; ++i)
            {
                dir[i] = 0.0;
                for (int j = 0; j < kAllocDim; ++j) dir[i] -= H[i][j] * state.gradient[j];
            }
            double dd = 0.0, maxStep = 0.0;
            for (int i = 0; i < kAllocDim; ++i)
            {
                dd += state.gradient[i] * dir[i];
                maxStep = maxd(maxStep, ::fabs(dir[i]));
            }
            if (!(dd < 0.0) || !finite(dd))
            {
                dd = 0.0;
                maxStep = 0.0;
                for (int i = 0; i < kAllocDim; ++i)
                {
                    for (int j = 0; j < kAllocDim; ++j) H[i][j] = (i == j) ? 1.0 : 0.0;
                    dir[i] = -state.gradient[i];
                    dd += state.gradient[i] * dir[i];
                    maxStep = maxd(maxStep, ::fabs(dir[i]));
                }
            }
            if (maxStep > 4.0)
                for (int i = 0; i < kAllocDim; ++i) dir[i] *= 4.0 / maxStep;

            double trialU[kAllocDim];
            double step = 1.0;
            bool accepted = false;
            for (int ls = 0; ls < 24; ++ls)
            {
                for (int i = 0; i < kAllocDim; ++i)
                    trialU[i] = clampd(u[i] + step * dir[i], -40.0, 40.0);
                evaluateAllocation(P, T, z, trialU, trial);
                if (err.bad())
                    return;
                if (trial.valid && trial.gibbs <= state.gibbs + 1.0e-4 * step * dd)
                {
                    accepted = true;
                    break;
                }
                step *= 0.5;
            }
            if (!accepted)
                return;

            double sv[kAllocDim], yv[kAllocDim], hy[kAllocDim];
            double ys = 0.0;
            for (int i = 0; i < kAllocDim; ++i)
            {
                sv[i] = trialU[i] - u[i];
                yv[i] = trial.gradient[i] - state.gradient[i];
            }
            for (int i = 0; i < kAllocDim; ++i) ys += yv[i] * sv[i];
            if (ys > 1.0e-12)
            {
                double yhy = 0.0;
                for (int i = 0; i < kAllocDim; ++i)
                {
                    hy[i] = 0.0;
                    for (int j = 0; j < kAllocDim; ++j) hy[i] += H[i][j] * yv[j];
                }
                for (int i = 0; i < kAllocDim; ++i) yhy += yv[i] * hy[i];
                const double coefficient = (ys + yhy) / (ys * ys);
                for (int i = 0; i < kAllocDim; ++i)
                    for (int j = 0; j < kAllocDim; ++j)
                        H[i][j] += coefficient * sv[i] * sv[j] - (hy[i] * sv[j] + sv[i] * hy[j]) / ys;
            }
            else
            {
                for (int i = 0; i < kAllocDim; ++i)
                    for (int j = 0; j < kAllocDim; ++j) H[i][j] = (i == j) ? 1.0 : 0.0;
            }
            for (int i = 0; i < kAllocDim; ++i) u[i] = trialU[i];
            state = trial;
        }
    }

    // allocation.inc:245-343: Levenberg-Marquardt on (mu_g - mu_o, mu_w - mu_o).
    __host__ __device__ bool refineAllocationChemicalPotentials(double P, double T, const double *z,
                                                                double *u, Allocation &state)
    {
        evaluateAllocation(P, T, z, u, state);
        if (err.bad() || !state.valid)
            return false;
        double lambda = 1.0e-5;
        Allocation ps, ms;
        for (int iteration = 0; iteration < 160; ++iteration)
        {
            if (state.closure < o.fugTol)
                return true;
            double J[kAllocDim][kAllocDim];
            for (int col = 0; col < kAllocDim; ++col)
            {
                const double h = 2.0e-5 * maxd(1.0, ::fabs(u[col]));
                double plus[kAllocDim], minus[kAllocDim];
                for (int i = 0; i < kAllocDim; ++i) plus[i] = minus[i] = u[i];
                plus[col] = mind(40.0, u[col] + h);
                minus[col] = maxd(-40.0, u[col] - h);
                evaluateAllocation(P, T, z, plus, ps);
                if (err.bad())
                    return false;
                evaluateAllocation(P, T, z, minus, ms);
                if (err.bad())
                    return false;
                if (!ps.valid || !ms.valid)
                    return false;
                const double den = plus[col] - minus[col];
                for (int row = 0; row < kAllocDim; ++row)
                    J[row][col] = (ps.residual[row] - ms.residual[row]) / den;
            }
            bool accepted = false;
            for (int lm = 0; lm < 14 && !accepted; ++lm)
            {
                double normal[kAllocDim][kAllocDim], rhs[kAllocDim], delta[kAllocDim];
                for (int a = 0; a < kAllocDim; ++a)
                {
                    for (int b = 0; b < kAllocDim; ++b)
                    {
                        double value = 0.0;
                        for (int row = 0; row < kAllocDim; ++row) value += J[row][a] * J[row][b];
                        normal[a][b] = value;
                    }
                    double value = 0.0;
                    for (int row = 0; row < kAllocDim; ++row) value += J[row][a] * state.residual[row];
                    rhs[a] = -value;
                    normal[a][a] += lambda;
                }
                if (!solveDense<kAllocDim>(normal, rhs, delta))
                {
                    lambda *= 10.0;
                    continue;
                }
                double maxDelta = 0.0;
                for (int i = 0; i < kAllocDim; ++i) maxDelta = maxd(maxDelta, ::fabs(delta[i]));
                if (maxDelta > 6.0)
                    for (int i = 0; i < kAllocDim; ++i) delta[i] *= 6.0 / maxDelta;
                double damping = 1.0;
                for (int ls = 0; ls < 22; ++ls)
                {
                    double trialU[kAllocDim];
                    for (int i = 0; i < kAllocDim; ++i)
                        trialU[i] = clampd(u[i] + damping * delta[i], -40.0, 40.0);
                    evaluateAllocation(P, T, z, trialU, ps);
                    if (err.bad())
                        return false;
                    if (ps.valid && ps.norm2 < state.norm2 * (1.0 - 1.0e-5 * damping))
                    {
                        for (int i = 0; i < kAllocDim; ++i) u[i] = trialU[i];
                        state = ps;
                        lambda = maxd(1.0e-14, lambda * 0.2);
                        accepted = true;
                        break;
                    }
                    damping *= 0.5;
                }
                if (!accepted)
                    lambda *= 10.0;
            }
            if (!accepted)
                return false;
        }
        return state.closure < o.fugTol;
    }

    // u from per-phase component amounts (the common tail of the seed builders).
    __host__ __device__ static void logitsFromAmounts(const double *z, const double (*amount)[NC],
                                                      int c, double *u)
    {
        const double qo = maxd(amount[0][c] / z[c], 1.0e-18);
        const double qg = maxd(amount[1][c] / z[c], 1.0e-18);
        const double qw = maxd(amount[2][c] / z[c], 1.0e-18);
        u[c] = ::log(qg / qo);
        u[NC + c] = ::log(qw / qo);
    }

    // allocation.inc:359-443
    __host__ __device__ bool allocationSeedFromPair(double P, double T, const double *z,
                                                    uint8_t allowed, double *u)
    {
        const Result pair = flashRestrictedRecovery(P, T, z, allowed);
        if (err.bad() || !pair.converged || popcount3(pair.presence) != 2)
            return false;
        const Stab st = stabilityTest(P, T, z, pair.presence, pair.x);
        if (err.bad())
            return false;
        int missing = -1;
        for (int p = 0; p < 3; ++p)
            if (!(pair.presence & (1u << p)))
            {
                missing = p;
                break;
            }
        if (missing < 0 || !st.valid || !st.unstable[missing])
            return false;
        const double *inc = st.incipient[missing];
        double eps = 1.0e-3;
        for (int c = 0; c < NC; ++c)
            if (inc[c] > 1.0e-30)
                eps = mind(eps, 0.15 * z[c] / inc[c]);
        eps = maxd(eps, 1.0e-9);
        double amount[3][NC];
        for (int p = 0; p < 3; ++p)
            for (int c = 0; c < NC; ++c) amount[p][c] = 0.0;
        for (int c = 0; c < NC; ++c)
        {
            if (!(z[c] > o.floor))
                continue;
            amount[missing][c] = eps * inc[c];
            const double remaining = z[c] - amount[missing][c];
            if (!(remaining > 0.0))
                return false;
            double pairAmount = 0.0;
            for (int p = 0; p < 3; ++p)
                if (p != missing)
                    pairAmount += pair.beta[p] * pair.x[p][c];
            if (!(pairAmount > 0.0))
                return false;
            for (int p = 0; p < 3; ++p)
                if (p != missing)
                    amount[p][c] = pair.beta[p] * pair.x[p][c] * remaining / pairAmount;
        }
        for (int c = 0; c < NC; ++c)
        {
            if (!(z[c] > o.floor))
            {
                u[c] = 0.0;
                u[NC + c] = 0.0;
                continue;
            }
            logitsFromAmounts(z, amount, c, u);
        }
        return true;
    }

    // allocation.inc:446-501
    __host__ __device__ bool allocationSeedFromRawOg(double P, double T, const double *z, double *u)
    {
        double K[NC];
        wilsonK(P, T, K);
        Result pair = flashTwo(P, T, z, kOil, kGas, K);
        if (err.bad() || !pair.converged || popcount3(presenceFromFractions(pair.beta)) != 2)
            return false;
        pair.presence = presenceFromFractions(pair.beta);
        const Stab st = stabilityTest(P, T, z, pair.presence, pair.x);
        if (err.bad())
            return false;
        if (!st.valid || !st.unstable[kWater])
            return false;
        const double *inc = st.incipient[kWater];
        double eps = 1.0e-3;
        for (int c = 0; c < NC; ++c)
            if (inc[c] > 1.0e-30)
                eps = mind(eps, 0.2 * z[c] / inc[c]);
        eps = maxd(eps, 1.0e-9);
        double amount[3][NC];
        for (int c = 0; c < NC; ++c)
        {
            if (!(z[c] > o.floor))
            {
                u[c] = 0.0;
                u[NC + c] = 0.0;
                continue;
            }
            amount[2][c] = eps * inc[c];
            const double remaining = z[c] - amount[2][c];
            const double no = pair.beta[0] * pair.x[0][c];
            const double ng = pair.beta[1] * pair.x[1][c];
            const double sum = no + ng;
            if (!(remaining > 0.0) || !(sum > 0.0))
                return false;
            amount[0][c] = no * remaining / sum;
            amount[1][c] = ng * remaining / sum;
            logitsFromAmounts(z, amount, c, u);
        }
        return true;
    }

    // allocation.inc:503-582
    __host__ __device__ bool allocationSeedFromSingle(double P, double T, const double *z, double *u)
    {
        const uint8_t active = chooseSinglePhase(P, T, z);
        if (err.bad())
            return false;
        const Result single = singlePhaseResult(P, T, z, active);
        if (err.bad())
            return false;
        const Stab st = stabilityTest(P, T, z, active, single.x);
        if (err.bad())
            return false;
        if (!st.valid || st.stable)
            return false;
        double eps[3] = {0.0, 0.0, 0.0};
        int missingCount = 0;
        for (int p = 0; p < 3; ++p)
            if (!(active & (1u << p)) && st.unstable[p])
            {
                eps[p] = 1.0e-4;
                ++missingCount;
            }
        if (missingCount == 0)
            return false;
        for (int c = 0; c < NC; ++c)
            for (int p = 0; p < 3; ++p)
            {
                if (!(eps[p] > 0.0))
                    continue;
                eps[p] = mind(eps[p], 0.08 * z[c] / maxd(st.incipient[p][c], 1.0e-30));
            }
        double amount[3][NC];
        for (int p = 0; p < 3; ++p)
            for (int c = 0; c < NC; ++c) amount[p][c] = 0.0;
        const int activeSlot = firstActive(active);
        for (int c = 0; c < NC; ++c)
        {
            if (!(z[c] > o.floor))
            {
                u[c] = 0.0;
                u[NC + c] = 0.0;
                continue;
            }
            double removed = 0.0;
            for (int p = 0; p < 3; ++p)
            {
                if (!(eps[p] > 0.0))
                    continue;
                amount[p][c] = eps[p] * st.incipient[p][c];
                removed += amount[p][c];
            }
            if (!(z[c] > removed))
                return false;
            amount[activeSlot][c] = z[c] - removed;
            logitsFromAmounts(z, amount, c, u);
        }
        return true;
    }

    // allocation.inc:584-602
    __host__ __device__ void thermodynamicAllocationSeed(double P, double T, double scale,
                                                         double gasShift, double waterShift,
                                                         double *u) const
    {
        double kg[NC], kw[NC];
        wilsonK(P, T, kg);
        initialWaterOilK(kw);
        for (int c = 0; c < NC; ++c)
        {
            u[c] = clampd(scale * ::log(clampd(kg[c], 1.0e-30, 1.0e30)) + gasShift, -30.0, 30.0);
            u[NC + c] = clampd(scale * ::log(clampd(kw[c], 1.0e-30, 1.0e30)) + waterShift, -30.0, 30.0);
        }
    }

    // allocation.inc:609-700: 23 seeds (raw O/G pair, single phase, the three
    // restricted pairs, 6 Wilson-scaled and 12 uniform logit biases), each
    // preconditioned and refined; the lowest-Gibbs three-phase solution wins.
    __host__ __device__ __attribute__((noinline)) Result flashThreeAllocationFallback(
        double P, double T, const double *z)
    {
        Result best;
        double bestG = __builtin_inf();
        if constexpr (Mode == kModeDefer && kDeferAllocation)
        {
            err.raise(kErrDefer);
            return best;
        }
        constexpr double thermoSeeds[6][3] = {{1.0, 0.0, 0.0},  {0.5, 0.0, 0.0},
                                              {1.5, 0.0, 0.0},  {1.0, -1.5, -1.5},
                                              {1.0, 1.5, -1.5}, {1.0, -1.5, 1.5}};
        constexpr double generic[12][2] = {{0.0, 0.0},  {1.0, -2.0},  {2.0, -4.0}, {2.0, -2.0},
                                           {1.0, -4.0}, {-1.0, -1.0}, {-2.0, -2.0}, {-2.0, 1.0},
                                           {-1.0, 2.0}, {2.0, 1.0},   {1.0, 2.0},   {-2.0, 2.0}};
        constexpr uint8_t pairs[3] = {kOilBit | kGasBit, kOilBit | kWaterBit, kGasBit | kWaterBit};
        // Seeds that need flashes are built first, in the reference's order,
        // and kept; the closed-form seeds are generated when their turn comes.
        double built[5][kAllocDim];
        bool builtValid[5];
        builtValid[0] = allocationSeedFromRawOg(P, T, z, built[0]);
        if (err.bad())
            return best;
        builtValid[1] = allocationSeedFromSingle(P, T, z, built[1]);
        if (err.bad())
            return best;
        for (int k = 0; k < 3; ++k)
        {
            builtValid[2 + k] = allocationSeedFromPair(P, T, z, pairs[k], built[2 + k]);
            if (err.bad())
                return best;
        }
        Allocation state;
        auto seedVector = [&](int seed, double *u) {
            if (seed < 5)
            {
                if (!builtValid[seed])
                    return false;
                for (int i = 0; i < kAllocDim; ++i) u[i] = built[seed][i];
            }
            else if (seed < 11)
            {
                const double *spec = thermoSeeds[seed - 5];
                thermodynamicAllocationSeed(P, T, spec[0], spec[1], spec[2], u);
            }
            else
            {
                for (int c = 0; c < NC; ++c)
                {
                    u[c] = generic[seed - 11][0];
                    u[NC + c] = generic[seed - 11][1];
                }
            }
            return true;
        };
        if constexpr (Mode == kModeCoop)
        {
            // Seed s is searched on lane s; the selection below then walks the
            // seeds in order with the serial loop's rules (first error aborts,
            // strict Gibbs improvement, invalid candidates skipped).
            const int lane = waveLane();
            bool ok = false;
            int seedErr = kErrNone;
            if (lane < 23)
            {
                double u[kAllocDim];
                if (seedVector(lane, u))
                {
                    gibbsPreconditionAllocation(P, T, z, u, state);
                    if (!err.bad() && state.valid)
                    {
                        const bool refined = refineAllocationChemicalPotentials(P, T, z, u, state);
                        ok = !err.bad() && refined && state.beta[0] > o.betaTol &&
                             state.beta[1] > o.betaTol && state.beta[2] > o.betaTol;
                    }
                }
                seedErr = err.code;
                err.code = kErrNone;
            }
            for (int seed = 0; seed < 23; ++seed)
            {
                const int e = waveShfl(seedErr, seed);
                if (e != kErrNone)
                {
                    err.raise(e);
                    return best;
                }
                if (!waveShfl(int(ok), seed))
                    continue;
                const double g = waveShfl(state.gibbs, seed);
                noteGibbsTie(g, bestG);
                if (!(g < bestG))
                    continue;
                Result cand;
                cand.converged = true;
                cand.presence = kAllBits;
                for (int p = 0; p < 3; ++p)
                {
                    cand.beta[p] = waveShfl(state.beta[p], seed);
                    cand.Z[p] = waveShfl(state.Z[p], seed);
                    for (int c = 0; c < NC; ++c) cand.x[p][c] = waveShfl(state.x[p][c], seed);
                }
                deriveReferenceK(cand);
                updateSaturationsFromMoles(P, T, cand);
                if (err.bad())
                    return best;
                if (!cand.converged)
                    continue;
                best = cand;
                bestG = g;
            }
            return best;
        }
        for (int seed = 0; seed < 23; ++seed)
        {
            double u[kAllocDim];
            if (!seedVector(seed, u))
                continue;
            gibbsPreconditionAllocation(P, T, z, u, state);
            if (err.bad())
                return best;
            if (!state.valid)
                continue;
            const bool refined = refineAllocationChemicalPotentials(P, T, z, u, state);
            if (err.bad())
                return best;
            if (!refined)
                continue;
            if (!(state.beta[0] > o.betaTol) || !(state.beta[1] > o.betaTol) ||
                !(state.beta[2] > o.betaTol))
                continue;
            noteGibbsTie(state.gibbs, bestG);
            if (!(state.gibbs < bestG))
                continue;
            Result cand;
            cand.converged = true;
            cand.presence = kAllBits;
            for (int p = 0; p < 3; ++p)
            {
                cand.beta[p] = state.beta[p];
                cand.Z[p] = state.Z[p];
                for (int c = 0; c < NC; ++c) cand.x[p][c] = state.x[p][c];
            }
            deriveReferenceK(cand);
            updateSaturationsFromMoles(P, T, cand);
            if (err.bad())
                return best;
            if (!cand.converged)
                continue;
            best = cand;
            bestG = state.gibbs;
        }
        return best;
    }

    // ------------------------------------------------------------- flash
    // three_phase_flash.hpp:83-175: everything up to (and including) the SW
    // allocation stage.  `done` reports whether a result was returned early.
    __host__ __device__ Result flashMain(double P, double T, double *z, bool &done)
    {
        done = true;
        Result failure;
        if (!validatePT(P, T))
            return failure;
        normalize(z);
        if (err.bad())
            return failure;
        double kg[NC], kw[NC];
        wilsonK(P, T, kg);
        initialWaterOilK(kw);

        Result seed;
        if (hint != NN_HINT_SINGLE_PHASE)
        {
            double k0[NC];
            for (int c = 0; c < NC; ++c)
                k0[c] = (hint == NN_HINT_LOG_K && hintLogK) ? ::exp(clampd(hintLogK[c], -230.0, 230.0))
                                                           : kg[c];
            seed = flashTwo(P, T, z, kOil, kGas, k0);
        }
        NN_TRACE("seed conv=%d it=%d beta=%.17g %.17g %.17g\n", (int)seed.converged, seed.iterations, seed.beta[0], seed.beta[1], seed.beta[2]);
        if (err.bad())
            return seed;
        uint8_t active = presenceFromFractions(seed.beta);
        if (!seed.converged || active == 0)
        {
            active = chooseSinglePhase(P, T, z);
            seed = singlePhaseResult(P, T, z, active);
            if (err.bad())
                return seed;
        }
        Result selected = equilibrateActiveSet(P, T, z, active, seed);
        NN_TRACE("selected conv=%d pres=%d beta=%.17g %.17g %.17g\n", (int)selected.converged, (int)selected.presence, selected.beta[0], selected.beta[1], selected.beta[2]);
        if (err.bad())
            return selected;
        if (selected.converged)
        {
            canonicalizeFinal(P, T, selected);
            if (err.bad())
                return selected;
            if (selected.converged && validatePhysical(P, T, z, selected))
            {
                Result pref = preferStableSwReducedOverSingle(P, T, z, selected);
                if (err.bad() || pref.converged)
                    return pref;
            }
            if (err.bad())
                return selected;
            selected.converged = false;
        }

        Result full = flashThree(P, T, z, kg, kw);
        if (err.bad())
            return full;
        if (full.converged)
        {
            canonicalize(P, T, full);
            canonicalizeFinal(P, T, full);
            if (err.bad())
                return full;
            if (full.converged && validatePhysical(P, T, z, full))
            {
                Result pref = preferStableSwReducedOverSingle(P, T, z, full);
                if (err.bad() || pref.converged)
                    return pref;
            }
            if (err.bad())
                return full;
        }

        Result reduced = stableReducedSetFallback(P, T, z);
        if (err.bad())
            return reduced;
        if (reduced.converged)
        {
            canonicalizeFinal(P, T, reduced);
            if (err.bad())
                return reduced;
            if (reduced.converged && validatePhysical(P, T, z, reduced))
                return reduced;
            if (err.bad())
                return reduced;
        }

        if (sw())
        {
            // three_phase_flash.hpp:153-170: exact-mass allocation basin search
            allocationFallbackReached = true;
            Result allocated = flashThreeAllocationFallback(P, T, z);
            if (err.bad())
                return allocated;
            if (allocated.converged)
            {
                canonicalizeFinal(P, T, allocated);
                if (err.bad())
                    return allocated;
                if (allocated.converged && validatePhysical(P, T, z, allocated))
                {
                    const Result preferred = preferStableSwReducedOverSingle(P, T, z, allocated);
                    if (err.bad())
                        return preferred;
                    if (preferred.converged)
                        return preferred;
                }
                if (err.bad())
                    return allocated;
            }
        }
        done = false;
        return selected;
    }

    __host__ __device__ Result flash(double P, double T, const double *zin)
    {
        double z[NC];
        for (int c = 0; c < NC; ++c) z[c] = zin[c];
        bool done = false;
        Result selected = flashMain(P, T, z, done);
        if (done || err.bad())
            return selected;

        // three_phase_flash.hpp:168-204: SW role permutations
        if (sw() && enablePermutation)
        {
            const int perms[6][3] = {{kOil, kGas, kWater}, {kOil, kWater, kGas},
                                     {kGas, kOil, kWater}, {kGas, kWater, kOil},
                                     {kWater, kOil, kGas}, {kWater, kGas, kOil}};
            Result bestPermuted;
            double bestG = __builtin_inf();
            for (int k = 0; k < 6; ++k)
            {
                if (perms[k][0] == roles[0] && perms[k][1] == roles[1] && perms[k][2] == roles[2])
                    continue;
                Flash cand(f, o, err, perms[k], false, false);
                double zc[NC];
                for (int c = 0; c < NC; ++c) zc[c] = zin[c];
                bool d2 = false;
                cand.work = work;
                Result permuted = cand.flashMain(P, T, zc, d2);
                work = cand.work;
                allocationFallbackReached = allocationFallbackReached || cand.allocationFallbackReached;
                nearTie = nearTie || cand.nearTie;
                if (err.bad())
                    return permuted;
                if (!permuted.converged)
                    continue;
                const double g = cand.dimensionlessGibbs(P, T, permuted);
                if (err.bad())
                    return permuted;
                noteGibbsTie(g, bestG);
                if (finite(g) && g < bestG)
                {
                    bestG = g;
                    bestPermuted = permuted;
                }
            }
            if (bestPermuted.converged)
            {
                Result pref = preferStableSwReducedOverSingle(P, T, z, bestPermuted);
                if (err.bad() || pref.converged)
                    return pref;
            }
        }

        // three_phase_flash.hpp:206-224: pressure continuation
        if (enableContinuation)
        {
            Result cont = pressureContinuation(P, T, z);
            if (err.bad())
                return cont;
            if (cont.converged)
            {
                canonicalizeFinal(P, T, cont);
                if (err.bad())
                    return cont;
                if (cont.converged && validatePhysical(P, T, z, cont))
                {
                    Result pref = preferStableSwReducedOverSingle(P, T, z, cont);
                    if (err.bad() || pref.converged)
                        return pref;
                }
            }
        }
        return selected;
    }

    // restricted.inc:554-622
    __host__ __device__ Result pressureContinuation(double P, double T, const double *z)
    {
        Result failure;
        if (!(P > 0.0))
            return failure;
        const double ratios[8] = {0.95, 1.05, 0.90, 1.10, 0.80, 1.25, 0.65, 1.50};
        const double maxStep = 0.025;
        const int maxSteps = 12;
        Flash anchorSolver(f, o, err, roles, false, false);
        const double logTarget = ::log(P);
        for (int k = 0; k < 8; ++k)
        {
            const double anchorP = P * ratios[k];
            if (!(anchorP > 0.0) || !finite(anchorP))
                continue;
            double za[NC];
            for (int c = 0; c < NC; ++c) za[c] = z[c];
            bool d2 = false;
            anchorSolver.work = work;
            Result anchor = anchorSolver.flashMain(anchorP, T, za, d2);
            work = anchorSolver.work;
            allocationFallbackReached = allocationFallbackReached || anchorSolver.allocationFallbackReached;
            nearTie = nearTie || anchorSolver.nearTie;
            if (err.bad())
                return anchor;
            if (!anchor.converged)
                continue;
            Result seed = sw() ? numericalSwRoleResult(anchor) : anchor;
            if (!seed.converged || seed.presence == 0)
                continue;
            const double logAnchor = ::log(anchorP);
            const double span = ::fabs(logTarget - logAnchor);
            const int steps = clampi(int(::ceil(span / maxStep)), 1, maxSteps);
            int totalIt = anchor.iterations;
            bool ok = true;
            for (int s = 1; s <= steps; ++s)
            {
                const double frac = double(s) / steps;
                const double stepP = ::exp(logAnchor + frac * (logTarget - logAnchor));
                Result current = equilibrateActiveSet(stepP, T, z, seed.presence, seed, false);
                if (err.bad())
                    return current;
                if (!current.converged)
                {
                    ok = false;
                    break;
                }
                totalIt += current.iterations;
                seed = current;
            }
            if (ok && seed.converged)
            {
                seed.iterations = totalIt;
                return seed;
            }
        }
        return failure;
    }
};

// restricted.inc:8-61
template <int NC, int Mode>
__host__ __device__ FlashResult<NC> Flash<NC, Mode>::flashRestrictedImpl(double P, double T, const double *zin,
                                                                 uint8_t allowed, const double (*seed)[NC])
{
    Result fail;
    if (!validatePT(P, T))
        return fail;
    double z[NC];
    for (int c = 0; c < NC; ++c) z[c] = zin[c];
    normalize(z);
    if (allowed == 0)
    {
        NN_TRACE("restricted flash with nothing allowed\n");
        err.raise(kErrInvalid);
        return fail;
    }
    if (err.bad())
        return fail;
    const int count = popcount3(allowed);
    if (count == 3)
    {
        // The reference forwards an all-phase restriction to the unrestricted
        // flash (restricted.inc:20-21).  None of its callers reaches this
        // branch (restricted flashes are requested for reduced sets only), and
        // the public all-phase request is routed to flash() by the kernel, so
        // it is rejected here: calling flash() from inside it would make the
        // device call graph recursive and its stack size unbounded.
        NN_TRACE("restricted flash with all phases allowed\n");
        err.raise(kErrInvalid);
        return fail;
    }
    if (count == 1)
        return singlePhaseResult(P, T, z, allowed);
    int a, b;
    activePair(allowed, a, b);
    double K[NC];
    pairInitialK(a, b, P, T, z, K);
    if (seed != nullptr)
    {
        double first[NC], second[NC];
        for (int c = 0; c < NC; ++c) { first[c] = seed[a][c]; second[c] = seed[b][c]; }
        normalize(first);
        normalize(second);
        for (int c = 0; c < NC; ++c)
            K[c] = clampd(second[c] / maxd(first[c], o.floor), 1.0e-100, 1.0e100);
    }
    Result r = flashTwo(P, T, z, a, b, K);
    if (err.bad())
        return r;
    if (!r.converged && seed != nullptr)
    {
        pairInitialK(a, b, P, T, z, K);
        r = flashTwo(P, T, z, a, b, K);
        if (err.bad())
            return r;
    }
    if (r.converged)
        canonicalize(P, T, r);
    return r;
}

} // namespace nn
