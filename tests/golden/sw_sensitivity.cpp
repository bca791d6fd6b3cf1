// Libm-conditioning probe for flash fixtures (test tooling, host only).
// (SW and CPA fluids; the file name is historical to its first use.)
//
// Runs this library's flash (nonenone_amd/csrc/flash.hpp, compiled for the
// host, where it reproduces the reference build to ~1e-15) once with the host
// libm and then under several seeded 1-ulp perturbations of exp/log/pow/cos/acos
// results.  A state whose result moves by more than round-off under such
// perturbations is conditioned by the math library itself: any two libms that
// are both within 1 ulp (glibc on the host, the device library on the GPU)
// can legitimately give values that differ there by more than the parity
// tolerance.  Output, one line per state:
//     <index> <presence> <label_flips> <worst_relative_change> <allocation_recovery> <near_tie>
//     <n_outcomes> <beta of outcome 1 (3 values)> ... <beta of outcome n>
// where allocation_recovery is 1 when the unperturbed flash reached the
// Soreide-Whitson exact-mass allocation recovery (allocation.inc:609-700) and
// near_tie is 1 when one of its lowest-Gibbs selections compared two
// candidates whose reduced Gibbs energies agree to 1e-12 (equivalent
// solutions, chosen by the last bits of the arithmetic).  Such states get
// `sensitive_salts` perturbed runs instead of `salts` (the alternative answers
// can be rare), and the distinct phase-fraction vectors of the runs that keep
// the unperturbed label are listed as outcomes (the first is the unperturbed
// one; at most kMaxOutcomes).  worst_relative_change is taken over the runs
// that stay in the unperturbed outcome (the spread of that answer).
//
// The same for the hysteretic phase update (phase_update.hpp), with the
// primary vector compared instead of the flash record.
//
// Usage: sw_sensitivity flash <fluid_desc.bin> <states.bin> [salts [sensitive_salts]]
//        sw_sensitivity update <fluid_desc.bin> <cells.bin> [salts [sensitive_salts]]
//   fluid_desc.bin: raw nn_fluid_desc; states.bin: int64 n, then P[n], T[n], z[n][nc];
//   cells.bin: int64 n, int64 has_well, then T[n], primary[n][NP], phase_state[n][NPS].
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

static uint64_t g_calls = 0, g_salt = 0;
static int g_mode = 0;
static int g_salts = 8;            // perturbed re-runs per state
static int g_sensitiveSalts = 64;  // for allocation-recovery and near-tie states
constexpr int kMaxOutcomes = 8;

// Distinct phase-fraction vectors (max abs difference above 1e-9).
struct Outcomes
{
    int n = 0;
    double beta[kMaxOutcomes][3];
    // index of the outcome b belongs to (appended when new; -1 when full)
    int add(const double *b)
    {
        for (int k = 0; k < n; ++k)
            if (std::fabs(beta[k][0] - b[0]) <= 1e-9 && std::fabs(beta[k][1] - b[1]) <= 1e-9 &&
                std::fabs(beta[k][2] - b[2]) <= 1e-9)
                return k;
        if (n == kMaxOutcomes)
            return -1;
        for (int p = 0; p < 3; ++p) beta[n][p] = b[p];
        return n++;
    }
    void print() const
    {
        std::printf(" %d", n);
        for (int k = 0; k < n; ++k)
            std::printf(" %.17g %.17g %.17g", beta[k][0], beta[k][1], beta[k][2]);
        std::printf("\n");
    }
};
static inline double bump(double r)
{
    if (!g_mode)
        return r;
    uint64_t h = (++g_calls + g_salt * 0x9E3779B1ull) * 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
    if ((h & 3) == 0)
        return std::nextafter(r, (h & 4) ? INFINITY : -INFINITY);
    return r;
}
static double noisyExp(double x) { return bump(std::exp(x)); }
static double noisyLog(double x) { return bump(std::log(x)); }
static double noisyPow(double x, double y) { return bump(std::pow(x, y)); }
static double noisyCos(double x) { return bump(std::cos(x)); }
static double noisyAcos(double x) { return bump(std::acos(x)); }
// the fluid set-up (host_params.hpp: LBC constants) uses the unperturbed libm,
// as the library's does
#include "../host_params.hpp"
#define exp noisyExp
#define log noisyLog
#define pow noisyPow
#define cos noisyCos
#define acos noisyAcos
#include "../../nonenone_amd/csrc/flash.hpp"
#include "../../nonenone_amd/csrc/phase_update.hpp"
#undef exp
#undef log
#undef pow
#undef cos
#undef acos

#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace nn;

template <int NC>
static double change(const FlashResult<NC> &a, const FlashResult<NC> &b)
{
    auto rel = [](double x, double y, double floor) {
        if (x == y)
            return 0.0;
        const double s = std::fmax(std::fmax(std::fabs(x), std::fabs(y)), floor);
        return std::fabs(x - y) / s;
    };
    double m = 0.0;
    for (int p = 0; p < 3; ++p)
    {
        m = std::fmax(m, rel(a.beta[p], b.beta[p], 1.0));
        m = std::fmax(m, rel(a.sat[p], b.sat[p], 1.0));
        if (!(a.presence & (1u << p)))
            continue;
        m = std::fmax(m, rel(a.Z[p], b.Z[p], 1.0));
        m = std::fmax(m, rel(a.rho[p], b.rho[p], 1.0));
        for (int c = 0; c < NC; ++c)
            m = std::fmax(m, rel(a.x[p][c], b.x[p][c], 1.0));
    }
    return m;
}

template <int NC>
static int run(const FluidParams &fp, const FlashOptions &op, int64_t n, const double *P,
               const double *T, const double *z)
{
    for (int64_t s = 0; s < n; ++s)
    {
        g_mode = 0;
        Err e0;
        Flash<NC> f0(fp, op, e0);
        const FlashResult<NC> r0 = f0.flash(P[s], T[s], z + s * NC);
        int flips = 0;
        double worst = 0.0;
        Outcomes out;
        out.add(r0.beta);
        const int salts =
            (f0.allocationFallbackReached || f0.nearTie) ? g_sensitiveSalts : g_salts;
        for (int salt = 1; salt <= salts; ++salt)
        {
            g_mode = 1;
            g_salt = salt;
            g_calls = 0;
            Err e1;
            Flash<NC> f1(fp, op, e1);
            const FlashResult<NC> r1 = f1.flash(P[s], T[s], z + s * NC);
            if (e1.code != e0.code || r1.converged != r0.converged || r1.presence != r0.presence)
            {
                ++flips;
                continue;
            }
            // the band is the spread within the unperturbed answer's outcome;
            // runs that land on another split are listed as outcomes instead
            if (out.add(r1.beta) == 0)
                worst = std::fmax(worst, change(r0, r1));
        }
        std::printf("%lld %d %d %.6e %d %d", (long long)s, (int)r0.presence, flips, worst,
                    int(f0.allocationFallbackReached), int(f0.nearTie));
        out.print();
    }
    return 0;
}

template <int NC, bool Well>
static int runUpdate(const FluidParams &fp, const FlashOptions &op, int64_t n, FILE *fs)
{
    using L = Layout<NC, Well>;
    std::vector<double> T(n), prim(size_t(n) * L::NP), ps(size_t(n) * L::NPS);
    if (std::fread(T.data(), sizeof(double), n, fs) != size_t(n) ||
        std::fread(prim.data(), sizeof(double), prim.size(), fs) != prim.size() ||
        std::fread(ps.data(), sizeof(double), ps.size(), fs) != ps.size())
        return 5;
    for (int64_t s = 0; s < n; ++s)
    {
        double p0[L::NP], p1[L::NP];
        int alloc = 0, tie = 0, flips = 0, label0 = -1;
        double worst = 0.0;
        Outcomes out;
        int salts = g_salts;
        for (int salt = 0; salt <= salts; ++salt)
        {
            g_mode = salt > 0;
            g_salt = salt;
            g_calls = 0;
            double *p = salt == 0 ? p0 : p1;
            for (int k = 0; k < L::NP; ++k) p[k] = prim[size_t(s) * L::NP + k];
            PhaseStateData<NC> st;
            if (!decodePhaseState<NC, Well>(&ps[size_t(s) * L::NPS], st))
                break;
            Err err;
            PhaseUpdater<NC, Well> up(fp, op, err, T[s]);
            up.sanitize(p);
            const int packed = up.update(p, st);
            const int label = err.bad() ? -1 - err.code : (packed >> 8) * 16 + st.presence;
            if (salt == 0)
            {
                label0 = label;
                alloc = int(up.flash.allocationFallbackReached);
                tie = int(up.flash.nearTie);
                if (alloc || tie)
                    salts = g_sensitiveSalts;
                out.add(st.beta);
                continue;
            }
            if (label != label0)
            {
                ++flips;
                continue;
            }
            if (out.add(st.beta) != 0)
                continue;
            for (int k = 0; k < L::NP; ++k)
            {
                if (p0[k] == p1[k])
                    continue;
                const double sc = std::fmax(std::fmax(std::fabs(p0[k]), std::fabs(p1[k])), 1.0);
                worst = std::fmax(worst, std::fabs(p0[k] - p1[k]) / sc);
            }
        }
        std::printf("%lld %d %d %.6e %d %d", (long long)s, label0, flips, worst, alloc, tie);
        out.print();
    }
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 4 || argc > 6)
        return 2;
    if (argc >= 5)
        g_salts = std::atoi(argv[4]);
    if (argc == 6)
        g_sensitiveSalts = std::atoi(argv[5]);
    const bool update = std::strcmp(argv[1], "update") == 0;
    nn_fluid_desc d;
    FILE *fd = std::fopen(argv[2], "rb");
    if (!fd || std::fread(&d, sizeof(d), 1, fd) != 1)
        return 3;
    std::fclose(fd);
    const FluidParams fp0 = paramsFrom(d);
    const FlashOptions op0{120, 80, 80, 8, 12, 8, 12, 10, 12, 1.0e-9, 1.0e-12, 1.0e-9,
                           1.0e-10, 1.0e-8, 1.0e-30, 2.0, 1.0e-6};
    if (update)
    {
        FILE *fc = std::fopen(argv[3], "rb");
        int64_t hdr[2];
        if (!fc || std::fread(hdr, sizeof(int64_t), 2, fc) != 2 || hdr[0] < 0)
            return 4;
        int rc = 6;
        if (d.nc == 3 && hdr[1]) rc = runUpdate<3, true>(fp0, op0, hdr[0], fc);
        else if (d.nc == 3) rc = runUpdate<3, false>(fp0, op0, hdr[0], fc);
        else if (d.nc == 6 && !hdr[1]) rc = runUpdate<6, false>(fp0, op0, hdr[0], fc);
        std::fclose(fc);
        return rc;
    }
    FILE *fs = std::fopen(argv[3], "rb");
    int64_t n = 0;
    if (!fs || std::fread(&n, sizeof(n), 1, fs) != 1 || n < 0)
        return 4;
    std::vector<double> P(n), T(n), z(size_t(n) * d.nc);
    if (std::fread(P.data(), sizeof(double), n, fs) != size_t(n) ||
        std::fread(T.data(), sizeof(double), n, fs) != size_t(n) ||
        std::fread(z.data(), sizeof(double), z.size(), fs) != z.size())
        return 5;
    std::fclose(fs);
    const FluidParams fp = paramsFrom(d);
    // Reference defaults (three_phase_flash_types.hpp:21-45).
    const FlashOptions op{120, 80, 80, 8, 12, 8, 12, 10, 12, 1.0e-9, 1.0e-12, 1.0e-9,
                          1.0e-10, 1.0e-8, 1.0e-30, 2.0, 1.0e-6};
    switch (d.nc)
    {
    case 3: return run<3>(fp, op, n, P.data(), T.data(), z.data());
    case 6: return run<6>(fp, op, n, P.data(), T.data(), z.data());
    default: return 6;
    }
}
