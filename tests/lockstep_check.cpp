// Host check of the phase update's three-pass stability test (test tooling).
//
// For every cell, runs this library's phase update (nonenone_amd/csrc/
// phase_update.hpp compiled for the host) twice:
//   serial:   phaseUpdateCell with the seed trials run inside stabilityTest;
//   lockstep: phaseUpdateCell in kStabRecord mode, then the trial state machine
//             of Flash::trialInit/trialPoint/trialStep/trialFold driven one
//             residual evaluation at a time (as one lane of the GPU's lockstep
//             pass), then phaseUpdateCell in kStabReplayHoldReflash mode and,
//             for a cell whose re-flash was held back, kStabReplay (the GPU's
//             re-flash pass);
// and prints one line per cell: <index> <pending> <identical>, where identical
// is 1 when primary, phase state, status and flags agree bit for bit.
//
// Usage: lockstep_check <fluid_desc.bin> <cells.bin>
//   cells.bin: int64 n, int64 has_well, then T[n], primary[n][NP], phase_state[n][NPS].
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../nonenone_amd/csrc/phase_update.hpp"
#include "host_params.hpp"

using namespace nn;

template <int NC, bool Well>
int run(const FluidParams &fp, const FlashOptions &op, int64_t n, const double *T,
        const double *prim0, const double *ps0)
{
    using L = Layout<NC, Well>;
    using F = Flash<NC, kModePlain>;
    std::vector<StabRecord<NC>> recs(static_cast<size_t>(n));
    int mismatches = 0, held = 0;
    for (int64_t s = 0; s < n; ++s)
    {
        // each cell on its own copies, as item 0 of a one-cell batch
        double pa[L::NP], sa[L::NPS], pb[L::NP], sb[L::NPS];
        std::memcpy(pa, prim0 + s * L::NP, sizeof pa);
        std::memcpy(sa, ps0 + s * L::NPS, sizeof sa);
        std::memcpy(pb, pa, sizeof pb);
        std::memcpy(sb, sa, sizeof sb);
        int32_t sta = -1, stb = -1;
        uint8_t fa = 0, fb = 0;
        StabRecord<NC> *rec = recs.data() + s;
        phaseUpdateCell<NC, Well, kModePlain>(fp, op, T[s], pa, sa, 0, &sta, &fa);
        int r = phaseUpdateCell<NC, Well, kModePlain>(fp, op, T[s], pb, sb, 0, &stb, &fb,
                                                      kStabRecord, rec);
        const bool pending = r == kItemPending;
        if (pending)
        {
            Err err;
            F fl(fp, op, err);
            typename F::TrialMachine m;
            bool done = fl.trialInit(*rec, m);
            while (!done)
            {
                double x[NC];
                fl.trialPoint(m, x);
                typename F::StabResidual res;
                fl.evaluateStabilityResidual(m.cand, rec->P, rec->T, rec->logRef, x, res);
                typename F::Trial t;
                done = fl.trialStep(*rec, m, res, t) && fl.trialFold(*rec, m, t);
            }
            r = phaseUpdateCell<NC, Well, kModePlain>(fp, op, T[s], pb, sb, 0, &stb, &fb,
                                                      kStabReplayHoldReflash, rec);
            if (r == kItemReflash)
            {
                ++held;
                r = phaseUpdateCell<NC, Well, kModePlain>(fp, op, T[s], pb, sb, 0, &stb, &fb,
                                                          kStabReplay, rec);
            }
        }
        const bool same = r == kItemDone && sta == stb && fa == fb &&
                          std::memcmp(pa, pb, sizeof pa) == 0 && std::memcmp(sa, sb, sizeof sa) == 0;
        if (!same)
            ++mismatches;
        std::printf("%lld %d %d\n", (long long)s, pending ? 1 : 0, same ? 1 : 0);
    }
    std::fprintf(stderr, "re-flashes held back and finished: %d\n", held);
    return mismatches;
}

template <bool Well>
int dispatch(const FluidParams &fp, const FlashOptions &op, int64_t n, const double *T,
             const double *prim, const double *ps)
{
    switch (fp.nc)
    {
    case 3: return run<3, Well>(fp, op, n, T, prim, ps);
    case 6: return run<6, Well>(fp, op, n, T, prim, ps);
    default: std::fprintf(stderr, "nc %d not built into the check\n", fp.nc); return -1;
    }
}

static std::vector<char> slurp(const char *path)
{
    std::vector<char> b;
    FILE *f = std::fopen(path, "rb");
    if (!f)
        return b;
    std::fseek(f, 0, SEEK_END);
    b.resize(size_t(std::ftell(f)));
    std::fseek(f, 0, SEEK_SET);
    if (std::fread(b.data(), 1, b.size(), f) != b.size())
        b.clear();
    std::fclose(f);
    return b;
}

int main(int argc, char **argv)
{
    if (argc < 3)
        return 2;
    const std::vector<char> fb = slurp(argv[1]), cb = slurp(argv[2]);
    if (fb.size() != sizeof(nn_fluid_desc) || cb.size() < 16)
        return 2;
    nn_fluid_desc d;
    std::memcpy(&d, fb.data(), sizeof d);
    const FluidParams fp = paramsFrom(d);
    const FlashOptions op = defaultOptions();
    int64_t n, well;
    std::memcpy(&n, cb.data(), 8);
    std::memcpy(&well, cb.data() + 8, 8);
    const int nc = fp.nc;
    const int np = 3 * nc + 1 + (well ? 1 : 0);
    const double *T = reinterpret_cast<const double *>(cb.data() + 16);
    const double *prim = T + n;
    const double *ps = prim + n * np;
    const int bad = well ? dispatch<true>(fp, op, n, T, prim, ps)
                         : dispatch<false>(fp, op, n, T, prim, ps);
    return bad == 0 ? 0 : 1;
}
