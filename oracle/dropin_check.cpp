// ORACLE — test infrastructure only.  Built from the reference's own headers
// (oracle/Makefile target `dropin`) into oracle/_ref/dropin_check.
//
// Drop-in check of the FlashBackend seam: the reference's
// FullyCompositionalThreePhaseEquilibrium::updatePhaseState
// (models/include/natural/state/three_phase_equilibrium.hpp:282-496) runs every
// cell twice — with its default CubicThreePhaseFlash backend and with
// GpuThreePhaseFlash (include/nonenone_amd_flash_backend.hpp), i.e. the
// reference's own phase-state logic calling this library's GPU flash,
// restricted flash and stability test through the C ABI — and compares the
// outcomes: statuses, flags, phase labels exactly, primary variables at 1e-10.
//
// Usage: dropin_check <cells.bin>
//   cells.bin: raw nn_fluid_desc, int64 n, then T[n], primary[n][NP], phase_state[n][NPS]
//   (nc = 3, no well unknown).  Prints one summary line; exit 0 when all match.
#include <indices/indices.hpp>
#include <natural/fluid_system.hpp>
#include <natural/petsc/phase_state_codec.hpp>
#include <natural/state/three_phase_equilibrium.hpp>
#include <natural/thermo/three_phase_flash.hpp>

#include "../include/nonenone_amd.h"
#include "../include/nonenone_amd_flash_backend.hpp"
#include "ref_fluid.hpp"

#include <cmath>
#include <cstdio>
#include <vector>

namespace
{

using I = oracle::AdIdx<3, false>;
constexpr int NP = I::numPrimaryVariables;
constexpr int NPS = I::numPhaseStateVariables;

struct Outcome
{
    int status = 0;
    int flags = 0;
    std::array<double, NP> prim{};
    std::array<double, NPS> ps{};
};

template <class Eq>
Outcome run(Eq &eq, MPMC::FluidSystem<I> &fluid, double T, const double *prim0, const double *ps0)
{
    Outcome o;
    fluid.temperature = T;
    for (int k = 0; k < NP; ++k) o.prim[k] = prim0[k];
    for (int k = 0; k < NPS; ++k) o.ps[k] = ps0[k];
    try
    {
        auto data = MPMC::PetscPhaseStateCodec<I>::decode(o.ps);
        eq.sanitizePrimaryBeforeFlash(o.prim, false);
        const auto r = eq.updatePhaseState(o.prim, data);
        o.ps = MPMC::PetscPhaseStateCodec<I>::encode(data);
        o.status = r.status == MPMC::PhaseUpdateStatus::RecoverableThermodynamicFailure ? 4 : 0;
        o.flags = (r.phaseRemoved ? 1 : 0) | (r.stabilityInvalid ? 2 : 0) |
                  (r.missingPhaseUnstable ? 4 : 0) | (r.restrictedFlashFailed ? 8 : 0) |
                  (r.unrestrictedFlashFailed ? 16 : 0) |
                  (r.status == MPMC::PhaseUpdateStatus::StableReducedSet ? 32 : 0);
    }
    catch (const std::invalid_argument &)
    {
        o.status = 3;
    }
    catch (const std::exception &)
    {
        o.status = 2;
    }
    return o;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc != 2)
        return 2;
    FILE *f = std::fopen(argv[1], "rb");
    nn_fluid_desc desc;
    int64_t n = 0;
    if (!f || std::fread(&desc, sizeof(desc), 1, f) != 1 || std::fread(&n, sizeof(n), 1, f) != 1 ||
        n <= 0 || desc.nc != 3)
        return 3;
    std::vector<double> T(n), prim(size_t(n) * NP), ps(size_t(n) * NPS);
    if (std::fread(T.data(), 8, n, f) != size_t(n) ||
        std::fread(prim.data(), 8, prim.size(), f) != prim.size() ||
        std::fread(ps.data(), 8, ps.size(), f) != ps.size())
        return 4;
    std::fclose(f);

    GpuThreePhaseFlash<I>::setFluid(desc);
    auto fluidRef = oracle::makeFluid<I>(desc, 300.0);
    auto fluidGpu = oracle::makeFluid<I>(desc, 300.0);
    MPMC::FullyCompositionalThreePhaseEquilibrium<I> reference(fluidRef);
    MPMC::FullyCompositionalThreePhaseEquilibrium<I, GpuThreePhaseFlash<I>> gpu(fluidGpu);

    int64_t statusMismatch = 0, flagMismatch = 0, labelMismatch = 0, valueMismatch = 0;
    double worst = 0.0;
    for (int64_t s = 0; s < n; ++s)
    {
        const Outcome a = run(reference, fluidRef, T[s], &prim[s * NP], &ps[s * NPS]);
        const Outcome b = run(gpu, fluidGpu, T[s], &prim[s * NP], &ps[s * NPS]);
        if (a.status != b.status)
        {
            ++statusMismatch;
            continue;
        }
        if (a.flags != b.flags)
            ++flagMismatch;
        if (a.ps[0] != b.ps[0] || a.ps[NPS - 1] != b.ps[NPS - 1])
        {
            ++labelMismatch;
            continue;
        }
        double d = 0.0;
        for (int k = 0; k < NP; ++k)
        {
            const double scale = std::fmax(std::fmax(std::fabs(a.prim[k]), std::fabs(b.prim[k])), 1.0);
            d = std::fmax(d, std::fabs(a.prim[k] - b.prim[k]) / scale);
        }
        worst = std::fmax(worst, d);
        if (d > 1.0e-10)
            ++valueMismatch;
    }
    std::printf("dropin cells=%lld status_mismatch=%lld flag_mismatch=%lld label_mismatch=%lld "
                "value_mismatch=%lld max_rel=%.3e\n",
                (long long)n, (long long)statusMismatch, (long long)flagMismatch,
                (long long)labelMismatch, (long long)valueMismatch, worst);
    return (statusMismatch || flagMismatch || labelMismatch || valueMismatch) ? 1 : 0;
}
