// wells.hpp — perforation sources of the Natural runtime's wells.
//
// Follows (paths relative to the reference checkout):
//   computePerforationWellSourceInto  models/include/natural/physics/well_source.hpp:37-133
//                                     (the fully compositional three-phase branch)
//   wellControlResidual (BHP)         well_source.hpp:310-327
// A well's bottom-hole pressure is the well-pressure unknown of its
// representative cell (natural_petsc_runtime_wells.inc:8-49).
#pragma once

#include "common.hpp"

namespace nn
{

// One well on the device (nn_well without the perforation range).
struct DeviceWell
{
    int type;             // NN_WELL_INJECTOR / NN_WELL_PRODUCER
    int64_t bhpCell;
    double target;
    double injPhase[3];
    double injMass[kMaxNc];
};

// Component mass sources of one perforation (injection positive, production
// negative): Delta p_a = p - p_bhp, b_a = rho_a / rho_a,sc; a phase that flows
// into the reservoir through an injector carries WI * total mobility * Delta p /
// sum(f_a / b_a) * f_a at surface conditions, any other phase WI * lambda_a *
// Delta p * b_a (both with a minus sign); reservoir rate q_sc / b_a, mass rate
// q_res * rho_a; a component's source takes the injection mass fractions for
// inflowing phases and the cell's phase mass fractions otherwise.
// rho: phase mass densities, mob: phase mobilities, w: phase mass fractions.
template <int NC, class S>
__host__ __device__ void perforationSource(const DeviceWell &wl, double wellIndex,
                                           const double *surfaceDensity, const S &pressure,
                                           const S &bhp, const S *rho, const S *mob,
                                           const S (*w)[NC], S *source, Err &err)
{
    S dP[3], b[3];
    S volumeRatio = 0.0;
    S totalMobility = 0.0;
    for (int p = 0; p < 3; ++p)
    {
        dP[p] = pressure - bhp;
        b[p] = rho[p] / surfaceDensity[p];
        volumeRatio += wl.injPhase[p] / b[p];
        totalMobility += mob[p];
    }
    S massRate[3];
    for (int p = 0; p < 3; ++p)
    {
        const bool injecting = wl.type == NN_WELL_INJECTOR && sval(dP[p]) < 0.0;
        S surfaceRate;
        if (injecting)
        {
            if (!(sval(volumeRatio) > 0.0))
            {
                err.raise(kErrThermo);   // the reference throws std::runtime_error
                return;
            }
            surfaceRate = -wellIndex * totalMobility * dP[p] / volumeRatio * wl.injPhase[p];
        }
        else
        {
            surfaceRate = -wellIndex * mob[p] * dP[p] * b[p];
        }
        const S reservoirRate = surfaceRate / b[p];
        massRate[p] = reservoirRate * rho[p];
    }
    for (int c = 0; c < NC; ++c)
    {
        S s = 0.0;
        for (int p = 0; p < 3; ++p)
        {
            if (sval(massRate[p]) > 0.0)
                s += massRate[p] * wl.injMass[c];
            else
                s += massRate[p] * w[p][c];
        }
        source[c] = s;
    }
}

} // namespace nn
