// host_params.hpp — host-side nn_fluid_desc -> FluidParams conversion and the
// reference's default flash options, for the host builds of the flash used by
// the test tooling (tests/golden/sw_sensitivity.cpp, tests/lockstep_check.cpp,
// tests/coop_emulation.cpp, tests/phase_update_dump.cpp).
#pragma once

#include <cmath>

#include "../nonenone_amd/csrc/common.hpp"

namespace nn
{

inline FluidParams paramsFrom(const nn_fluid_desc &d)
{
    FluidParams p{};
    p.nc = d.nc;
    p.model = d.thermo_model;
    p.eosType = d.eos_type;
    p.water = d.water_component;
    p.omegaA = d.omega_a;
    p.omegaB = d.omega_b;
    p.m1 = d.m1;
    p.m2 = d.m2;
    for (int i = 0; i < NN_MAX_NC; ++i)
    {
        p.tc[i] = d.tc[i];
        p.pc[i] = d.pc[i];
        p.vc[i] = d.vc[i];
        p.omega[i] = d.acentric[i];
        p.mw[i] = d.molar_mass[i];
        p.swAqKij[i] = d.sw_aqueous_kij[i];
        p.swAqSet[i] = d.sw_aqueous_kij_set[i];
        p.cpaA0[i] = d.cpa_a0[i];
        p.cpaB[i] = d.cpa_b[i];
        p.cpaC1[i] = d.cpa_c1[i];
        p.cpaEps[i] = d.cpa_epsilon[i];
        p.cpaBeta[i] = d.cpa_beta[i];
        p.cpaDonors[i] = d.cpa_donors[i];
        p.cpaAcceptors[i] = d.cpa_acceptors[i];
        for (int j = 0; j < NN_MAX_NC; ++j)
        {
            p.kij[i][j] = d.kij[i * NN_MAX_NC + j];
            p.cpaCrossEps[i][j] = d.cpa_cross_epsilon[i * NN_MAX_NC + j];
            p.cpaCrossBeta[i][j] = d.cpa_cross_beta[i * NN_MAX_NC + j];
        }
    }
    // LBC constants (compositional_properties.hpp:166-182), as the library's set-up
    for (int i = 0; i < d.nc; ++i)
    {
        const double mwSqrt = std::sqrt(1000.0 * d.molar_mass[i]);
        p.lbcMwSqrt[i] = mwSqrt;
        p.lbcEps[i] = 5.4402 * std::pow(d.tc[i] / kKelvinPerRankine, 1.0 / 6.0) /
                      (mwSqrt * std::pow(d.pc[i] / kPsi, 2.0 / 3.0) * kCentipoise);
    }
    for (int k = 0; k < 7; ++k) p.relperm[k] = d.relperm[k];
    p.swSalinity = d.sw_salinity;
    p.cpaTerm = d.cpa_physical_term;
    p.cpaMaxIter = d.cpa_max_iterations;
    p.cpaTol = d.cpa_tolerance;
    p.cpaDamping = d.cpa_damping;
    p.aqVolume = d.aqueous_volume;
    p.aqViscosity = d.aqueous_viscosity;
    p.co2 = d.co2_component;
    p.waterMolarMass = d.water_molar_mass;
    p.aqVolumeMax = d.aqueous_volume_max_unsupported;
    p.aqViscosityMax = d.aqueous_viscosity_max_unsupported;
    return p;
}

// Reference defaults (three_phase_flash_types.hpp:21-45).
inline FlashOptions defaultOptions()
{
    return FlashOptions{120, 80, 80, 8, 12, 8, 12, 10, 12, 1.0e-9, 1.0e-12, 1.0e-9,
                        1.0e-10, 1.0e-8, 1.0e-30, 2.0, 1.0e-6};
}

} // namespace nn
