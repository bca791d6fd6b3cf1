// ORACLE — test infrastructure only (see ref_harness.cpp).
//
// ref_fluid.hpp: the reference's FluidSystem / flash options for a
// nn_fluid_desc, shared by the oracle's reference harness and its drop-in
// check.  Include after the reference's headers.
#pragma once

#include "../include/nonenone_amd.h"

#include <array>
#include <string>

namespace oracle
{

template <int NC, bool Well>
using Cfg = MPMC::CompositionalModelConfig<
    NC, true, Well, false, false, false,
    MPMC::PhaseBehaviorModel::FullyCompositionalThreePhase>;
template <int NC, bool Well>
using AdIdx = MPMC::ADIndices<Cfg<NC, Well>>;
template <int NC>
using ScIdx = MPMC::ScalarIndices<Cfg<NC, false>>;

template <class Indices>
MPMC::FluidSystem<Indices> makeFluid(const nn_fluid_desc &d, double temperature)
{
    constexpr int N = Indices::numComponents;
    using Arr = std::array<double, N>;
    Arr tc{}, pc{}, vc{}, om{}, mw{};
    std::array<Arr, N> kij{};
    for (int i = 0; i < N; ++i)
    {
        tc[i] = d.tc[i];
        pc[i] = d.pc[i];
        vc[i] = d.vc[i];
        om[i] = d.acentric[i];
        mw[i] = d.molar_mass[i];
        for (int j = 0; j < N; ++j)
            kij[i][j] = d.kij[i * NN_MAX_NC + j];
    }
    MPMC::CompositionalMixture<Indices> mixture(tc, pc, vc, om, mw, kij);
    MPMC::CubicEquationOfState<Indices> eos(
        d.omega_a, d.omega_b, std::move(mixture), d.eos_type, d.m1, d.m2);
    if (d.thermo_model == NN_MODEL_SW)
    {
        typename MPMC::CubicEquationOfState<Indices>::SoreideWhitsonOptions sw;
        sw.waterComponent = d.water_component;
        sw.salinityMolality = d.sw_salinity;
        for (int c = 0; c < N; ++c)
        {
            if (d.sw_aqueous_kij_set[c])
            {
                const double k = d.sw_aqueous_kij[c];
                sw.aqueousWaterBip[c] = [k](double, double) { return k; };
            }
        }
        eos.configureSoreideWhitson(std::move(sw));
    }
    else if (d.thermo_model == NN_MODEL_CPA)
    {
        typename MPMC::CubicEquationOfState<Indices>::CubicPlusAssociationOptions cpa;
        for (int c = 0; c < N; ++c)
        {
            cpa.a0[c] = d.cpa_a0[c];
            cpa.b[c] = d.cpa_b[c];
            cpa.c1[c] = d.cpa_c1[c];
            cpa.associationEnergy[c] = d.cpa_epsilon[c];
            cpa.associationVolume[c] = d.cpa_beta[c];
            cpa.donorSites[c] = d.cpa_donors[c];
            cpa.acceptorSites[c] = d.cpa_acceptors[c];
            for (int e = 0; e < N; ++e)
            {
                cpa.crossAssociationEnergy[c][e] = d.cpa_cross_epsilon[c * NN_MAX_NC + e];
                cpa.crossAssociationVolume[c][e] = d.cpa_cross_beta[c * NN_MAX_NC + e];
            }
        }
        cpa.physicalTerm = d.cpa_physical_term == 1
            ? MPMC::CpaCubicPhysicalTerm::PengRobinson
            : MPMC::CpaCubicPhysicalTerm::SoaveRedlichKwong;
        cpa.radialDistribution = MPMC::CpaRadialDistribution::Simplified;
        if (d.cpa_max_iterations > 0)
            cpa.maximumAssociationIterations = d.cpa_max_iterations;
        if (d.cpa_tolerance > 0.0)
            cpa.associationTolerance = d.cpa_tolerance;
        if (d.cpa_damping > 0.0)
            cpa.associationDamping = d.cpa_damping;
        eos.configureCubicPlusAssociation(std::move(cpa));
    }
    // Water-role closures, as case/include/case/fluid_factory.hpp:278-296, 447-461
    // configures them
    if (d.aqueous_volume)
    {
        typename MPMC::CubicEquationOfState<Indices>::AqueousVolumeOptions aqueous;
        aqueous.waterComponent = d.water_component;
        aqueous.co2Component = d.co2_component;
        aqueous.waterMolarMass = d.water_molar_mass;
        aqueous.maximumUnsupportedMoleFraction = d.aqueous_volume_max_unsupported;
        eos.configureAqueousVolume(aqueous);
    }
    if (d.aqueous_viscosity == NN_AQ_VISCOSITY_IAPWS2008)
        eos.configureAqueousCompositionDomain(d.water_component,
                                              d.aqueous_viscosity_max_unsupported);

    typename MPMC::FluidSystem<Indices>::ComponentNames names{};
    for (int i = 0; i < N; ++i)
        names[i] = "c" + std::to_string(i);
    typename MPMC::FluidSystem<Indices>::PhaseNames phases{"Oil", "Gas", "Water"};
    MPMC::FluidSystem<Indices> fluid(
        {730.0, 1.8, 985.4}, {2.4e-4, 1.86e-5, 4.69e-4}, std::move(eos),
        names, phases, temperature);
    if (d.water_component >= 0)
        fluid.configureFullyCompositionalThreePhase(d.water_component);
    if (d.aqueous_viscosity == NN_AQ_VISCOSITY_MCBRIDE_WRIGHT)
        fluid.configureMcBrideWrightAqueousViscosity(d.water_component, d.co2_component,
                                                     d.aqueous_viscosity_max_unsupported);
    else if (d.aqueous_viscosity == NN_AQ_VISCOSITY_IAPWS2008)
        fluid.configureIapws2008AqueousViscosity(d.water_component,
                                                 d.aqueous_viscosity_max_unsupported);

    using Eval = typename Indices::ValueType;
    const double swc = d.relperm[0], sorw = d.relperm[1], sgc = d.relperm[2],
                 sorg = d.relperm[3], krwMax = d.relperm[4], krgMax = d.relperm[5],
                 kroMax = d.relperm[6];
    // Corey forms of case/h2o_co2_nc10_2d_benchmark/benchmark_common.hpp:97-126.
    fluid.waterRelativePermeability = [=](Eval sw) {
        const double value = MPMC::scalarValue(sw);
        if (value <= swc) return Eval(0.0);
        if (value >= 1.0 - sorw) return Eval(krwMax);
        const Eval se = (sw - swc) / (1.0 - swc - sorw);
        return Eval(krwMax) * se * se;
    };
    fluid.gasRelativePermeability = [=](Eval sg) {
        const double value = MPMC::scalarValue(sg);
        if (value <= sgc) return Eval(0.0);
        if (value >= 1.0 - sorg - swc) return Eval(krgMax);
        const Eval se = (sg - sgc) / (1.0 - sorg - swc - sgc);
        return Eval(krgMax) * se * se;
    };
    fluid.threePhaseOilRelativePermeability = [=](Eval, Eval so, Eval) {
        const double value = MPMC::scalarValue(so);
        if (value <= sorg) return Eval(0.0);
        if (value >= 1.0 - swc) return Eval(kroMax);
        const Eval se = (so - sorg) / (1.0 - swc - sorg);
        return Eval(kroMax) * se * se;
    };
    return fluid;
}

MPMC::ThreePhaseFlashOptions makeOptions(const nn_flash_options &o, int water)
{
    MPMC::ThreePhaseFlashOptions r;
    r.waterComponent = water;
    r.maximumIterations = o.maximum_iterations;
    r.maximumRachfordRiceIterations = o.maximum_rachford_rice_iterations;
    r.maximumStabilityIterations = o.maximum_stability_iterations;
    r.stabilitySsiIterationsBeforeNewton = o.stability_ssi_before_newton;
    r.maximumStabilityNewtonIterations = o.maximum_stability_newton;
    r.twoPhaseSsiIterationsBeforeNewton = o.two_phase_ssi_before_newton;
    r.maximumTwoPhaseNewtonIterations = o.maximum_two_phase_newton;
    r.threePhaseSsiIterationsBeforeNewton = o.three_phase_ssi_before_newton;
    r.maximumThreePhaseNewtonIterations = o.maximum_three_phase_newton;
    r.fugacityTolerance = o.fugacity_tolerance;
    r.rachfordRiceTolerance = o.rachford_rice_tolerance;
    r.stabilityTolerance = o.stability_tolerance;
    r.phaseFractionTolerance = o.phase_fraction_tolerance;
    r.coincidentPhaseCompositionTolerance = o.coincident_composition_tolerance;
    r.compositionFloor = o.composition_floor;
    r.logKStepLimit = o.log_k_step_limit;
    r.numericalJacobianStep = o.numerical_jacobian_step;
    return r;
}


} // namespace oracle
