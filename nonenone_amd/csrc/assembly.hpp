// assembly.hpp — per-cell AD properties, local residual block and TPFA face
// fluxes of the Natural formulation, written straight into BSR blocks.
//
// Follows:
//   CellStateCodec::decode            state/state_codec.hpp:99-176
//   evaluateFullyCompositional_       state/cell_property_evaluator.hpp:236-393
//   FluidSystem::eosFlowProperties    fluid_system.hpp:97-133
//   CompositionalPropertyModel        properties/compositional_properties.hpp:37-150
//   massFractions                     thermo/cubic_eos.hpp:636-653
//   Corey relative permeability       case/h2o_co2_nc10_2d_benchmark/benchmark_common.hpp:97-126
//   accumulation                      physics/accumulation.hpp:28-53
//   assembleCellLocalResidual         assembly/cell_residual.hpp:71-425, local_equations.hpp:88-96
//   computeFaceMassFluxValue          physics/face_flux.hpp:330-430 (residual, records upwind)
//   computeFaceMassFlux               physics/face_flux.hpp:174-321 (AD, exterior frozen,
//                                     fixed upwind decision)
//   residual / Jacobian drivers       petsc/detail/natural_petsc_runtime_residual.inc:59-120,
//                                     natural_petsc_runtime_jacobian.inc:8-299 (scaling off)
// Scalar `S` is Dual<NP> for the Jacobian and Dual<0> for value-only passes, so
// both round exactly like the reference's AD evaluation.
#pragma once

#include "common.hpp"
#include "dual.hpp"
#include "aqueous.hpp"
#include "cpa.hpp"
#include "eos.hpp"
#include "phase_update.hpp"

namespace nn
{

constexpr double kTraceComposition = 1.0e-14;        // numerics.hpp:77
constexpr double kDependentCancellation = 1.0e-13;   // numerics.hpp:90

template <int NC, class S>
struct CellStateT
{
    S p;
    S sat[3];
    S x[3][NC];
    S well;
    uint8_t presence;
};

template <int NC, class S>
struct CellPropsT
{
    S rho[3], mu[3], mob[3], sat[3];
    S w[3][NC];
    S fug[3][NC];
    S porosity;
};

// state_codec.hpp:99-176.  The derivative components of S seed primary
// variables col0 .. col0 + S::size - 1 (a derivative chunk; col0 = 0 and
// S::size = N_p seed them all).
template <int NC, bool Well, class S>
__host__ __device__ void decodeState(const double *prim, uint8_t presence, CellStateT<NC, S> &st,
                                     int col0 = 0)
{
    using L = Layout<NC, Well>;
    auto var = [&](int idx) {
        S r(prim[idx]);
        if constexpr (S::size > 0)
            if (idx >= col0 && idx < col0 + S::size)
                r.d[idx - col0] = 1.0;
        return r;
    };
    st.presence = presence;
    st.p = var(L::pressure);
    for (int q = 0; q < 3; ++q) st.sat[q] = var(L::satBegin + q);
    if constexpr (Well)
        st.well = var(L::well);
    else
        st.well = S(0.0);
    for (int q = 0; q < 3; ++q)
    {
        S dep = 1.0;
        for (int i = 0; i < NC - 1; ++i)
        {
            st.x[q][i] = var(L::compBegin(q) + i);
            dep -= st.x[q][i];
        }
        const double dv = sval(dep);
        if (::fabs(dv) <= kDependentCancellation)
            dep += -dv;
        st.x[q][NC - 1] = dep;
    }
}

// Corey curves (benchmark_common.hpp:97-126); rp = swc, sorw, sgc, sorg, krw, krg, kro.
template <class S>
__host__ __device__ S coreyWater(const double *rp, const S &sw)
{
    const double v = sval(sw);
    if (v <= rp[0]) return S(0.0);
    if (v >= 1.0 - rp[1]) return S(rp[4]);
    const S se = (sw - rp[0]) / (1.0 - rp[0] - rp[1]);
    return S(rp[4]) * se * se;
}
template <class S>
__host__ __device__ S coreyGas(const double *rp, const S &sg)
{
    const double v = sval(sg);
    if (v <= rp[2]) return S(0.0);
    if (v >= 1.0 - rp[3] - rp[0]) return S(rp[5]);
    const S se = (sg - rp[2]) / (1.0 - rp[3] - rp[0] - rp[2]);
    return S(rp[5]) * se * se;
}
template <class S>
__host__ __device__ S coreyOil(const double *rp, const S &so)
{
    const double v = sval(so);
    if (v <= rp[3]) return S(0.0);
    if (v >= 1.0 - rp[0]) return S(rp[6]);
    const S se = (so - rp[3]) / (1.0 - rp[0] - rp[3]);
    return S(rp[6]) * se * se;
}

// LBC viscosity from molar density (compositional_properties.hpp:99-152).
template <int NC, class S>
__host__ __device__ S lbcViscosity(const FluidParams &f, const S *x, const S &rhoMolar, const S &T)
{
    S num = 0.0, den = 0.0;
    for (int i = 0; i < NC; ++i)
    {
        const double mwSqrt = f.lbcMwSqrt[i];
        const S tr = T / f.tc[i];
        const double eps = f.lbcEps[i];
        const S cv = sval(tr) > 1.5 ? 17.78e-5 * pow(4.58 * tr - 1.67, 0.625) / eps
                                    : 34.0e-5 * pow(tr, 0.94) / eps;
        num += x[i] * cv * mwSqrt;
        den += x[i] * mwSqrt;
    }
    const S atmospheric = num / den;
    S pcMix = 0.0, tcMix = 0.0, vcMix = 0.0, mwMix = 0.0;
    for (int i = 0; i < NC; ++i)
    {
        pcMix += f.pc[i] * x[i];
        tcMix += f.tc[i] * x[i];
        vcMix += f.vc[i] * x[i];
        mwMix += f.mw[i] * x[i];
    }
    const S epsMix = 5.4402 * pow(tcMix / kKelvinPerRankine, 1.0 / 6.0) /
                     (sqrt(1000.0 * mwMix) * pow(pcMix / kPsi, 2.0 / 3.0) * kCentipoise);
    const S rd = vcMix * rhoMolar;
    const S rd2 = rd * rd;
    const S rd3 = rd2 * rd;
    const S rd4 = rd2 * rd2;
    const S poly = 0.1023 + 0.023364 * rd + 0.058533 * rd2 + -0.040758 * rd3 + 0.0093324 * rd4;
    const S poly2 = poly * poly;
    const S poly4 = poly2 * poly2;
    return atmospheric + (poly4 + -1.0e-4) / epsMix;
}

// cell_property_evaluator.hpp:236-393 (PR / SW / PR-CPA, with the Water-role
// closures when configured; no Land, no pore-volume multiplier).  roots (PR-CPA, may be null): 3 x 3 doubles, per
// phase the scalar density root and its two correction slopes of an earlier
// evaluation of the same cell values (root > 0), or slots for them (root <= 0).
template <int NC, class S>
__host__ __device__ void evaluateProperties(const FluidParams &f, const CellStateT<NC, S> &st,
                                            double porosity, double Tcell,
                                            CellPropsT<NC, S> &pr, Err &err,
                                            double *roots = nullptr)
{
    pr.porosity = S(porosity);
    S comp[3][NC];
    for (int q = 0; q < 3; ++q)
        for (int c = 0; c < NC; ++c) comp[q][c] = st.x[q][c];
    if (!(st.presence & kWaterBit))
        for (int c = 0; c < NC; ++c) comp[kWater][c] = S(c == f.water ? 1.0 : 0.0);
    const S T(Tcell);
    const bool cpa = f.model == NN_MODEL_CPA;
    MixParams<NC, S> nonAq;
    CpaTemp<NC> cpaT;
    if (cpa)
        cpaTemperature<NC>(f, Tcell, cpaT);  // CPA: composition-dependent roots (:338-357)
    else
        mixingParameters<NC, S>(f, st.p, T, kOil, nonAq);
    const bool sw = f.model == NN_MODEL_SW;
    for (int q = 0; q < 3; ++q)
    {
        pr.sat[q] = st.sat[q];
        S Z;
        S lnPhi[NC];
        if (cpa)
        {
            cpaPhase<NC, S>(f, cpaT, st.p, T, comp[q], q != kGas, Z, lnPhi, err,
                            roots ? roots + 3 * q : nullptr);
            if (err.bad())
                return;
        }
        else
        {
            MixParams<NC, S> aq;
            const MixParams<NC, S> *mp = &nonAq;
            if (q == kWater && sw)
            {
                mixingParameters<NC, S>(f, st.p, T, kWater, aq);
                mp = &aq;
            }
            PhaseMix<NC, S> mix;
            phaseMixing<NC, S>(comp[q], *mp, mix);
            Z = selectRoot(f, mix.A, mix.B, q != kGas, err);
            if (err.bad())
                return;
            lnPhiAtRoot<NC, S>(f, *mp, mix, Z, lnPhi, err);
            if (err.bad())
                return;
        }
        for (int c = 0; c < NC; ++c)
            pr.fug[q][c] = exp(lnPhi[c]) * st.p * comp[q][c];
        // eosFlowProperties: mass fractions, density, LBC viscosity
        S totalMass = 0.0;
        S mass[NC];
        for (int c = 0; c < NC; ++c)
        {
            mass[c] = comp[q][c] * f.mw[c];
            totalMass += mass[c];
        }
        if (!(sval(totalMass) > 0.0))
        {
            err.raise(kErrThermo);
            return;
        }
        for (int c = 0; c < NC; ++c) pr.w[q][c] = mass[c] / totalMass;
        // fluid_system.hpp:97-133: the phase molar density (Water role: the
        // IAPWS-Garcia volume when configured) gives the mass density; the
        // viscosity is LBC at the cubic molar density, or the Water role's closure
        const S cubicRhoMolar = st.p / (kGasConstant * T * Z);
        const S rhoMolar = f.aqVolume ? phaseMolarDensity<NC, S>(f, st.p, T, comp[q], Z,
                                                                 q == kWater, err)
                                      : cubicRhoMolar;
        if (err.bad())
            return;
        S mixMolarMass = 0.0;
        for (int c = 0; c < NC; ++c) mixMolarMass += comp[q][c] * f.mw[c];
        pr.rho[q] = rhoMolar * mixMolarMass;
        if (q == kWater && f.aqViscosity != NN_AQ_VISCOSITY_NONE)
            pr.mu[q] = aqueousViscosity<NC, S>(f, st.p, T, comp[q], err);
        else
            pr.mu[q] = lbcViscosity<NC, S>(f, comp[q], cubicRhoMolar, T);
        if (err.bad())
            return;
    }
    S sgRelperm = st.sat[kGas];
    if (!(st.presence & kGasBit))
        clearDeriv(sgRelperm);
    const S freeGas = sgRelperm - S(0.0);
    const S kr[3] = {coreyOil(f.relperm, st.sat[kOil]), coreyGas(f.relperm, freeGas),
                     coreyWater(f.relperm, st.sat[kWater])};
    for (int q = 0; q < 3; ++q)
        pr.mob[q] = (st.presence & (1u << q)) ? kr[q] / pr.mu[q] : S(0.0);
}

// Property record a neighbour reads (values of the AD evaluation).
template <int NC>
struct FaceRecord
{
    static constexpr int kFields = 10 + 3 * NC;  // p, rho[3], sat[3], mob[3], w[3][NC]
};

// physics/accumulation.hpp:28-53: M_c = phi * sum_p rho_p S_p w_pc
template <int NC, class S>
__host__ __device__ S componentAccumulation(const CellPropsT<NC, S> &pr, int c)
{
    S mass = 0.0;
    for (int q = 0; q < 3; ++q) mass += pr.rho[q] * pr.sat[q] * pr.w[q][c];
    return pr.porosity * mass;
}

// cell_residual.hpp:71-425 local/base block (no adsorption, no legacy water).
template <int NC, bool Well, class S>
__host__ __device__ void localResidual(const CellStateT<NC, S> &st, const CellPropsT<NC, S> &pr,
                                       const double *prevAccum, double volume, double dt,
                                       double fugScale, S *res)
{
    const double storage = volume / dt;
    for (int c = 0; c < NC; ++c)
        res[c] = (componentAccumulation<NC, S>(pr, c) - prevAccum[c]) * storage;
    const bool oil = st.presence & kOilBit;
    const bool gas = st.presence & kGasBit;
    const bool water = st.presence & kWaterBit;
    const int fugBegin = NC, waterFugBegin = 2 * NC;

    auto fugacityBlock = [&](int eqBegin, int ref, int other, int alt, bool altActive) {
        for (int c = 0; c < NC; ++c)
        {
            const int eq = eqBegin + c;
            const bool refTrace = sval(st.x[ref][c]) <= kTraceComposition;
            const bool otherTrace = sval(st.x[other][c]) <= kTraceComposition;
            if (otherTrace)
            {
                res[eq] = st.x[other][c];
                continue;
            }
            if (refTrace)
            {
                if (altActive && alt >= 0 && sval(st.x[alt][c]) > kTraceComposition)
                {
                    res[eq] = (pr.fug[alt][c] - pr.fug[other][c]) / kBar / fugScale;
                    continue;
                }
                res[eq] = st.x[ref][c];
                continue;
            }
            res[eq] = (pr.fug[ref][c] - pr.fug[other][c]) / kBar / fugScale;
        }
    };
    auto inactiveBlock = [&](int eqBegin, int phase) {
        for (int i = 0; i < NC - 1; ++i)
            res[eqBegin + i] = st.x[phase][i] - sval(st.x[phase][i]);
        res[eqBegin + NC - 1] = st.sat[phase];
    };
    if (oil)
    {
        if (gas)
            fugacityBlock(fugBegin, kOil, kGas, -1, false);
        else
            inactiveBlock(fugBegin, kGas);
        if (water)
            fugacityBlock(waterFugBegin, kOil, kWater, kGas, gas);
        else
            inactiveBlock(waterFugBegin, kWater);
    }
    else
    {
        inactiveBlock(fugBegin, kOil);
        if (gas && water)
            fugacityBlock(waterFugBegin, kGas, kWater, -1, false);
        else if (gas)
            inactiveBlock(waterFugBegin, kWater);
        else
            inactiveBlock(waterFugBegin, kGas);
    }
    res[3 * NC] = st.sat[0] + st.sat[1] + st.sat[2] - 1.0;
    if constexpr (Well)
        res[3 * NC + 1] = st.well - sval(st.well);
}

// face_flux.hpp:65-87
__host__ __device__ inline bool interiorUpwind(double dPhi, double vIn, double vEx, int64_t idIn,
                                               int64_t idEx)
{
    if (dPhi > 0.0) return false;
    if (dPhi < 0.0) return true;
    if (vIn > vEx) return true;
    if (vEx > vIn) return false;
    return idIn < idEx;
}

} // namespace nn
