// ORACLE — test infrastructure only.  Never linked into, imported by or
// called from the product library; only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load the shared object built from this file.
//
// ref_harness.cpp: a C-ABI wrapper around the reference implementation
// itself.  It contains no reference code; it #includes the reference's
// header-only C++17 sources where they lie (include paths set by
// oracle/Makefile) and restates only the PETSc-bound driver loops that cannot
// be compiled without PETSc:
//   * the residual / Jacobian per-cell loops of
//     models/include/natural/petsc/detail/natural_petsc_runtime_residual.inc:44-133
//     models/include/natural/petsc/detail/natural_petsc_runtime_jacobian.inc:8-299
//   * the cell-cache build of natural_petsc_runtime_cell_cache.inc:43-126
//   * the previous-accumulation refresh of natural_petsc_runtime_cache_history.inc:177-260
// Everything per cell (EOS, flash, stability, phase update, AD properties,
// local residual, face flux) is the reference's own code.
//
// Output goes to oracle/_ref/libnn_ref.so (never committed).

#include <indices/indices.hpp>
#include <output/core/format.hpp>
#include <natural/assembly/cell_residual.hpp>
#include <natural/fluid_system.hpp>
#include <natural/kernel/cell_kernel.hpp>
#include <natural/physics/accumulation.hpp>
#include <natural/physics/face_flux.hpp>
#include <natural/properties/aqueous_viscosity.hpp>
#include <natural/petsc/phase_state_codec.hpp>
#include <natural/physics/well_source.hpp>
#include <natural/petsc/property_conversion.hpp>
#include <natural/state/newton_limiter.hpp>
#include <natural/state/state_codec.hpp>
#include <natural/state/three_phase_equilibrium.hpp>
#include <natural/thermo/three_phase_flash.hpp>

#include "../include/nonenone_amd.h"
#include "ref_fluid.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <thread>
#include <sstream>
#include <vector>

namespace
{

using oracle::AdIdx;
using oracle::Cfg;
using oracle::ScIdx;
using oracle::makeFluid;
using oracle::makeOptions;

struct Handle
{
    nn_fluid_desc desc;
    nn_flash_options opts;
};

template <class F>
void parallelFor(int64_t n, int threads, F &&body)
{
    if (threads <= 1 || n < 2)
    {
        body(0, int64_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    const int64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t)
    {
        const int64_t b = t * chunk;
        const int64_t e = std::min(n, b + chunk);
        if (b >= e)
            break;
        pool.emplace_back([&, t, b, e] { body(t, b, e); });
    }
    for (auto &th : pool)
        th.join();
}

// ---------------------------------------------------------------- EOS
template <int NC>
int eosPhase(const Handle &h, int64_t n, const double *P, const double *T,
             const double *x, const int32_t *role, double *Z, double *lnphi,
             double *fug, int32_t *status)
{
    using I = ScIdx<NC>;
    auto fluid = makeFluid<I>(h.desc, 300.0);
    for (int64_t s = 0; s < n; ++s)
    {
        std::array<double, NC> comp{};
        for (int c = 0; c < NC; ++c)
            comp[c] = x[s * NC + c];
        try
        {
            const auto r = fluid.eos.phaseResult(
                P[s], T[s], comp, static_cast<MPMC::CompositionalPhase>(role[s]));
            Z[s] = r.compressibility;
            for (int c = 0; c < NC; ++c)
            {
                lnphi[s * NC + c] = std::log(r.fugacityCoefficient[c]);
                fug[s * NC + c] = r.fugacity[c];
            }
            status[s] = NN_STATUS_OK;
        }
        catch (const std::exception &)
        {
            status[s] = NN_STATUS_THERMO_ERROR;
        }
    }
    return NN_OK;
}

template <class Result, int NC>
void storeFlash(const Result &r, int64_t s, uint8_t *presence, uint8_t *converged,
                int32_t *iterations, double *beta, double *sat, double *x,
                double *Z, double *rho, double *kgo, double *kwo)
{
    presence[s] = r.presence.bits();
    converged[s] = r.converged ? 1 : 0;
    iterations[s] = r.iterations;
    for (int p = 0; p < 3; ++p)
    {
        beta[s * 3 + p] = r.phaseMoleFraction[p];
        sat[s * 3 + p] = r.saturation[p];
        Z[s * 3 + p] = r.compressibility[p];
        rho[s * 3 + p] = r.molarDensity[p];
        for (int c = 0; c < NC; ++c)
            x[(s * 3 + p) * NC + c] = r.composition[p][c];
    }
    for (int c = 0; c < NC; ++c)
    {
        kgo[s * NC + c] = r.vaporOilK[c];
        kwo[s * NC + c] = r.waterOilK[c];
    }
}

template <int NC>
int flashBatch(const Handle &h, int64_t n, const double *P, const double *T,
               const double *z, const uint8_t *allowed, const double *seed,
               uint8_t *presence, uint8_t *converged, int32_t *status,
               int32_t *iterations, double *beta, double *sat, double *x,
               double *Z, double *rho, double *kgo, double *kwo, int threads)
{
    using I = ScIdx<NC>;
    parallelFor(n, threads, [&](int, int64_t b, int64_t e) {
        auto fluid = makeFluid<I>(h.desc, 300.0);
        MPMC::CubicThreePhaseFlash<I> flash(
            fluid.eos, makeOptions(h.opts, h.desc.water_component));
        for (int64_t s = b; s < e; ++s)
        {
            std::array<double, NC> comp{};
            for (int c = 0; c < NC; ++c)
                comp[c] = z[s * NC + c];
            try
            {
                typename MPMC::CubicThreePhaseFlash<I>::Result r;
                if (allowed == nullptr)
                    r = flash.flash(P[s], T[s], comp);
                else if (seed == nullptr)
                    r = flash.flashRestricted(P[s], T[s], comp,
                                              MPMC::PhasePresence(allowed[s]));
                else
                {
                    std::array<std::array<double, NC>, 3> sd{};
                    for (int p = 0; p < 3; ++p)
                        for (int c = 0; c < NC; ++c)
                            sd[p][c] = seed[(s * 3 + p) * NC + c];
                    r = flash.flashRestricted(P[s], T[s], comp,
                                              MPMC::PhasePresence(allowed[s]), sd);
                }
                storeFlash<decltype(r), NC>(r, s, presence, converged, iterations,
                                            beta, sat, x, Z, rho, kgo, kwo);
                status[s] = r.converged ? NN_STATUS_OK : NN_STATUS_NOT_CONVERGED;
            }
            catch (const std::invalid_argument &)
            {
                status[s] = NN_STATUS_INVALID_INPUT;
                converged[s] = 0;
            }
            catch (const std::exception &)
            {
                status[s] = NN_STATUS_THERMO_ERROR;
                converged[s] = 0;
            }
        }
    });
    return NN_OK;
}

template <int NC>
int stabilityBatch(const Handle &h, int64_t n, const double *P, const double *T,
                   const double *z, const uint8_t *active, const double *xin,
                   uint8_t *stable, uint8_t *valid, uint8_t *mask, double *trial,
                   double *incipient, int32_t *status)
{
    using I = ScIdx<NC>;
    auto fluid = makeFluid<I>(h.desc, 300.0);
    MPMC::CubicThreePhaseFlash<I> flash(
        fluid.eos, makeOptions(h.opts, h.desc.water_component));
    for (int64_t s = 0; s < n; ++s)
    {
        std::array<double, NC> comp{};
        std::array<std::array<double, NC>, 3> xs{};
        for (int c = 0; c < NC; ++c)
            comp[c] = z[s * NC + c];
        for (int p = 0; p < 3; ++p)
            for (int c = 0; c < NC; ++c)
                xs[p][c] = xin[(s * 3 + p) * NC + c];
        try
        {
            const auto r = flash.stabilityTest(
                P[s], T[s], comp, MPMC::PhasePresence(active[s]), xs);
            stable[s] = r.stable;
            valid[s] = r.valid;
            uint8_t m = 0;
            for (int p = 0; p < 3; ++p)
            {
                if (r.missingPhaseUnstable[p])
                    m |= uint8_t(1u << p);
                trial[s * 3 + p] = r.trialSum[p];
                for (int c = 0; c < NC; ++c)
                    incipient[(s * 3 + p) * NC + c] = r.incipientComposition[p][c];
            }
            mask[s] = m;
            status[s] = NN_STATUS_OK;
        }
        catch (const std::invalid_argument &)
        {
            status[s] = NN_STATUS_INVALID_INPUT;
        }
        catch (const std::exception &)
        {
            status[s] = NN_STATUS_THERMO_ERROR;
        }
    }
    return NN_OK;
}

// ---------------------------------------------------------- phase update
template <int NC, bool Well>
int phaseUpdate(const Handle &h, int64_t n, const double *T, double *primary,
                double *phaseState, int32_t *status, uint8_t *flags, int threads)
{
    using I = AdIdx<NC, Well>;
    constexpr int NP = I::numPrimaryVariables;
    constexpr int NPS = I::numPhaseStateVariables;
    parallelFor(n, threads, [&](int, int64_t b, int64_t e) {
        auto fluid = makeFluid<I>(h.desc, 300.0);
        MPMC::FullyCompositionalThreePhaseEquilibrium<I> eq(fluid);
        for (int64_t s = b; s < e; ++s)
        {
            fluid.temperature = T[s];
            std::array<double, NP> prim{};
            std::array<double, NPS> ps{};
            for (int k = 0; k < NP; ++k)
                prim[k] = primary[s * NP + k];
            for (int k = 0; k < NPS; ++k)
                ps[k] = phaseState[s * NPS + k];
            try
            {
                auto data = MPMC::PetscPhaseStateCodec<I>::decode(ps);
                // natural_petsc_runtime_state.inc:215-221 (useVariableBounds = false)
                eq.sanitizePrimaryBeforeFlash(prim, false);
                const auto r = eq.updatePhaseState(prim, data);
                const auto enc = MPMC::PetscPhaseStateCodec<I>::encode(data);
                for (int k = 0; k < NP; ++k)
                    primary[s * NP + k] = prim[k];
                for (int k = 0; k < NPS; ++k)
                    phaseState[s * NPS + k] = enc[k];
                status[s] = r.status == MPMC::PhaseUpdateStatus::RecoverableThermodynamicFailure
                    ? NN_STATUS_RECOVERABLE_FAILURE
                    : NN_STATUS_OK;
                if (flags)
                {
                    uint8_t f = 0;
                    if (r.phaseRemoved) f |= 1;
                    if (r.stabilityInvalid) f |= 2;
                    if (r.missingPhaseUnstable) f |= 4;
                    if (r.restrictedFlashFailed) f |= 8;
                    if (r.unrestrictedFlashFailed) f |= 16;
                    if (r.status == MPMC::PhaseUpdateStatus::StableReducedSet) f |= 32;
                    flags[s] = f;
                }
            }
            catch (const std::invalid_argument &)
            {
                status[s] = NN_STATUS_INVALID_INPUT;
            }
            catch (const std::exception &)
            {
                status[s] = NN_STATUS_THERMO_ERROR;
            }
        }
    });
    return NN_OK;
}

// ------------------------------------------------------------ assembly
struct Topology
{
    int64_t ncell;
    const int64_t *connOffsets;
    const int64_t *connNeighbor;
    const double *connTrans;
    const double *connGdz;
    const int64_t *rowOffsets;
    const int64_t *colBlocks;
    const double *volume;
    const double *porosity;
    const double *T;
};

// Wells of an assembly (nn_well), host arrays; nwells = 0: none.
struct WellSet
{
    int32_t nwells = 0;
    const nn_well *wells = nullptr;
    const int64_t *perfCells = nullptr;
    const double *perfIndex = nullptr;
    const double *surfaceDensity = nullptr;
};

int64_t findSlot(const Topology &t, int64_t row, int64_t col)
{
    const int64_t *b = t.colBlocks + t.rowOffsets[row];
    const int64_t *e = t.colBlocks + t.rowOffsets[row + 1];
    const int64_t *it = std::lower_bound(b, e, col);
    if (it == e || *it != col)
        throw std::logic_error("BSR pattern lacks a required block");
    return it - t.colBlocks;
}

// Equation and variable scale factors of the Natural runtime
// (natural_petsc_runtime_scaling.inc:8-89), restated because they are private
// members of the PETSc-bound runtime.  `params` = {pressureScale,
// compositionScale, saturationScale, massResidualScale, fugacityResidualScale,
// closureResidualScale}; nullptr = scaling disabled (every factor 1).  The
// well-control row's factor depends on the cell's representative well
// (rate-controlled: 1 / max(|target|, floor), otherwise 1 / pressureScale):
// `wellRow` gives it per cell (nullptr = 1 / pressureScale everywhere).
template <class I>
struct ScalingRule
{
    const double *params = nullptr;
    const double *wellRow = nullptr;

    double variable(int v) const
    {
        if (!params)
            return 1.0;
        if (v == I::Primary::pressure || (I::hasWellUnknown && v == I::Primary::wellPressure))
            return params[0];
        if (v == I::Primary::liquidSaturation || v == I::Primary::vaporSaturation ||
            (I::hasWater && v == I::Primary::waterSaturation))
            return params[2];
        return params[1];
    }

    double equation(int e, int64_t cell) const
    {
        if (!params)
            return 1.0;
        if (e >= I::Equation::massConservationBegin && e < I::Equation::afterMassConservation)
            return 1.0 / params[3];
        if (e >= I::Equation::fugacityBegin && e < I::Equation::fugacityBegin + I::numComponents)
            return 1.0 / params[4];
        if constexpr (I::fullyCompositionalThreePhase)
            if (e >= I::Equation::waterFugacityBegin &&
                e < I::Equation::waterFugacityBegin + I::numComponents)
                return 1.0 / params[4];
        if (e == I::Equation::volumeClosure)
            return 1.0 / params[5];
        if constexpr (I::hasWellUnknown)
            if (e == I::Equation::wellControl)
                return wellRow ? wellRow[cell] : 1.0 / params[0];
        return 1.0;
    }
};

template <int NC, bool Well>
struct AssemblyRunner
{
    using I = AdIdx<NC, Well>;
    using Kernel = MPMC::NaturalCellKernel<I>;
    using State = typename Kernel::State;
    using Props = typename Kernel::Properties;
    using Scalar = typename I::ValueType;
    static constexpr int NP = I::numPrimaryVariables;
    static constexpr int NPS = I::numPhaseStateVariables;

    // One fluid + kernel per thread: the reference objects hold mutable state.
    struct Worker
    {
        MPMC::FluidSystem<I> fluid;
        Kernel kernel;
        explicit Worker(const nn_fluid_desc &d) : fluid(makeFluid<I>(d, 300.0)), kernel(fluid) {}
    };

    const Handle &h;
    const Topology &t;
    std::vector<State> states;
    std::vector<Props> props;

    AssemblyRunner(const Handle &hh, const Topology &tt) : h(hh), t(tt)
    {
        states.resize(t.ncell);
        props.resize(t.ncell);
    }

    // natural_petsc_runtime_cell_cache.inc:58-116 for cells [b, e).
    void buildCache(Worker &w, const double *primary, const double *phaseState, int64_t b,
                    int64_t e)
    {
        for (int64_t c = b; c < e; ++c)
        {
            w.fluid.temperature = t.T[c];
            std::array<double, NP> prim{};
            std::array<double, NPS> ps{};
            for (int k = 0; k < NP; ++k)
                prim[k] = primary[c * NP + k];
            for (int k = 0; k < NPS; ++k)
                ps[k] = phaseState[c * NPS + k];
            const auto data = MPMC::PetscPhaseStateCodec<I>::decode(ps);
            states[c] = w.kernel.decodeState(prim, data);
            std::array<Scalar, NC> overall{};
            for (int k = 0; k < NC; ++k)
                overall[k] = Scalar(data.overallComposition[k]);
            props[c] = w.kernel.evaluateProperties(
                states[c], Scalar(t.porosity[c]), 0.0, 2.0, &overall);
        }
    }

    void accumulation(double *accum, int64_t b, int64_t e)
    {
        for (int64_t c = b; c < e; ++c)
        {
            const auto scalar = MPMC::scalarizeCellProperties<I>(props[c]);
            const auto a = MPMC::computeFluidAccumulation<I>(scalar);
            for (int k = 0; k < NC; ++k)
                accum[c * NC + k] = a.componentMass[k];
        }
    }

    // residual.inc:59-120 and jacobian.inc:20-237 for cells [b, e); each cell
    // writes its residual row, its diagonal block and its reciprocal blocks.
    // natural_petsc_runtime_wells.inc:54-91: one perforation's source with the
    // reference's own computePerforationWellSourceInto, bhp the representative
    // cell's well-pressure unknown (broadcastWellBottomHolePressures_, :8-49).
    MPMC::PerforationWellResult<I, Scalar> perforation(const WellSet &ws, int32_t well,
                                                       int64_t k, const double *primary) const
    {
        const nn_well &wl = ws.wells[well];
        const int64_t c = ws.perfCells[k];
        const Scalar bhp = MPMC::PrimaryVariables<I>::make(
            primary[wl.bhp_cell * NP + I::Primary::wellPressure], I::Primary::wellPressure);
        std::array<Scalar, 3> phasePressure{states[c].pressure, states[c].pressure,
                                            states[c].pressure};
        std::array<double, 3> surface{ws.surfaceDensity[0], ws.surfaceDensity[1],
                                      ws.surfaceDensity[2]};
        std::array<double, 3> injPhase{wl.injection_phase_fraction[0],
                                       wl.injection_phase_fraction[1],
                                       wl.injection_phase_fraction[2]};
        std::array<double, NC> injMass{};
        for (int i = 0; i < NC; ++i)
            injMass[i] = wl.injection_component_mass_fraction[i];
        MPMC::PerforationWellResult<I, Scalar> result;
        MPMC::PerforationWellWorkspace<I, Scalar> work;
        MPMC::computePerforationWellSourceInto<I, Scalar>(
            result, work, wl.type == NN_WELL_INJECTOR ? MPMC::WellType::Injector
                                                      : MPMC::WellType::Producer,
            ws.perfIndex[k], bhp, phasePressure, props[c].density, props[c].mobility, surface,
            injPhase, injMass, props[c].massFraction, props[c].aqueousCO2MassFraction, -1);
        return result;
    }

    // evaluateWellFunctionContributions_ (:93-165): per-cell sums in well, then
    // perforation order
    std::vector<std::array<Scalar, NC>> cellSources(const WellSet &ws, const double *primary) const
    {
        std::vector<std::array<Scalar, NC>> src;
        if (ws.nwells == 0)
            return src;
        src.assign(static_cast<std::size_t>(t.ncell), std::array<Scalar, NC>{});
        for (int32_t wl = 0; wl < ws.nwells; ++wl)
            for (int64_t k = ws.wells[wl].perforation_begin; k < ws.wells[wl].perforation_end; ++k)
            {
                const auto r = perforation(ws, wl, k, primary);
                for (int i = 0; i < NC; ++i)
                    src[static_cast<std::size_t>(ws.perfCells[k])][i] += r.componentMassSource[i];
            }
        return src;
    }

    // insertWellJacobian_ (jacobian.inc:301-457) for active BHP-controlled wells:
    // each perforation adds its rows, the well-pressure column moved to the
    // representative cell's block; and overwriteWellControlResiduals_
    // (wells.inc:167-243): R_w = p_bhp - target on the representative cell's row.
    void addWells(const WellSet &ws, const ScalingRule<I> &sc, const double *primary,
                  double *residual, double *bsr) const
    {
        constexpr int N = NP;
        for (int32_t wl = 0; wl < ws.nwells; ++wl)
        {
            const nn_well &well = ws.wells[wl];
            if (residual)
            {
                const int row = I::Equation::wellControl;
                const double bhp = primary[well.bhp_cell * N + I::Primary::wellPressure];
                residual[well.bhp_cell * N + row] =
                    (bhp - well.target) * sc.equation(row, well.bhp_cell);
            }
            if (!bsr)
                continue;
            for (int64_t k = well.perforation_begin; k < well.perforation_end; ++k)
            {
                const int64_t c = ws.perfCells[k];
                const auto r = perforation(ws, wl, k, primary);
                for (int q = 0; q < NC; ++q)
                {
                    const int row = I::Equation::massConservation[q];
                    const Scalar eq = -r.componentMassSource[q];
                    const double rowScale = sc.equation(row, c);
                    for (int col = 0; col < N; ++col)
                    {
                        const int64_t colBlock = col == I::Primary::wellPressure ? well.bhp_cell : c;
                        bsr[findSlot(t, c, colBlock) * N * N + row * N + col] +=
                            eq.derivative(col) * rowScale * sc.variable(col);
                    }
                }
            }
        }
    }

    void evaluate(Worker &w, double dt, double fugScale, const ScalingRule<I> &sc,
                  const double *accumPrev, double *residual, double *bsr,
                  uint8_t *decisionsOut, int64_t b, int64_t e,
                  const std::vector<std::array<Scalar, NC>> *wellSource = nullptr)
    {
        constexpr int N = NP;
        std::vector<MPMC::FaceFluxLinearizationDecision<I>> decisions;
        std::vector<double> selfBlock(N * N);
        for (int64_t c = b; c < e; ++c)
        {
            MPMC::CellProperties<I, double> prevProps{};
            MPMC::AccumulationResult<I, double> prevAccum{};
            for (int k = 0; k < NC; ++k)
                prevAccum.componentMass[k] = accumPrev[c * NC + k];
            const int64_t cb = t.connOffsets[c], ce = t.connOffsets[c + 1];
            decisions.assign(static_cast<std::size_t>(ce - cb), {});

            MPMC::FaceMassFlux<I, double> outward{};
            for (int64_t k = cb; k < ce; ++k)
            {
                const int64_t nb = t.connNeighbor[k];
                const auto flux = MPMC::computeFaceMassFluxValue<I, Scalar>(
                    states[c], props[c], states[nb], props[nb], t.connTrans[k],
                    t.connGdz[k], t.volume[c], t.volume[nb], c, nb, -1,
                    &decisions[static_cast<std::size_t>(k - cb)]);
                for (int q = 0; q < NC; ++q)
                    outward.component[q] += flux.component[q];
                if (decisionsOut)
                    for (int p = 0; p < 3; ++p)
                        decisionsOut[k * 3 + p] = static_cast<uint8_t>(
                            decisions[static_cast<std::size_t>(k - cb)].upwind[p]);
            }
            auto block = w.kernel.assembleLocalResidual(
                states[c], props[c], prevProps, t.volume[c], dt, fugScale, -1,
                2650.0, nullptr, &prevAccum, nullptr);
            if (residual)
            {
                auto withFlux = block;
                std::array<Scalar, NC> noWell{};
                MPMC::addCellConservationTransportAndWell<I>(
                    withFlux, outward,
                    wellSource ? (*wellSource)[static_cast<std::size_t>(c)] : noWell,
                    Scalar(0.0));
                // natural_petsc_runtime_common.inc:345-356
                for (int r = 0; r < N; ++r)
                    residual[c * N + r] = MPMC::scalarValue(withFlux.value[r]) * sc.equation(r, c);
            }
            if (!bsr)
                continue;
            // jacobian.inc:262-268 (self block), :196-213 (face terms)
            for (int r = 0; r < N; ++r)
            {
                const double rowScale = sc.equation(r, c);
                for (int q = 0; q < N; ++q)
                    selfBlock[r * N + q] = block.value[r].derivative(q) * rowScale * sc.variable(q);
            }
            for (int64_t k = cb; k < ce; ++k)
            {
                const int64_t nb = t.connNeighbor[k];
                const auto flux = MPMC::computeFaceMassFlux<I, Scalar>(
                    states[c], props[c], states[nb], props[nb], t.connTrans[k],
                    t.connGdz[k], t.volume[c], t.volume[nb], c, nb, -1,
                    &decisions[static_cast<std::size_t>(k - cb)]);
                double *recip = bsr + findSlot(t, nb, c) * N * N;
                for (int q = 0; q < NC; ++q)
                {
                    const int row = I::Equation::massConservation[q];
                    const double selfRowScale = sc.equation(row, c);
                    const double recipRowScale = sc.equation(row, nb);
                    for (int col = 0; col < N; ++col)
                    {
                        const double dv = flux.component[q].derivative(col);
                        const double varScale = sc.variable(col);
                        selfBlock[row * N + col] += dv * selfRowScale * varScale;
                        recip[row * N + col] = -dv * recipRowScale * varScale;
                    }
                }
            }
            double *self = bsr + findSlot(t, c, c) * N * N;
            for (int q = 0; q < N * N; ++q)
                self[q] += selfBlock[q];
        }
    }
};

template <int NC, bool Well>
int assemblyRun(const Handle &h, const Topology &t, double dt, double fugScale,
                const double *scaling, const double *wellRowScale,
                const double *primary, const double *phaseState,
                const double *accumPrev, double *residual, double *bsr,
                uint8_t *decisions, double *accumOut, int threads,
                const WellSet &wells = WellSet{})
{
    using Runner = AssemblyRunner<NC, Well>;
    using I = typename Runner::I;
    const ScalingRule<I> sc{scaling, wellRowScale};
    constexpr int N = Runner::NP;
    Runner runner(h, t);
    if (bsr)
        std::fill(bsr, bsr + t.rowOffsets[t.ncell] * N * N, 0.0);
    // Properties of every cell first (neighbours read them), then the cell loop.
    parallelFor(t.ncell, threads, [&](int, int64_t b, int64_t e) {
        typename Runner::Worker w(h.desc);
        runner.buildCache(w, primary, phaseState, b, e);
    });
    if (accumOut)
        runner.accumulation(accumOut, 0, t.ncell);
    if (wells.nwells > 0 && !Well)
        throw std::invalid_argument("wells need a well unknown per cell");
    const auto wellSource = runner.cellSources(wells, primary);
    if (residual || bsr)
        parallelFor(t.ncell, threads, [&](int, int64_t b, int64_t e) {
            typename Runner::Worker w(h.desc);
            runner.evaluate(w, dt, fugScale, sc, accumPrev, residual, bsr, decisions, b, e,
                            wellSource.empty() ? nullptr : &wellSource);
        });
    if constexpr (Well)
        if (wells.nwells > 0)
            runner.addWells(wells, sc, primary, residual, bsr);
    return NN_OK;
}

// NaturalPetscRuntime::limitNewtonStep (natural_petsc_runtime_state.inc:452-545)
// restated around the reference's own limiters: delta = -Y, the scaled increment
// recovered with the variable scales (`params`: the six scaling factors, null =
// scaling off), limitNaturalNewtonIncrement, limitAqueousCompositionNewtonIncrement
// when maxAqueous >= 0, then Y = -delta.
template <int NC, bool Well>
int newtonLimit(int64_t n, const double *primary, double *step, const double *params,
                double maxAqueous, int water, int co2)
{
    using I = AdIdx<NC, Well>;
    constexpr int NP = I::numPrimaryVariables;
    const ScalingRule<I> sc{params, nullptr};
    for (int64_t c = 0; c < n; ++c)
    {
        std::array<double, NP> state{}, delta{};
        for (int v = 0; v < NP; ++v)
        {
            state[v] = primary[c * NP + v];
            delta[v] = -step[c * NP + v];
            if (params)
                delta[v] *= sc.variable(v);
        }
        MPMC::limitNaturalNewtonIncrement<I>(delta, state, MPMC::PhasePresence{});
        if (maxAqueous >= 0.0)
            MPMC::limitAqueousCompositionNewtonIncrement<I>(delta, state, water, co2, maxAqueous);
        for (int v = 0; v < NP; ++v)
            step[c * NP + v] = -delta[v];
    }
    return NN_OK;
}

template <class F>
int dispatchNc(int nc, F &&f)
{
    switch (nc)
    {
    case 2: return f(std::integral_constant<int, 2>{});
    case 3: return f(std::integral_constant<int, 3>{});
    case 4: return f(std::integral_constant<int, 4>{});
    case 5: return f(std::integral_constant<int, 5>{});
    case 6: return f(std::integral_constant<int, 6>{});
    case 8: return f(std::integral_constant<int, 8>{});
    default: return NN_E_UNSUPPORTED;
    }
}

thread_local char g_error[512] = "";

} // namespace

extern "C" {

void *ref_create(const nn_fluid_desc *desc, const nn_flash_options *opts)
{
    auto *h = new Handle;
    h->desc = *desc;
    h->opts = *opts;
    return h;
}

void ref_destroy(void *h) { delete static_cast<Handle *>(h); }

const char *ref_last_error(void) { return g_error; }

// The reference's sample CSV text (output/include/output/core/format.hpp:19-42):
// the header for `count` values, then one row of `values`; returns the text
// length (written into buf when it fits, NUL-terminated).
int64_t ref_csv_samples(int64_t count, const double *values, char *buf, int64_t cap)
{
    std::ostringstream os;
    MPMC::writeIndexedCsvHeader(os, static_cast<std::size_t>(count));
    MPMC::writeCsvValues(os, std::vector<double>(values, values + count));
    const std::string text = os.str();
    if (buf && int64_t(text.size()) < cap)
        std::memcpy(buf, text.c_str(), text.size() + 1);
    return int64_t(text.size());
}

int ref_newton_limit(void *hv, int64_t n, int hasWell, const double *primary, double *step,
                     const double *scaling, double maxAqueous, int water, int co2)
{
    const Handle &h = *static_cast<Handle *>(hv);
    return dispatchNc(h.desc.nc, [&](auto nc) {
        constexpr int NC = decltype(nc)::value;
        return hasWell ? newtonLimit<NC, true>(n, primary, step, scaling, maxAqueous, water, co2)
                       : newtonLimit<NC, false>(n, primary, step, scaling, maxAqueous, water, co2);
    });
}

// The reference's Water-role closures at (P, T) for an H2O/CO2 aqueous phase of
// CO2 mole fraction xco2: IF97 density (aqueous_volume.hpp:270-295), Garcia CO2
// volume, IAPWS-2008 and McBride-Wright viscosities; NaN where the reference throws.
int ref_water_closures(int64_t n, const double *P, const double *T, const double *xco2,
                       double *density, double *co2Volume, double *muIapws, double *muMbw)
{
    const MPMC::Iapws2008IndustrialAqueousViscosity<2> iapws(0, 1.0);
    const MPMC::McBrideWright2015AqueousViscosity<2> mbw(0, 1, 1.0);
    for (int64_t i = 0; i < n; ++i)
    {
        const std::array<double, 2> x{1.0 - xco2[i], xco2[i]};
        const std::array<double, 2> pure{1.0, 0.0};
        auto guarded = [&](auto &&f) {
            try { return double(f()); }
            catch (const std::exception &) { return std::nan(""); }
        };
        density[i] = guarded([&] { return MPMC::IapwsIf97WaterDensity::density(P[i], T[i]); });
        co2Volume[i] = MPMC::Garcia2001Co2ApparentVolume::molarVolume(T[i]);
        muIapws[i] = guarded([&] { return iapws.viscosity(P[i], T[i], pure); });
        muMbw[i] = guarded([&] { return mbw.viscosity(P[i], T[i], x); });
    }
    return NN_OK;
}

int ref_eos_phase(void *hv, int64_t n, const double *P, const double *T,
                  const double *x, const int32_t *role, double *Z, double *lnphi,
                  double *fug, int32_t *status)
{
    const Handle &h = *static_cast<Handle *>(hv);
    return dispatchNc(h.desc.nc, [&](auto nc) {
        return eosPhase<decltype(nc)::value>(h, n, P, T, x, role, Z, lnphi, fug, status);
    });
}

int ref_flash(void *hv, int64_t n, const double *P, const double *T, const double *z,
              const uint8_t *allowed, const double *seed, uint8_t *presence,
              uint8_t *converged, int32_t *status, int32_t *iterations,
              double *beta, double *sat, double *x, double *Z, double *rho,
              double *kgo, double *kwo, int threads)
{
    const Handle &h = *static_cast<Handle *>(hv);
    return dispatchNc(h.desc.nc, [&](auto nc) {
        return flashBatch<decltype(nc)::value>(
            h, n, P, T, z, allowed, seed, presence, converged, status, iterations,
            beta, sat, x, Z, rho, kgo, kwo, threads);
    });
}

int ref_stability(void *hv, int64_t n, const double *P, const double *T,
                  const double *z, const uint8_t *active, const double *x,
                  uint8_t *stable, uint8_t *valid, uint8_t *mask, double *trial,
                  double *incipient, int32_t *status)
{
    const Handle &h = *static_cast<Handle *>(hv);
    return dispatchNc(h.desc.nc, [&](auto nc) {
        return stabilityBatch<decltype(nc)::value>(
            h, n, P, T, z, active, x, stable, valid, mask, trial, incipient, status);
    });
}

int ref_phase_update(void *hv, int64_t n, int hasWell, const double *T,
                     double *primary, double *phaseState, int32_t *status,
                     uint8_t *flags, int threads)
{
    const Handle &h = *static_cast<Handle *>(hv);
    return dispatchNc(h.desc.nc, [&](auto nc) {
        constexpr int NC = decltype(nc)::value;
        return hasWell
            ? phaseUpdate<NC, true>(h, n, T, primary, phaseState, status, flags, threads)
            : phaseUpdate<NC, false>(h, n, T, primary, phaseState, status, flags, threads);
    });
}

// ref_assembly with the runtime's equation/variable scaling: `scaling` = the six
// factors of ScalingRule (nullptr = disabled), `wellRowScale` per cell or nullptr.
int ref_assembly_scaled(void *hv, int64_t ncell, int hasWell, const int64_t *connOffsets,
                        const int64_t *connNeighbor, const double *connTrans,
                        const double *connGdz, const int64_t *rowOffsets,
                        const int64_t *colBlocks, const double *volume,
                        const double *porosity, const double *T, double dt,
                        double fugScale, const double *scaling, const double *wellRowScale,
                        const double *primary, const double *phaseState,
                        const double *accumPrev, double *residual, double *bsr,
                        uint8_t *decisions, double *accumOut, int threads)
{
    const Handle &h = *static_cast<Handle *>(hv);
    Topology t{ncell, connOffsets, connNeighbor, connTrans, connGdz, rowOffsets,
               colBlocks, volume, porosity, T};
    try
    {
        return dispatchNc(h.desc.nc, [&](auto nc) {
            constexpr int NC = decltype(nc)::value;
            return hasWell
                ? assemblyRun<NC, true>(h, t, dt, fugScale, scaling, wellRowScale, primary,
                                        phaseState, accumPrev, residual, bsr, decisions,
                                        accumOut, threads)
                : assemblyRun<NC, false>(h, t, dt, fugScale, scaling, wellRowScale, primary,
                                         phaseState, accumPrev, residual, bsr, decisions,
                                         accumOut, threads);
        });
    }
    catch (const std::exception &e)
    {
        std::snprintf(g_error, sizeof(g_error), "%s", e.what());
        return NN_E_INVALID_ARGUMENT;
    }
}

int ref_assembly_wells(void *hv, int64_t ncell, int hasWell, const int64_t *connOffsets,
                       const int64_t *connNeighbor, const double *connTrans,
                       const double *connGdz, const int64_t *rowOffsets,
                       const int64_t *colBlocks, const double *volume,
                       const double *porosity, const double *T, double dt,
                       double fugScale, const double *scaling, const double *wellRowScale,
                       int32_t nwells, const nn_well *wells, const int64_t *perfCells,
                       const double *perfIndex, const double *surfaceDensity,
                       const double *primary, const double *phaseState,
                       const double *accumPrev, double *residual, double *bsr,
                       uint8_t *decisions, int threads)
{
    const Handle &h = *static_cast<Handle *>(hv);
    Topology t{ncell, connOffsets, connNeighbor, connTrans, connGdz, rowOffsets,
               colBlocks, volume, porosity, T};
    const WellSet ws{nwells, wells, perfCells, perfIndex, surfaceDensity};
    try
    {
        return dispatchNc(h.desc.nc, [&](auto nc) {
            constexpr int NC = decltype(nc)::value;
            return hasWell
                ? assemblyRun<NC, true>(h, t, dt, fugScale, scaling, wellRowScale, primary,
                                        phaseState, accumPrev, residual, bsr, decisions,
                                        nullptr, threads, ws)
                : assemblyRun<NC, false>(h, t, dt, fugScale, scaling, wellRowScale, primary,
                                         phaseState, accumPrev, residual, bsr, decisions,
                                         nullptr, threads, ws);
        });
    }
    catch (const std::exception &e)
    {
        std::snprintf(g_error, sizeof(g_error), "%s", e.what());
        return NN_E_INVALID_ARGUMENT;
    }
}

int ref_assembly(void *hv, int64_t ncell, int hasWell, const int64_t *connOffsets,
                 const int64_t *connNeighbor, const double *connTrans,
                 const double *connGdz, const int64_t *rowOffsets,
                 const int64_t *colBlocks, const double *volume,
                 const double *porosity, const double *T, double dt,
                 double fugScale, const double *primary, const double *phaseState,
                 const double *accumPrev, double *residual, double *bsr,
                 uint8_t *decisions, double *accumOut, int threads)
{
    return ref_assembly_scaled(hv, ncell, hasWell, connOffsets, connNeighbor, connTrans, connGdz,
                               rowOffsets, colBlocks, volume, porosity, T, dt, fugScale, nullptr,
                               nullptr, primary, phaseState, accumPrev, residual, bsr, decisions,
                               accumOut, threads);
}

} // extern "C"
