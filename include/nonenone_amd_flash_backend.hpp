// nonenone_amd_flash_backend.hpp — a FlashBackend for the reference's
// FullyCompositionalThreePhaseEquilibrium<Indices, FlashBackend>
// (models/include/natural/state/three_phase_equilibrium.hpp:98-117), backed by
// the GPU flash of libnonenone_amd.
//
// Compiled inside the reference's build (its headers on the include path):
//
//   GpuThreePhaseFlash<Indices>::setFluid(desc);   // once: the GPU fluid description
//   MPMC::FullyCompositionalThreePhaseEquilibrium<Indices, GpuThreePhaseFlash<Indices>> eq(fluid);
//
// The surface is the one the equilibrium manager uses (and the reference's
// test fakes implement, test/src/unit/three_phase_flash_test.cpp:319-413):
// construction from (EOS, ThreePhaseFlashOptions), options(), flash(P,T,z),
// flashRestricted(P,T,z,allowed[,seed]) and stabilityTest(P,T,z,active,x).
// Each call runs one state through the batched C ABI (device buffers owned by
// the backend); callers with many cells use the batched entry points directly.
// The fluid description is passed once through setFluid() because the
// reference's EOS keeps some of its parameters (SW options, cubic constants)
// private.  Errors map to the reference's exceptions: NN_STATUS_THERMO_ERROR ->
// std::runtime_error, NN_STATUS_INVALID_INPUT -> std::invalid_argument.
#pragma once

#include <natural/thermo/three_phase_flash.hpp>

#include <array>
#include <memory>
#include <stdexcept>

#include "nonenone_amd.hpp"

template <class Indices>
class GpuThreePhaseFlash
{
public:
    static constexpr int N = Indices::numComponents;
    using Result = MPMC::ThreePhaseFlashResult<Indices>;
    using StabilityResult = MPMC::ThreePhaseStabilityResult<Indices>;
    using EosType = MPMC::CubicEquationOfState<Indices>;
    using Composition = std::array<double, N>;

    static void setFluid(const nn_fluid_desc &desc) { fluidStore() = desc; }

    explicit GpuThreePhaseFlash(const EosType &, MPMC::ThreePhaseFlashOptions options = {})
        : options_(options), state_(std::make_shared<State>(fluidStore(), options))
    {
    }

    const MPMC::ThreePhaseFlashOptions &options() const noexcept { return options_; }

    Result flash(double P, double T, Composition z) const
    {
        State &s = *state_;
        s.upload(P, T, z);
        const nn_flash_out view = s.out.view();
        nonenone_amd::check(nn_flash_batch(s.ctx.get(), 1, s.P.get(), s.T.get(), s.z.get(), &view,
                                           nullptr),
                            "nn_flash_batch");
        return s.result();
    }

    Result flashRestricted(double P, double T, Composition z, MPMC::PhasePresence allowed) const
    {
        return restricted(P, T, z, allowed, nullptr);
    }

    Result flashRestricted(double P, double T, Composition z, MPMC::PhasePresence allowed,
                           const std::array<Composition, 3> &seed) const
    {
        return restricted(P, T, z, allowed, &seed);
    }

    StabilityResult stabilityTest(double P, double T, const Composition &z,
                                  MPMC::PhasePresence active,
                                  const std::array<Composition, 3> &x) const
    {
        State &s = *state_;
        s.upload(P, T, z);
        const uint8_t bits = active.bits();
        s.mask.upload(&bits, 1);
        s.seed.upload(flatten(x).data(), 3 * N);
        check(nn_stability_batch(s.ctx.get(), 1, s.P.get(), s.T.get(), s.z.get(), s.mask.get(),
                                 s.seed.get(), s.stable.get(), s.valid.get(), s.unstable.get(),
                                 s.trial.get(), s.incipient.get(), s.status.get(), nullptr),
              "nn_stability_batch");
        throwOnStatus(s.status.download(1)[0]);
        StabilityResult r;
        r.stable = s.stable.download(1)[0] != 0;
        r.valid = s.valid.download(1)[0] != 0;
        r.testedPresence = active;
        const uint8_t unstable = s.unstable.download(1)[0];
        const std::vector<double> trial = s.trial.download(3);
        const std::vector<double> inc = s.incipient.download(3 * N);
        for (int p = 0; p < 3; ++p)
        {
            r.missingPhaseUnstable[p] = (unstable >> p) & 1u;
            r.trialSum[p] = trial[p];
            for (int c = 0; c < N; ++c) r.incipientComposition[p][c] = inc[p * N + c];
        }
        return r;
    }

private:
    static nn_fluid_desc &fluidStore()
    {
        static nn_fluid_desc desc{};
        return desc;
    }

    static std::array<double, 3 * N> flatten(const std::array<Composition, 3> &x)
    {
        std::array<double, 3 * N> f{};
        for (int p = 0; p < 3; ++p)
            for (int c = 0; c < N; ++c) f[p * N + c] = x[p][c];
        return f;
    }

    static void check(int rc, const char *call) { nonenone_amd::check(rc, call); }

    static void throwOnStatus(int32_t status)
    {
        if (status == NN_STATUS_THERMO_ERROR)
            throw std::runtime_error("nonenone_amd: thermodynamic evaluation failed");
        if (status == NN_STATUS_INVALID_INPUT)
            throw std::invalid_argument("nonenone_amd: invalid pressure, temperature or composition");
    }

    static nn_flash_options translate(const MPMC::ThreePhaseFlashOptions &o)
    {
        nn_flash_options r;
        r.maximum_iterations = o.maximumIterations;
        r.maximum_rachford_rice_iterations = o.maximumRachfordRiceIterations;
        r.maximum_stability_iterations = o.maximumStabilityIterations;
        r.stability_ssi_before_newton = o.stabilitySsiIterationsBeforeNewton;
        r.maximum_stability_newton = o.maximumStabilityNewtonIterations;
        r.two_phase_ssi_before_newton = o.twoPhaseSsiIterationsBeforeNewton;
        r.maximum_two_phase_newton = o.maximumTwoPhaseNewtonIterations;
        r.three_phase_ssi_before_newton = o.threePhaseSsiIterationsBeforeNewton;
        r.maximum_three_phase_newton = o.maximumThreePhaseNewtonIterations;
        r.fugacity_tolerance = o.fugacityTolerance;
        r.rachford_rice_tolerance = o.rachfordRiceTolerance;
        r.stability_tolerance = o.stabilityTolerance;
        r.phase_fraction_tolerance = o.phaseFractionTolerance;
        r.coincident_composition_tolerance = o.coincidentPhaseCompositionTolerance;
        r.composition_floor = o.compositionFloor;
        r.log_k_step_limit = o.logKStepLimit;
        r.numerical_jacobian_step = o.numericalJacobianStep;
        return r;
    }

    // Device buffers for one state; shared by copies of the backend.
    struct State
    {
        State(const nn_fluid_desc &desc, const MPMC::ThreePhaseFlashOptions &o)
            : ctx(desc, &optionsFor(o)), P(1), T(1), z(N), mask(1), seed(3 * N), out(1, N),
              stable(1), valid(1), unstable(1), trial(3), incipient(3 * N), status(1)
        {
        }
        static const nn_flash_options &optionsFor(const MPMC::ThreePhaseFlashOptions &o)
        {
            static thread_local nn_flash_options r;
            r = translate(o);
            return r;
        }
        void upload(double p, double t, const Composition &c)
        {
            P.upload(&p, 1);
            T.upload(&t, 1);
            z.upload(c.data(), N);
        }
        Result result() const
        {
            const nonenone_amd::FlashBatch b = out.download();
            throwOnStatus(b.status[0]);
            Result r;
            r.converged = b.converged[0] != 0;
            r.iterations = b.iterations[0];
            r.presence = MPMC::PhasePresence(b.presence[0]);
            for (int p = 0; p < 3; ++p)
            {
                r.phaseMoleFraction[p] = b.beta[p];
                r.saturation[p] = b.saturation[p];
                r.compressibility[p] = b.Z[p];
                r.molarDensity[p] = b.rho_m[p];
                for (int c = 0; c < N; ++c) r.composition[p][c] = b.x[p * N + c];
            }
            for (int c = 0; c < N; ++c)
            {
                r.vaporOilK[c] = b.k_go[c];
                r.waterOilK[c] = b.k_wo[c];
            }
            return r;
        }

        nonenone_amd::Context ctx;
        nonenone_amd::DeviceArray<double> P, T, z;
        nonenone_amd::DeviceArray<uint8_t> mask;
        nonenone_amd::DeviceArray<double> seed;
        nonenone_amd::DeviceFlashBatch out;
        nonenone_amd::DeviceArray<uint8_t> stable, valid, unstable;
        nonenone_amd::DeviceArray<double> trial, incipient;
        nonenone_amd::DeviceArray<int32_t> status;
    };

    Result restricted(double P, double T, const Composition &z, MPMC::PhasePresence allowed,
                      const std::array<Composition, 3> *seed) const
    {
        State &s = *state_;
        s.upload(P, T, z);
        const uint8_t bits = allowed.bits();
        s.mask.upload(&bits, 1);
        if (seed)
            s.seed.upload(flatten(*seed).data(), 3 * N);
        const nn_flash_out view = s.out.view();
        check(nn_flash_restricted_batch(s.ctx.get(), 1, s.P.get(), s.T.get(), s.z.get(),
                                        s.mask.get(), seed ? s.seed.get() : nullptr, &view, nullptr),
              "nn_flash_restricted_batch");
        return s.result();
    }

    MPMC::ThreePhaseFlashOptions options_;
    std::shared_ptr<State> state_;
};
