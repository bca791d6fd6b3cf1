// newton_limiter.hpp — per-cell Newton-step limiter of the Natural runtime.
//
// Follows (paths relative to the reference checkout):
//   NaturalPetscRuntime::limitNewtonStep  models/include/natural/petsc/detail/
//                                         natural_petsc_runtime_state.inc:452-545
//   limitNaturalNewtonIncrement (fully compositional three-phase)
//                                         models/include/natural/state/newton_limiter.hpp:124-179
//   limitAqueousCompositionNewtonIncrement newton_limiter.hpp:181-245
//   limits                                models/include/natural/numerics.hpp:55-84
// on the primary layout of Layout<NC, Well> (pressure, three composition blocks
// of NC - 1 fractions, three saturations, well pressure).
#pragma once

#include "common.hpp"
#include "phase_update.hpp"

namespace nn
{

constexpr double kMaxSaturationNewtonChange = 0.1;         // numerics.hpp:81
constexpr double kMaxCompositionNewtonChange = 0.1;        // numerics.hpp:82
constexpr double kMaxRelativePressureNewtonChange = 0.25;  // numerics.hpp:84
constexpr double kPhaseEquilibriumTrace = 1.0e-14;         // numerics.hpp:67

// newton_limiter.hpp:18-24
__host__ __device__ inline double limiterScale(double magnitude, double maximum)
{
    if (!(magnitude > 0.0) || !finite(magnitude))
        return 1.0;
    return mind(maximum / magnitude, 1.0);
}

// One cell: `step` (PETSc's correction Y) becomes -(limited delta), delta = -Y.
// scale: the variable scale factors when scaling is on, else null.
// maxAqueous < 0: no aqueous composition-domain limiting.
template <int NC, bool Well>
__host__ __device__ void limitNewtonCell(const double *state, double *step, const double *scale,
                                         double maxAqueous, int water, int co2)
{
    using L = Layout<NC, Well>;
    double delta[L::NP];
    for (int q = 0; q < L::NP; ++q) delta[q] = -step[q];
    if (scale)
        for (int q = 0; q < L::NP; ++q) delta[q] *= scale[q];

    // saturations share one scale (newton_limiter.hpp:133-143)
    double maxSat = ::fabs(delta[L::satBegin]);
    for (int p = 1; p < 3; ++p)
    {
        const double v = ::fabs(delta[L::satBegin + p]);
        if (maxSat < v)
            maxSat = v;
    }
    const double satScale = limiterScale(maxSat, kMaxSaturationNewtonChange);
    for (int p = 0; p < 3; ++p) delta[L::satBegin + p] *= satScale;

    // each phase's composition block on its own (:145-171)
    for (int p = 0; p < 3; ++p)
    {
        double maximum = 0.0, dependent = 0.0;
        for (int i = 0; i < NC - 1; ++i)
        {
            const double v = delta[L::compBegin(p) + i];
            maximum = maxd(maximum, ::fabs(v));
            dependent -= v;
        }
        maximum = maxd(maximum, ::fabs(dependent));
        const double s = limiterScale(maximum, kMaxCompositionNewtonChange);
        for (int i = 0; i < NC - 1; ++i) delta[L::compBegin(p) + i] *= s;
    }

    // pressure, relative (:173-179)
    const double pressure = ::fabs(state[L::pressure]);
    if (pressure > kMinNormDenominator)
    {
        const double relative = ::fabs(delta[L::pressure]) / pressure;
        delta[L::pressure] *= limiterScale(relative, kMaxRelativePressureNewtonChange);
    }

    // Water composition block inside the aqueous closures' domain (:181-245)
    if (maxAqueous >= 0.0)
    {
        double unsupported = 0.0, unsupportedDelta = 0.0;
        double dependent = 1.0, dependentDelta = 0.0;
        for (int i = 0; i < NC - 1; ++i)
        {
            const double value = state[L::compBegin(kWater) + i];
            const double change = delta[L::compBegin(kWater) + i];
            dependent -= value;
            dependentDelta -= change;
            if (i != water && i != co2)
            {
                unsupported += value;
                unsupportedDelta += change;
            }
        }
        if (NC - 1 != water && NC - 1 != co2)
        {
            unsupported += dependent;
            unsupportedDelta += dependentDelta;
        }
        if (unsupportedDelta > 0.0)
        {
            const double margin = maxd(kPhaseEquilibriumTrace, maxAqueous * 1.0e-10);
            const double target = maxd(0.0, maxAqueous - margin);
            const double s = clampd((target - unsupported) / unsupportedDelta, 0.0, 1.0);
            for (int i = 0; i < NC - 1; ++i) delta[L::compBegin(kWater) + i] *= s;
        }
    }
    for (int q = 0; q < L::NP; ++q) step[q] = -delta[q];
}

} // namespace nn
