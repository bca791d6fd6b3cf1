// common.hpp — device-side fluid parameters, options and small helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nonenone_amd.h"

namespace nn
{

constexpr int kMaxNc = NN_MAX_NC;
constexpr double kGasConstant = 8.31446261815324;     // common/include/common/units.hpp
constexpr double kBar = 1.0e5;
// Unit constants evaluated with the same expressions as common/units.hpp:19-53
// so that they round identically.
constexpr double kPsi = (0.45359237 * 9.80665) / (0.0254 * 0.0254);
constexpr double kKelvinPerRankine = 5.0 / 9.0;
constexpr double kCentipoise = 1.0e-2 * 0.1;
constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kDblMax = 1.7976931348623157e308;
constexpr double kLogDblMax = 709.782712893384;

// Phase slots / thermodynamic roles (phase_behavior.hpp:17-22).
constexpr int kOil = 0;
constexpr int kGas = 1;
constexpr int kWater = 2;
constexpr uint8_t kOilBit = 1, kGasBit = 2, kWaterBit = 4, kAllBits = 7;

// Error codes raised where the reference throws; they abort the enclosing
// public call exactly as the exception would.
enum : int
{
    kErrNone = 0,
    kErrThermo = NN_STATUS_THERMO_ERROR,
    kErrInvalid = NN_STATUS_INVALID_INPUT,
    // Internal: the state exceeded the lane-per-state work budget and is handed
    // to the wave-cooperative pass; never reported to callers.
    kErrDefer = 64,
    // Internal: the phase update recorded its stability test's seed-trial inputs
    // for the lockstep trial pass and stops; never reported to callers.
    kErrPending = 65,
    // Internal: the replayed phase update found a missing phase unstable and
    // stops before its unrestricted re-flash, which a later pass runs for all
    // such cells together; never reported to callers.
    kErrReflash = 66,
};

// Stability-test record of one cell (phase update, three passes): the
// lane-per-cell pass records the inputs of the seed trials, the lockstep trial
// pass fills in the best trial of every missing candidate phase, and the
// finishing pass replays the stability test from them.
enum : int
{
    kStabNone = 0,
    kStabRecord = 1,
    kStabReplay = 2,
    kStabReplayHoldReflash = 3,  // replay, and hand the re-flash to a later pass
};

template <int NC>
struct StabRecord
{
    double P, T;
    double z[NC], refComp[NC], logRef[NC];
    int32_t active;
    int32_t errCode;  // error a trial raised: the replay raises it where the serial test stops
    double sum[3];
    double comp[3][NC];
    uint8_t valid[3], trivial[3];
};

// Outcome of one item of a lane-per-item pass.
enum : int
{
    kItemDone = 0,
    kItemDeferred = 1,
    kItemPending = 2,
    kItemReflash = 3,
};

// Execution modes of the flash code.
//   kModePlain: one work-item per state, everything serial in that work-item.
//   kModeDefer: as plain, but a state whose Rachford-Rice recovery work passes
//               kDeferBudget, or (with kDeferAllocation) that enters the SW
//               allocation recovery, stops with kErrDefer so the wave-cooperative
//               pass can redo it.
//   kModeCoop:  one wavefront per state; every lane runs the state's serial
//               control flow on identical data, and the independent candidates
//               of the recovery searches (RR3 starting guesses, allocation
//               seeds) run on separate lanes.  Results are bit-identical to the
//               serial order: the reductions keep the reference's first-wins rules.
enum : int
{
    kModePlain = 0,
    kModeDefer = 1,
    kModeCoop = 2,
};
#ifndef NN_DEFER_BUDGET
#define NN_DEFER_BUDGET 20000
#endif
constexpr int kDeferBudget = NN_DEFER_BUDGET;  // RR3 residual evaluations per state
// Whether kModeDefer hands a state entering the SW allocation recovery to the
// wavefront-per-state pass (1) or runs the recovery's seeds serially on its lane (0).
#ifndef NN_DEFER_ALLOCATION
#define NN_DEFER_ALLOCATION 1
#endif
constexpr bool kDeferAllocation = NN_DEFER_ALLOCATION != 0;

// Wavefront helpers for kModeCoop.  The host build sees single-lane stand-ins,
// or, with NN_HOST_WAVE_EMULATION (test tooling: tests/coop_emulation.cpp), a
// 64-lane wavefront emulated by 64 host threads that meet at every wave operation.
#if !defined(__HIP_DEVICE_COMPILE__) && defined(NN_HOST_WAVE_EMULATION)
int nnHostWaveLane();
unsigned long long nnHostWaveBallot(bool pred);
double nnHostWaveShfl(double v, int src);
int nnHostWaveShfl(int v, int src);
#endif
__host__ __device__ inline int waveLane()
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __lane_id();
#elif defined(NN_HOST_WAVE_EMULATION)
    return nnHostWaveLane();
#else
    return 0;
#endif
}
__host__ __device__ inline unsigned long long waveBallot(bool pred)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __ballot(pred);
#elif defined(NN_HOST_WAVE_EMULATION)
    return nnHostWaveBallot(pred);
#else
    return pred ? 1ull : 0ull;
#endif
}
__host__ __device__ inline double waveShfl(double v, int src)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __shfl(v, src);
#elif defined(NN_HOST_WAVE_EMULATION)
    return nnHostWaveShfl(v, src);
#else
    (void)src;
    return v;
#endif
}
__host__ __device__ inline int waveShfl(int v, int src)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __shfl(v, src);
#elif defined(NN_HOST_WAVE_EMULATION)
    return nnHostWaveShfl(v, src);
#else
    (void)src;
    return v;
#endif
}
__host__ __device__ inline int firstSetLane(unsigned long long mask)
{
    return mask ? __builtin_ctzll(mask) : -1;
}

// Fluid parameters resident in HBM; kernels read them through uniform loads.
struct FluidParams
{
    int nc;
    int model;        // NN_MODEL_*
    int eosType;      // 1 or 5
    int water;        // H2O component (-1 if none)
    double omegaA, omegaB, m1, m2;
    double tc[kMaxNc], pc[kMaxNc], vc[kMaxNc], omega[kMaxNc], mw[kMaxNc];
    double kij[kMaxNc][kMaxNc];
    double swSalinity;
    double swAqKij[kMaxNc];
    int swAqSet[kMaxNc];
    // LBC constants (compositional_properties.hpp:166-182), computed on the host
    // with the reference's expression order.
    double lbcMwSqrt[kMaxNc];
    double lbcEps[kMaxNc];
    double relperm[7];
    // CPA
    int cpaTerm, cpaMaxIter;
    double cpaTol, cpaDamping;
    double cpaA0[kMaxNc], cpaB[kMaxNc], cpaC1[kMaxNc], cpaEps[kMaxNc], cpaBeta[kMaxNc];
    int cpaDonors[kMaxNc], cpaAcceptors[kMaxNc];
    double cpaCrossEps[kMaxNc][kMaxNc], cpaCrossBeta[kMaxNc][kMaxNc];
    // Water-role closures (nn_fluid_desc.aqueous_*)
    int aqVolume, aqViscosity, co2;
    double waterMolarMass, aqVolumeMax, aqViscosityMax;
};

struct FlashOptions
{
    int maxIterations, maxRrIterations, maxStabilityIterations;
    int stabilitySsi, maxStabilityNewton, twoPhaseSsi, maxTwoPhaseNewton;
    int threePhaseSsi, maxThreePhaseNewton;
    double fugTol, rrTol, stabTol, betaTol, coincidentTol, floor, logKStep, fdStep;
};

__host__ __device__ inline double clampd(double v, double lo, double hi)
{
    return v < lo ? lo : (hi < v ? hi : v);
}
__host__ __device__ inline double maxd(double a, double b) { return (a < b) ? b : a; }
__host__ __device__ inline double mind(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ inline int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (hi < v ? hi : v);
}
__host__ __device__ inline bool finite(double v) { return __builtin_isfinite(v); }

__host__ __device__ inline int popcount3(uint8_t b)
{
    return (b & 1) + ((b >> 1) & 1) + ((b >> 2) & 1);
}

} // namespace nn
