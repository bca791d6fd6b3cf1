// nn_api.hip — the C ABI declared in include/nonenone_amd.h (kernels: kernels.hpp).
//
// One HIP work-item owns one state (flash, stability, phase update) or one
// cell (assembly).  Flash-like kernels are persistent: a grid sized to the
// chip pulls state indices from a device counter, so work-items that finish
// early take new states instead of idling while slow states iterate.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/nonenone_amd.h"
#include "common.hpp"
#include "cpa.hpp"
#include "launch.hpp"
#include "newton_limiter.hpp"

using namespace nn;

struct nn_context
{
    int device = 0;
    FluidParams host{};
    FlashOptions hostOpts{};
    FluidParams *dFluid = nullptr;
    FlashOptions *dOpts = nullptr;
    WorkQueue *dQueue = nullptr;       // counters of the two-pass persistent launches
    unsigned int *dList = nullptr;     // deferred item indices (capacity listCap)
    int64_t listCap = 0;
    // Lockstep stability trials of the phase update (NN_STAB_LOCKSTEP=0 disables):
    // recorded cell indices and one StabRecord per cell.
    bool lockstep = true;
    unsigned int *dPending = nullptr;
    void *dRecords = nullptr;
    int64_t recordCap = 0;
    size_t recordBytes = 0;
    int numCUs = 256;
    int64_t maxBlocks = 0;  // NN_PERSISTENT_BLOCKS override of the persistent grid (0 = auto)
    int64_t coopBlocks = 0; // NN_COOP_BLOCKS override of the cooperative grid (0 = auto)
    bool coopAll = false;   // NN_COOP_ALL=1: every item through the cooperative pass (testing)
    // Launch shapes of flash-like batches (nn_context_set_schedule / nn_flash_autotune).
    // Processing order of large batches (radix-sorted condition keys): sched.ordered,
    // NN_ORDER=0 disables.
    nn_schedule sched{16, 4, 1, 0, 0};
    unsigned int *dKeys[2] = {nullptr, nullptr};
    unsigned int *dIdx[2] = {nullptr, nullptr};
    void *dSortTemp = nullptr;
    size_t sortTempBytes = 0;
    int64_t sortCap = 0;
    // Kernel timing: (start, end) event pairs per kernel kind while enabled.
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events[NN_KERNEL_KINDS];
};

namespace
{

thread_local char g_error[512] = "";

int fail(int code, const char *msg)
{
    std::snprintf(g_error, sizeof(g_error), "%s", msg);
    return code;
}

int hipFail(hipError_t e, const char *what)
{
    std::snprintf(g_error, sizeof(g_error), "%s: %s", what, hipGetErrorString(e));
    return NN_E_HIP;
}

#define NN_HIP(call)                                  \
    do                                                \
    {                                                 \
        hipError_t e_ = (call);                       \
        if (e_ != hipSuccess) return hipFail(e_, #call); \
    } while (0)


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
    default:
        return fail(NN_E_UNSUPPORTED,
                    "component count not compiled into this library (2, 3, 4, 5, 6, 8)");
    }
}


// Work-queue counters reset by a kernel rather than hipMemsetAsync, so that
// the launch sequence stays valid inside a captured HIP graph.
template <int NC, bool Well>
__global__ void newtonLimitKernel(int64_t n, const double *__restrict__ primary,
                                  double *__restrict__ step, AssemblyView v, double maxAqueous,
                                  int water, int co2)
{
    using L = Layout<NC, Well>;
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n)
        return;
    limitNewtonCell<NC, Well>(primary + c * L::NP, step + c * L::NP, v.varScale, maxAqueous,
                              water, co2);
}

// Output sample gather (cpgrid_output_access.hpp:42-140): value (i, k) is
// variable vars[k] of cell cells[i]; one thread per value, consecutive threads
// on consecutive variables of one cell.
__global__ void stateSampleKernel(const double *__restrict__ vec, int64_t ncell, int dof,
                                  const int64_t *__restrict__ cells, int64_t nsel,
                                  const int32_t *__restrict__ vars, int nvar,
                                  double *__restrict__ out, int *__restrict__ bad)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nsel * nvar)
        return;
    const int64_t i = t / nvar;
    const int k = int(t - i * nvar);
    const int64_t c = cells[i];
    const int v = vars[k];
    if (c < 0 || c >= ncell || v < 0 || v >= dof)
    {
        atomicOr(bad, 1);
        out[t] = __builtin_nan("");
        return;
    }
    out[t] = vec[c * dof + v];
}

__global__ void clearFlagKernel(int *flag) { *flag = 0; }

// Blocks of the pattern that no face or diagonal term writes (the wells'
// couplings): zeros before the assembly.
__global__ void zeroBlocksKernel(double *__restrict__ bsr, const int64_t *__restrict__ slots,
                                 int64_t count, int np)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t per = int64_t(np) * np;
    if (t >= count * per)
        return;
    bsr[slots[t / per] * per + t % per] = 0.0;
}

__global__ void resetQueueKernel(WorkQueue *q, unsigned int deferAll, unsigned int claim)
{
    q->head = 0u;
    q->deferred = 0u;
    q->coopHead = 0u;
    q->deferAll = deferAll;
    q->claim = claim;
    q->pending = 0u;
    q->trialHead = 0u;
    q->replayHead = 0u;
    q->reflash = 0u;
    q->reflashHead = 0u;
}

void releaseEvents(nn_context *ctx)
{
    for (auto &list : ctx->events)
    {
        for (auto &pr : list)
        {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
        list.clear();
    }
}

// Brackets one launch with events on its stream when the context is timing.
struct LaunchTimer
{
    nn_context *ctx;
    int kind;
    hipStream_t st;
    hipEvent_t start = nullptr;
    LaunchTimer(nn_context *c, int k, hipStream_t s) : ctx(c), kind(k), st(s)
    {
        if (ctx->timing && hipEventCreate(&start) == hipSuccess)
            (void)hipEventRecord(start, st);
    }
    void done()
    {
        if (!start)
            return;
        hipEvent_t end = nullptr;
        if (hipEventCreate(&end) == hipSuccess && hipEventRecord(end, st) == hipSuccess)
            ctx->events[kind].emplace_back(start, end);
        else
            (void)hipEventDestroy(start);
        start = nullptr;
    }
    // Events between the two passes of a two-pass launch (null when not
    // timing): `splitA` ends the first interval, `splitB` starts the second.
    hipEvent_t splitA = nullptr, splitB = nullptr;
    void makeSplit()
    {
        if (!start)
            return;
        if (hipEventCreate(&splitA) != hipSuccess)
            splitA = nullptr;
        if (hipEventCreate(&splitB) != hipSuccess)
            splitB = nullptr;
    }
    // Ends a two-pass launch: (start, splitA) under `kind`, (splitB, end) under `coopKind`.
    void doneSplit(int coopKind)
    {
        if (!start || !splitA || !splitB)
        {
            done();
            return;
        }
        hipEvent_t end = nullptr;
        if (hipEventCreate(&end) == hipSuccess && hipEventRecord(end, st) == hipSuccess)
        {
            ctx->events[kind].emplace_back(start, splitA);
            ctx->events[coopKind].emplace_back(splitB, end);
            start = splitA = splitB = nullptr;
        }
    }
    ~LaunchTimer()
    {
        if (start)
            (void)hipEventDestroy(start);
        if (splitA)
            (void)hipEventDestroy(splitA);
        if (splitB)
            (void)hipEventDestroy(splitB);
    }
};

// Makes room for n deferred indices (synchronous allocation when it grows).
// Pending list and stability records for n cells of `bytes` each (synchronous when it grows).
int ensureRecords(nn_context *ctx, int64_t n, size_t bytes)
{
    if (n <= ctx->recordCap && bytes == ctx->recordBytes)
        return NN_OK;
    if (ctx->dPending) (void)hipFree(ctx->dPending);
    if (ctx->dRecords) (void)hipFree(ctx->dRecords);
    ctx->dPending = nullptr;
    ctx->dRecords = nullptr;
    ctx->recordCap = 0;
    if (hipMalloc(&ctx->dPending, 2 * sizeof(unsigned int) * size_t(n)) != hipSuccess ||
        hipMalloc(&ctx->dRecords, bytes * size_t(n)) != hipSuccess)
        return fail(NN_E_OUT_OF_MEMORY, "device allocation of the stability records failed");
    ctx->recordCap = n;
    ctx->recordBytes = bytes;
    return NN_OK;
}

int ensureList(nn_context *ctx, int64_t n)
{
    if (n <= ctx->listCap)
        return NN_OK;
    if (ctx->dList)
    {
        (void)hipFree(ctx->dList);
        ctx->dList = nullptr;
        ctx->listCap = 0;
    }
    if (hipMalloc(&ctx->dList, sizeof(unsigned int) * size_t(n)) != hipSuccess)
        return fail(NN_E_OUT_OF_MEMORY, "device allocation of the deferred list failed");
    ctx->listCap = n;
    return NN_OK;
}

constexpr int64_t kOrderMinItems = 4096;  // smaller batches run in index order
// hipcub's radix sort takes an int item count: larger batches run in index order.
constexpr int64_t kOrderMaxItems = int64_t(0x7FFFFFFF);

void releaseSort(nn_context *ctx)
{
    for (int k = 0; k < 2; ++k)
    {
        if (ctx->dKeys[k]) (void)hipFree(ctx->dKeys[k]);
        if (ctx->dIdx[k]) (void)hipFree(ctx->dIdx[k]);
        ctx->dKeys[k] = ctx->dIdx[k] = nullptr;
    }
    if (ctx->dSortTemp) (void)hipFree(ctx->dSortTemp);
    ctx->dSortTemp = nullptr;
    ctx->sortTempBytes = 0;
    ctx->sortCap = 0;
}

// Key/index buffers and radix-sort scratch for n items (synchronous when it grows).
int ensureSort(nn_context *ctx, int64_t n)
{
    if (n > kOrderMaxItems)
        return fail(NN_E_INVALID_ARGUMENT, "ordered batches hold at most 2^31-1 items");
    if (n <= ctx->sortCap)
        return NN_OK;
    releaseSort(ctx);
    const size_t bytes = sizeof(unsigned int) * size_t(n);
    for (int k = 0; k < 2; ++k)
        if (hipMalloc(&ctx->dKeys[k], bytes) != hipSuccess || hipMalloc(&ctx->dIdx[k], bytes) != hipSuccess)
        {
            releaseSort(ctx);
            return fail(NN_E_OUT_OF_MEMORY, "device allocation of the ordering buffers failed");
        }
    size_t temp = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, temp, ctx->dKeys[0], ctx->dKeys[1], ctx->dIdx[0],
                                           ctx->dIdx[1], int(n)) != hipSuccess ||
        hipMalloc(&ctx->dSortTemp, temp > 0 ? temp : 1) != hipSuccess)
    {
        releaseSort(ctx);
        return fail(NN_E_OUT_OF_MEMORY, "device allocation of the sort scratch failed");
    }
    ctx->sortTempBytes = temp;
    ctx->sortCap = n;
    return NN_OK;
}

// Sorts the keys written into dKeys[0]/dIdx[0]; returns the ordered indices.
hipError_t sortOrder(nn_context *ctx, int64_t n, hipStream_t st, const unsigned int **order)
{
    size_t temp = ctx->sortTempBytes;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(ctx->dSortTemp, temp, ctx->dKeys[0],
                                                      ctx->dKeys[1], ctx->dIdx[0], ctx->dIdx[1],
                                                      int(n), 0, 32, st);
    *order = ctx->dIdx[1];
    return e;
}

struct ReflashSortArg
{
    nn_context *ctx;
    int64_t n;
};

hipError_t reflashSort(void *arg, hipStream_t st, const unsigned int **order)
{
    const auto *a = static_cast<const ReflashSortArg *>(arg);
    return sortOrder(a->ctx, a->n, st, order);
}

// Ordered batches sort with hipCUB, whose radix sort issues hipMemsetAsync; a
// memset recorded into a captured graph is not replayed reliably here, so a
// batch issued on a capturing stream runs in index order.
bool streamCapturing(hipStream_t st)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive;
}

unsigned int coopGrid(const nn_context *ctx)
{
    return static_cast<unsigned int>(ctx->coopBlocks > 0
                                         ? ctx->coopBlocks
                                         : int64_t(ctx->numCUs) * ctx->sched.coop_blocks_per_cu);
}

unsigned int persistentGrid(const nn_context *ctx, int64_t n)
{
    // Default 16 wavefront-blocks per CU keeps every SIMD busy while the queue drains.
    const int64_t cap = ctx->maxBlocks > 0
                            ? ctx->maxBlocks
                            : int64_t(ctx->numCUs) * ctx->sched.persistent_blocks_per_cu;
    const int64_t need = (n + kBlock - 1) / kBlock;
    return static_cast<unsigned int>(std::max<int64_t>(1, std::min(cap, need)));
}

// 32-bit work-queue positions: the batch plus the overshoot of the last claims
// (at most one 64-item tile per block) must stay below 2^32.
constexpr int64_t kMaxQueueItems = int64_t(0xF0000000u);

// Allocation-free flash (nn_flash_batch_static): its work queue and deferred-index
// list live in this file-scope device array, one copy per GPU, held until the
// process ends.  Layout: WorkQueue (padded to 256 bytes), then up to
// kStaticMaxStates deferred indices.  Host code uses only the per-device address
// from hipGetSymbolAddress, never the array's name.
constexpr int64_t kStaticMaxStates = NN_STATIC_MAX_STATES;
static_assert(NN_STATIC_MAX_STATES == 1048576, "error text states the limit");
constexpr size_t kStaticQueueBytes = 256;
static_assert(sizeof(WorkQueue) <= kStaticQueueBytes, "work queue exceeds its static slot");
constexpr size_t kStaticScratchBytes =
    kStaticQueueBytes + sizeof(unsigned int) * size_t(kStaticMaxStates);
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
void *g_scratchPtr[kMaxDevices] = {};

// Resolves this device's copy of `scratch` once (device current); NN_OK or NN_E_HIP.
int resolveScratch(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return fail(NN_E_UNSUPPORTED, "device index beyond the static scratch table");
    std::lock_guard<std::mutex> lk(g_scratchMutex);
    if (g_scratchPtr[device])
        return NN_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
    {
        std::snprintf(g_error, sizeof(g_error), "hipGetSymbolAddress(scratch): %s",
                      hipGetErrorString(e));
        return NN_E_HIP;
    }
    g_scratchPtr[device] = ptr;
    return NN_OK;
}

void *scratchPtr(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return nullptr;
    std::lock_guard<std::mutex> lk(g_scratchMutex);
    return g_scratchPtr[device];
}

FluidParams toParams(const nn_fluid_desc &d)
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
        for (int j = 0; j < NN_MAX_NC; ++j)
        {
            p.kij[i][j] = d.kij[i * NN_MAX_NC + j];
            p.cpaCrossEps[i][j] = d.cpa_cross_epsilon[i * NN_MAX_NC + j];
            p.cpaCrossBeta[i][j] = d.cpa_cross_beta[i * NN_MAX_NC + j];
        }
        p.swAqKij[i] = d.sw_aqueous_kij[i];
        p.swAqSet[i] = d.sw_aqueous_kij_set[i];
        p.cpaA0[i] = d.cpa_a0[i];
        p.cpaB[i] = d.cpa_b[i];
        p.cpaC1[i] = d.cpa_c1[i];
        p.cpaEps[i] = d.cpa_epsilon[i];
        p.cpaBeta[i] = d.cpa_beta[i];
        p.cpaDonors[i] = d.cpa_donors[i];
        p.cpaAcceptors[i] = d.cpa_acceptors[i];
        // LBC constants, compositional_properties.hpp:166-182 (host libm, as the reference).
        if (i < d.nc)
        {
            const double mwSqrt = std::sqrt(1000.0 * d.molar_mass[i]);
            const double tcR = d.tc[i] / kKelvinPerRankine;
            const double pcPsi = d.pc[i] / kPsi;
            p.lbcMwSqrt[i] = mwSqrt;
            p.lbcEps[i] = 5.4402 * std::pow(tcR, 1.0 / 6.0) /
                          (mwSqrt * std::pow(pcPsi, 2.0 / 3.0) * kCentipoise);
        }
    }
    p.swSalinity = d.sw_salinity;
    for (int k = 0; k < 7; ++k) p.relperm[k] = d.relperm[k];
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

FlashOptions toOptions(const nn_flash_options &o)
{
    FlashOptions r{};
    r.maxIterations = o.maximum_iterations;
    r.maxRrIterations = o.maximum_rachford_rice_iterations;
    r.maxStabilityIterations = o.maximum_stability_iterations;
    r.stabilitySsi = o.stability_ssi_before_newton;
    r.maxStabilityNewton = o.maximum_stability_newton;
    r.twoPhaseSsi = o.two_phase_ssi_before_newton;
    r.maxTwoPhaseNewton = o.maximum_two_phase_newton;
    r.threePhaseSsi = o.three_phase_ssi_before_newton;
    r.maxThreePhaseNewton = o.maximum_three_phase_newton;
    r.fugTol = o.fugacity_tolerance;
    r.rrTol = o.rachford_rice_tolerance;
    r.stabTol = o.stability_tolerance;
    r.betaTol = o.phase_fraction_tolerance;
    r.coincidentTol = o.coincident_composition_tolerance;
    r.floor = o.composition_floor;
    r.logKStep = o.log_k_step_limit;
    r.fdStep = o.numerical_jacobian_step;
    return r;
}

// ---- all-or-none phase update (natural_petsc_runtime_state.inc:166-334) ----
// The transaction's device record: the lowest failing cell index (atomicMin) and
// the report the caller reads.
__global__ void txnInitKernel(nn_update_report *r)
{
    r->first_failure_cell = -1;
    r->updated_cells = 0;
    r->stable_reduced_cells = 0;
    r->succeeded = 0;
    r->first_failure_status = NN_STATUS_OK;
    r->first_failure_flags = 0;
    r->reserved = 0;
}

// first failing cell = lowest index whose status is not OK; the 64-bit slot
// first_failure_cell holds n (no failure) before the reduction.
__global__ void txnFirstFailureKernel(int64_t n, const int32_t *__restrict__ status,
                                      unsigned long long *__restrict__ first)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c < n && status[c] != NN_STATUS_OK)
        atomicMin(first, static_cast<unsigned long long>(c));
}

// StableReducedSet updates before the first failure (the reference counts the
// cells it updated before it stopped).
__global__ void txnCountKernel(int64_t n, const uint8_t *__restrict__ flags,
                               const unsigned long long *__restrict__ first,
                               unsigned long long *__restrict__ stable)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t limit = static_cast<int64_t>(*first);
    if (c < n && c < limit && (flags[c] & 0x20))
        atomicAdd(stable, 1ull);
}

__global__ void txnReportKernel(int64_t n, const int32_t *__restrict__ status,
                                const uint8_t *__restrict__ flags, nn_update_report *r)
{
    const int64_t first = r->first_failure_cell;   // n when nothing failed
    r->updated_cells = first;
    if (first >= n)
    {
        r->first_failure_cell = -1;
        r->succeeded = 1;
        return;
    }
    r->succeeded = 0;
    r->first_failure_status = status[first];
    r->first_failure_flags = flags[first];
}

// Commit: the updated copies replace the caller's vectors only on success.
__global__ void txnCommitKernel(int64_t count, const nn_update_report *__restrict__ r,
                                const double *__restrict__ src, double *__restrict__ dst)
{
    if (!r->succeeded)
        return;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += int64_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

__global__ void txnSetKernel(unsigned long long *slot, unsigned long long value)
{
    *slot = value;
}

} // namespace

// ===================================================================== C ABI
extern "C" {

void nn_flash_options_default(nn_flash_options *o)
{
    if (!o)
        return;
    o->maximum_iterations = 120;
    o->maximum_rachford_rice_iterations = 80;
    o->maximum_stability_iterations = 80;
    o->stability_ssi_before_newton = 8;
    o->maximum_stability_newton = 12;
    o->two_phase_ssi_before_newton = 8;
    o->maximum_two_phase_newton = 12;
    o->three_phase_ssi_before_newton = 10;
    o->maximum_three_phase_newton = 12;
    o->fugacity_tolerance = 1.0e-9;
    o->rachford_rice_tolerance = 1.0e-12;
    o->stability_tolerance = 1.0e-9;
    o->phase_fraction_tolerance = 1.0e-10;
    o->coincident_composition_tolerance = 1.0e-8;
    o->composition_floor = 1.0e-30;
    o->log_k_step_limit = 2.0;
    o->numerical_jacobian_step = 1.0e-6;
}

const char *nn_last_error(void) { return g_error; }

int nn_context_create(const nn_fluid_desc *fluid, const nn_flash_options *opts, nn_context **out)
{
    if (!fluid || !out)
        return fail(NN_E_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (fluid->nc < 2 || fluid->nc > NN_MAX_NC)
        return fail(NN_E_INVALID_ARGUMENT, "nc out of range");
    // The five- and eight-component kernels are switched off: the five-component
    // phase update + assembly made a GPU memory fault whose cause in this code is
    // not identified (a host build of the same kernel bodies on the same inputs is
    // clean under ASan/UBSan and matches the reference bit for bit).
    if (fluid->nc == 5 || fluid->nc == 8)
        return fail(NN_E_UNSUPPORTED, "the five- and eight-component device builds are switched off");
    if (fluid->eos_type != 1 && fluid->eos_type != 5)
        return fail(NN_E_INVALID_ARGUMENT, "eos_type must be 1 or 5");
    if (!(fluid->omega_a > 0.0) || !(fluid->omega_b > 0.0) || fluid->m1 == fluid->m2)
        return fail(NN_E_INVALID_ARGUMENT, "invalid cubic EOS constants");
    if (fluid->thermo_model == NN_MODEL_CPA)
    {
        if (cpaScheme(toParams(*fluid)) == kCpaGeneral && fluid->nc > kCpaGeneralMaxNc)
            return fail(NN_E_UNSUPPORTED,
                        "CPA with several donor-bearing components is built for nc <= 4");
        if (fluid->cpa_physical_term != 0 && fluid->cpa_physical_term != 1)
            return fail(NN_E_INVALID_ARGUMENT, "cpa_physical_term must be 0 (SRK) or 1 (PR)");
        for (int i = 0; i < fluid->nc; ++i)
            if (!(fluid->cpa_b[i] > 0.0) || !(fluid->cpa_a0[i] > 0.0))
                return fail(NN_E_INVALID_ARGUMENT, "CPA a0 and b must be positive");
    }
    if (fluid->thermo_model == NN_MODEL_SW &&
        (fluid->water_component < 0 || fluid->water_component >= fluid->nc))
        return fail(NN_E_INVALID_ARGUMENT, "SW water-component index is out of range");
    // Water-role closures (IapwsGarciaAqueousVolume / McBrideWright2015AqueousViscosity /
    // Iapws2008IndustrialAqueousViscosity constructors: aqueous_volume.hpp:347-361,
    // aqueous_viscosity.hpp:33-41, 114-128)
    if (fluid->aqueous_volume != 0 && fluid->aqueous_volume != 1)
        return fail(NN_E_INVALID_ARGUMENT, "aqueous_volume must be 0 or 1");
    if (fluid->aqueous_viscosity < NN_AQ_VISCOSITY_NONE ||
        fluid->aqueous_viscosity > NN_AQ_VISCOSITY_IAPWS2008)
        return fail(NN_E_INVALID_ARGUMENT, "unknown aqueous_viscosity model");
    const bool needsCo2 = fluid->aqueous_volume != 0 ||
                          fluid->aqueous_viscosity == NN_AQ_VISCOSITY_MCBRIDE_WRIGHT;
    if ((needsCo2 || fluid->aqueous_viscosity != NN_AQ_VISCOSITY_NONE) &&
        (fluid->water_component < 0 || fluid->water_component >= fluid->nc))
        return fail(NN_E_INVALID_ARGUMENT, "aqueous closures need a water component");
    if (needsCo2 && (fluid->co2_component < 0 || fluid->co2_component >= fluid->nc ||
                     fluid->co2_component == fluid->water_component))
        return fail(NN_E_INVALID_ARGUMENT, "invalid aqueous H2O/CO2 component indices");
    if (fluid->aqueous_volume != 0 &&
        (!(fluid->water_molar_mass > 0.0) || !(fluid->aqueous_volume_max_unsupported >= 0.0)))
        return fail(NN_E_INVALID_ARGUMENT, "invalid IAPWS-Garcia configuration");
    if (fluid->aqueous_viscosity != NN_AQ_VISCOSITY_NONE &&
        !(fluid->aqueous_viscosity_max_unsupported >= 0.0))
        return fail(NN_E_INVALID_ARGUMENT, "invalid aqueous-viscosity configuration");
    for (int i = 0; i < fluid->nc; ++i)
        for (int j = 0; j < fluid->nc; ++j)
            if (fluid->kij[i * NN_MAX_NC + j] != fluid->kij[j * NN_MAX_NC + i])
                return fail(NN_E_INVALID_ARGUMENT, "k_ij must be symmetric");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(NN_E_NO_DEVICE, "no HIP device");
    nn_context *ctx = new (std::nothrow) nn_context;
    if (!ctx)
        return fail(NN_E_OUT_OF_MEMORY, "host allocation failed");
    NN_HIP(hipGetDevice(&ctx->device));
    if (const int rc = resolveScratch(ctx->device); rc != NN_OK)
    {
        delete ctx;
        return rc;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->numCUs = prop.multiProcessorCount;
    if (const char *env = std::getenv("NN_PERSISTENT_BLOCKS"))
        ctx->maxBlocks = std::max(0L, std::atol(env));
    if (const char *env = std::getenv("NN_COOP_BLOCKS"))
        ctx->coopBlocks = std::max(0L, std::atol(env));
    if (const char *env = std::getenv("NN_ORDER"))
        ctx->sched.ordered = std::atoi(env) != 0 ? 1 : 0;
    if (const char *env = std::getenv("NN_WAVE_CLAIM"))
        ctx->sched.wave_claim = std::atoi(env) != 0 ? 1 : 0;
    if (const char *env = std::getenv("NN_STAB_LOCKSTEP"))
        ctx->lockstep = std::atoi(env) != 0;
    if (const char *env = std::getenv("NN_COOP_ALL"))
        ctx->coopAll = std::atoi(env) != 0;
    nn_flash_options defaults;
    nn_flash_options_default(&defaults);
    ctx->host = toParams(*fluid);
    ctx->hostOpts = toOptions(opts ? *opts : defaults);
    if (hipMalloc(&ctx->dFluid, sizeof(FluidParams)) != hipSuccess ||
        hipMalloc(&ctx->dOpts, sizeof(FlashOptions)) != hipSuccess ||
        hipMalloc(&ctx->dQueue, sizeof(WorkQueue)) != hipSuccess)
    {
        nn_context_destroy(ctx);
        return fail(NN_E_OUT_OF_MEMORY, "device allocation failed");
    }
    NN_HIP(hipMemcpy(ctx->dFluid, &ctx->host, sizeof(FluidParams), hipMemcpyHostToDevice));
    NN_HIP(hipMemcpy(ctx->dOpts, &ctx->hostOpts, sizeof(FlashOptions), hipMemcpyHostToDevice));
    *out = ctx;
    return NN_OK;
}

int nn_context_set_kernel_timing(nn_context *ctx, int32_t enable)
{
    if (!ctx)
        return fail(NN_E_INVALID_ARGUMENT, "null context");
    NN_HIP(hipDeviceSynchronize());
    releaseEvents(ctx);
    ctx->timing = enable != 0;
    return NN_OK;
}

int nn_context_kernel_time(nn_context *ctx, int32_t kind, double *total_ms, int64_t *launches)
{
    if (!ctx || kind < 0 || kind >= NN_KERNEL_KINDS || !total_ms || !launches)
        return fail(NN_E_INVALID_ARGUMENT, "invalid argument");
    double sum = 0.0;
    for (auto &pr : ctx->events[kind])
    {
        NN_HIP(hipEventSynchronize(pr.second));
        float ms = 0.0f;
        NN_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
        sum += ms;
    }
    *total_ms = sum;
    *launches = int64_t(ctx->events[kind].size());
    return NN_OK;
}

int nn_context_queue_counts(nn_context *ctx, int64_t *deferred, int64_t *recorded,
                            int64_t *reflashed)
{
    if (!ctx || !deferred || !recorded || !reflashed)
        return fail(NN_E_INVALID_ARGUMENT, "null argument");
    NN_HIP(hipDeviceSynchronize());
    WorkQueue q{};
    NN_HIP(hipMemcpy(&q, ctx->dQueue, sizeof q, hipMemcpyDeviceToHost));
    *deferred = q.deferred;
    *recorded = q.pending;
    *reflashed = q.reflash;
    return NN_OK;
}

int nn_context_destroy(nn_context *ctx)
{
    if (!ctx)
        return NN_OK;
    releaseEvents(ctx);
    if (ctx->dFluid) (void)hipFree(ctx->dFluid);
    if (ctx->dOpts) (void)hipFree(ctx->dOpts);
    if (ctx->dQueue) (void)hipFree(ctx->dQueue);
    if (ctx->dList) (void)hipFree(ctx->dList);
    if (ctx->dPending) (void)hipFree(ctx->dPending);
    if (ctx->dRecords) (void)hipFree(ctx->dRecords);
    releaseSort(ctx);
    delete ctx;
    return NN_OK;
}

int nn_eos_phase_batch(nn_context *ctx, int64_t n, const double *d_P, const double *d_T,
                       const double *d_x, const int32_t *d_role, double *d_Z, double *d_lnphi,
                       double *d_fug, int32_t *d_status, void *stream)
{
    if (!ctx || n < 0)
        return fail(NN_E_INVALID_ARGUMENT, "invalid context or size");
    if (n == 0)
        return NN_OK;
    if (!d_P || !d_T || !d_x || !d_role || !d_Z || !d_lnphi || !d_fug || !d_status)
        return fail(NN_E_INVALID_ARGUMENT, "null device pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    return dispatchNc(ctx->host.nc, [&](auto ncTag) {
        constexpr int NC = decltype(ncTag)::value;
        const unsigned grid = static_cast<unsigned>((n + kBlock - 1) / kBlock);
        LaunchTimer t(ctx, NN_KERNEL_EOS, st);
        NN_HIP((Launch<NC>::eos(grid, st, ctx->dFluid, n, d_P, d_T, d_x, d_role, d_Z, d_lnphi,
                                d_fug, d_status)));
        t.done();
        return NN_OK;
    });
}

static int launchFlash(nn_context *ctx, int64_t n, const double *d_P, const double *d_T,
                       const double *d_z, const uint8_t *d_allowed, const double *d_seed,
                       FlashHint hint, const nn_flash_out *out, void *stream)
{
    if (!ctx || n < 0 || !out)
        return fail(NN_E_INVALID_ARGUMENT, "invalid argument");
    if (n == 0)
        return NN_OK;
    if (!d_P || !d_T || !d_z || !out->presence || !out->converged || !out->status ||
        !out->iterations || !out->beta || !out->saturation || !out->x || !out->Z ||
        !out->rho_m || !out->k_go || !out->k_wo)
        return fail(NN_E_INVALID_ARGUMENT, "null device pointer");
    if (n > kMaxQueueItems)
        return fail(NN_E_INVALID_ARGUMENT, "batch too large for the 32-bit work queue");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const nn_flash_out o = *out;
    if (const int rc = ensureList(ctx, n); rc != NN_OK)
        return rc;
    const bool plain = ctx->sched.plain != 0;
    const bool ordered = !plain && ctx->sched.ordered && n >= kOrderMinItems &&
                         n <= kOrderMaxItems && !streamCapturing(st);
    if (ordered)
        if (const int rc = ensureSort(ctx, n); rc != NN_OK)
            return rc;
    return dispatchNc(ctx->host.nc, [&](auto ncTag) {
        constexpr int NC = decltype(ncTag)::value;
        resetQueueKernel<<<1, 1, 0, st>>>(ctx->dQueue, ctx->coopAll ? 1u : 0u,
                                          ctx->sched.wave_claim ? 1u : 0u);
        NN_HIP(hipGetLastError());
        const unsigned int *order = nullptr;
        if (ordered)
        {
            NN_HIP((Launch<NC>::flashKeys(st, n, d_P, d_T, d_z, hint.hint, ctx->dKeys[0],
                                          ctx->dIdx[0])));
            NN_HIP(sortOrder(ctx, n, st, &order));
        }
        LaunchTimer t(ctx, NN_KERNEL_FLASH, st);
        t.makeSplit();
        // the baseline form: one work-item per state (ceil(n / 64) blocks), index order
        const unsigned grid = plain ? static_cast<unsigned>((n + kBlock - 1) / kBlock)
                                    : persistentGrid(ctx, n);
        NN_HIP((Launch<NC>::flash(grid, coopGrid(ctx), st, ctx->dFluid, ctx->dOpts, n, d_P, d_T,
                                  d_z, d_allowed, d_seed, hint, o, ctx->dQueue, ctx->dList, order,
                                  plain, t.splitA, t.splitB)));
        t.doneSplit(NN_KERNEL_FLASH_COOP);
        return NN_OK;
    });
}

int nn_flash_batch(nn_context *ctx, int64_t n, const double *d_P, const double *d_T,
                   const double *d_z, const nn_flash_out *out, void *stream)
{
    return launchFlash(ctx, n, d_P, d_T, d_z, nullptr, nullptr, FlashHint{nullptr, nullptr}, out,
                       stream);
}

int nn_flash_seeded_batch(nn_context *ctx, int64_t n, const double *d_P, const double *d_T,
                          const double *d_z, const uint8_t *d_hint, const double *d_log_k,
                          const nn_flash_out *out, void *stream)
{
    if (!d_hint && n > 0)
        return fail(NN_E_INVALID_ARGUMENT, "hint array required");
    return launchFlash(ctx, n, d_P, d_T, d_z, nullptr, nullptr, FlashHint{d_hint, d_log_k}, out,
                       stream);
}

int nn_flash_batch_static(nn_context *ctx, int64_t n, const double *d_P, const double *d_T,
                          const double *d_z, const nn_flash_out *out, void *stream)
{
    if (!ctx || n < 0 || !out)
        return fail(NN_E_INVALID_ARGUMENT, "invalid argument");
    if (n > kStaticMaxStates)
        return fail(NN_E_INVALID_ARGUMENT,
                    "nn_flash_batch_static: batch exceeds the static scratch "
                    "(NN_STATIC_MAX_STATES = 1048576 states)");
    if (n == 0)
        return NN_OK;
    if (!d_P || !d_T || !d_z || !out->presence || !out->converged || !out->status ||
        !out->iterations || !out->beta || !out->saturation || !out->x || !out->Z ||
        !out->rho_m || !out->k_go || !out->k_wo)
        return fail(NN_E_INVALID_ARGUMENT, "null device pointer");
    char *base = static_cast<char *>(scratchPtr(ctx->device));
    if (!base)
        return fail(NN_E_HIP, "static scratch not resolved for this device");
    WorkQueue *q = reinterpret_cast<WorkQueue *>(base);
    unsigned int *list = reinterpret_cast<unsigned int *>(base + kStaticQueueBytes);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const nn_flash_out o = *out;
    const FlashHint none{nullptr, nullptr};
    return dispatchNc(ctx->host.nc, [&](auto ncTag) {
        constexpr int NC = decltype(ncTag)::value;
        // the queue's counters start at zero (a kernel, not a memset: capturable)
        resetQueueKernel<<<1, 1, 0, st>>>(q, ctx->coopAll ? 1u : 0u,
                                          ctx->sched.wave_claim ? 1u : 0u);
        NN_HIP(hipGetLastError());
        const bool plain = ctx->sched.plain != 0;
        const unsigned grid = plain ? static_cast<unsigned>((n + kBlock - 1) / kBlock)
                                    : persistentGrid(ctx, n);
        NN_HIP((Launch<NC>::flash(grid, coopGrid(ctx), st, ctx->dFluid, ctx->dOpts, n, d_P, d_T,
                                  d_z, nullptr, nullptr, none, o, q, list, nullptr, plain,
                                  nullptr, nullptr)));
        return NN_OK;
    });
}

int nn_context_get_schedule(nn_context *ctx, nn_schedule *s)
{
    if (!ctx || !s)
        return fail(NN_E_INVALID_ARGUMENT, "null argument");
    *s = ctx->sched;
    return NN_OK;
}

int nn_context_set_schedule(nn_context *ctx, const nn_schedule *s)
{
    if (!ctx || !s)
        return fail(NN_E_INVALID_ARGUMENT, "null argument");
    if (s->persistent_blocks_per_cu < 1 || s->persistent_blocks_per_cu > 64 ||
        s->coop_blocks_per_cu < 1 || s->coop_blocks_per_cu > 64)
        return fail(NN_E_INVALID_ARGUMENT, "blocks per CU must be in 1..64");
    ctx->sched = *s;
    ctx->sched.ordered = s->ordered ? 1 : 0;
    ctx->sched.wave_claim = s->wave_claim ? 1 : 0;
    ctx->sched.plain = s->plain ? 1 : 0;
    ctx->maxBlocks = 0;
    ctx->coopBlocks = 0;
    return NN_OK;
}

int nn_flash_autotune(nn_context *ctx, int64_t n, const double *d_P, const double *d_T,
                      const double *d_z, const nn_flash_out *out, int32_t repeats,
                      nn_schedule *best, double *best_ms, void *stream)
{
    if (!ctx || n <= 0 || !out || repeats < 1)
        return fail(NN_E_INVALID_ARGUMENT, "invalid argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const nn_schedule saved = ctx->sched;
    const int64_t savedMax = ctx->maxBlocks, savedCoop = ctx->coopBlocks;
    const bool savedTiming = ctx->timing;
    ctx->timing = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    NN_HIP(hipEventCreate(&e0));
    NN_HIP(hipEventCreate(&e1));
    nn_schedule winner = saved;
    double winnerMs = 1.0e300;
    int rc = NN_OK;
    // Candidates: processing order x claim granularity x persistent grid; the
    // cooperative grid only serves the few deferred states and stays at 4/CU.
    const int grids[4] = {4, 8, 16, 32};
    for (int ordered = 1; ordered >= 0 && rc == NN_OK; --ordered)
        for (int claim = 0; claim <= 1 && rc == NN_OK; ++claim)
            for (int g = 0; g < 4 && rc == NN_OK; ++g)
            {
                const nn_schedule cand{grids[g], 4, ordered, claim, 0};
                ctx->sched = cand;
                ctx->maxBlocks = ctx->coopBlocks = 0;
                double bestRun = 1.0e300;
                for (int r = 0; r < repeats && rc == NN_OK; ++r)
                {
                    if (hipEventRecord(e0, st) != hipSuccess)
                    {
                        rc = hipFail(hipGetLastError(), "hipEventRecord");
                        break;
                    }
                    rc = launchFlash(ctx, n, d_P, d_T, d_z, nullptr, nullptr,
                                     FlashHint{nullptr, nullptr}, out, stream);
                    if (rc != NN_OK)
                        break;
                    float ms = 0.0f;
                    if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                    {
                        rc = hipFail(hipGetLastError(), "autotune timing");
                        break;
                    }
                    bestRun = std::min(bestRun, double(ms));
                }
                if (rc == NN_OK && bestRun < winnerMs)
                {
                    winnerMs = bestRun;
                    winner = cand;
                }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->timing = savedTiming;
    if (rc != NN_OK)
    {
        ctx->sched = saved;
        ctx->maxBlocks = savedMax;
        ctx->coopBlocks = savedCoop;
        return rc;
    }
    ctx->sched = winner;
    ctx->maxBlocks = ctx->coopBlocks = 0;
    if (best)
        *best = winner;
    if (best_ms)
        *best_ms = winnerMs;
    return NN_OK;
}

int nn_flash_restricted_batch(nn_context *ctx, int64_t n, const double *d_P, const double *d_T,
                              const double *d_z, const uint8_t *d_allowed, const double *d_seed,
                              const nn_flash_out *out, void *stream)
{
    if (!d_allowed)
        return fail(NN_E_INVALID_ARGUMENT, "allowed masks required");
    return launchFlash(ctx, n, d_P, d_T, d_z, d_allowed, d_seed, FlashHint{nullptr, nullptr}, out,
                       stream);
}

int nn_stability_batch(nn_context *ctx, int64_t n, const double *d_P, const double *d_T,
                       const double *d_z, const uint8_t *d_active, const double *d_x,
                       uint8_t *d_stable, uint8_t *d_valid, uint8_t *d_unstable_mask,
                       double *d_trial_sum, double *d_incipient, int32_t *d_status, void *stream)
{
    if (!ctx || n < 0)
        return fail(NN_E_INVALID_ARGUMENT, "invalid argument");
    if (n == 0)
        return NN_OK;
    if (!d_P || !d_T || !d_z || !d_active || !d_x || !d_stable || !d_valid || !d_unstable_mask ||
        !d_trial_sum || !d_incipient || !d_status)
        return fail(NN_E_INVALID_ARGUMENT, "null device pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    return dispatchNc(ctx->host.nc, [&](auto ncTag) {
        constexpr int NC = decltype(ncTag)::value;
        const unsigned grid = static_cast<unsigned>((n + kBlock - 1) / kBlock);
        LaunchTimer t(ctx, NN_KERNEL_STABILITY, st);
        NN_HIP((Launch<NC>::stability(grid, st, ctx->dFluid, ctx->dOpts, n, d_P, d_T, d_z, d_active,
                                      d_x, d_stable, d_valid, d_unstable_mask, d_trial_sum,
                                      d_incipient, d_status)));
        t.done();
        return NN_OK;
    });
}

int nn_phase_update_batch(nn_context *ctx, int64_t n, int32_t has_well_unknown, const double *d_T,
                          double *d_primary, double *d_phase_state, int32_t *d_status,
                          uint8_t *d_flags, void *stream)
{
    if (!ctx || n < 0)
        return fail(NN_E_INVALID_ARGUMENT, "invalid argument");
    if (n == 0)
        return NN_OK;
    if (!d_T || !d_primary || !d_phase_state || !d_status)
        return fail(NN_E_INVALID_ARGUMENT, "null device pointer");
    if (ctx->host.water < 0)
        return fail(NN_E_INVALID_ARGUMENT, "three-phase update needs a water component");
    if (n > kMaxQueueItems)
        return fail(NN_E_INVALID_ARGUMENT, "batch too large for the 32-bit work queue");
    if (const int rc = ensureList(ctx, n); rc != NN_OK)
        return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    // The plain form (nn_schedule.plain): one work-item per cell in index order
    // (ceil(n / 64) blocks), no condition sort and no lockstep stability trials;
    // cells past the lane budget still go to the cooperative pass.
    const bool plain = ctx->sched.plain != 0;
    const bool lockstep = ctx->lockstep && !plain;
    const bool ordered = !plain && ctx->sched.ordered && n >= kOrderMinItems &&
                         n <= kOrderMaxItems && !streamCapturing(st);
    if (ordered)
        if (const int rc = ensureSort(ctx, n); rc != NN_OK)
            return rc;
    return dispatchNc(ctx->host.nc, [&](auto ncTag) {
        constexpr int NC = decltype(ncTag)::value;
        resetQueueKernel<<<1, 1, 0, st>>>(ctx->dQueue, ctx->coopAll ? 1u : 0u,
                                          !plain && ctx->sched.wave_claim ? 1u : 0u);
        NN_HIP(hipGetLastError());
        const unsigned int *order = nullptr;
        if (ordered)
        {
            NN_HIP((Launch<NC>::phaseUpdateKeys(has_well_unknown != 0, st, n, d_T, d_primary,
                                                d_phase_state, ctx->dKeys[0], ctx->dIdx[0])));
            NN_HIP(sortOrder(ctx, n, st, &order));
        }
        if (lockstep)
            if (const int rc = ensureRecords(ctx, n, sizeof(StabRecord<NC>)); rc != NN_OK)
                return rc;
        // an ordered batch also sorts its re-flash pass (the sort buffers are free
        // again once the first pass has run)
        ReflashSortArg sortArg{ctx, n};
        const ReflashOrder reorder{ctx->dKeys[0], ctx->dIdx[0], reflashSort, &sortArg};
        LaunchTimer t(ctx, NN_KERNEL_PHASE_UPDATE, st);
        t.makeSplit();
        const unsigned grid = plain ? static_cast<unsigned>((n + kBlock - 1) / kBlock)
                                    : persistentGrid(ctx, n);
        NN_HIP((Launch<NC>::phaseUpdate(has_well_unknown != 0, grid,
                                        coopGrid(ctx), st, ctx->dFluid, ctx->dOpts, n, d_T,
                                        d_primary, d_phase_state, d_status, d_flags, ctx->dQueue,
                                        ctx->dList, order, lockstep ? ctx->dPending : nullptr,
                                        lockstep ? ctx->dRecords : nullptr,
                                        ordered && lockstep ? &reorder : nullptr, t.splitA,
                                        t.splitB)));
        t.doneSplit(NN_KERNEL_PHASE_UPDATE_COOP);
        return NN_OK;
    });
}

int nn_phase_update_transaction(nn_context *ctx, int64_t n, int32_t has_well_unknown,
                                const double *d_T, double *d_primary, double *d_phase_state,
                                double *d_work_primary, double *d_work_phase_state,
                                int32_t *d_status, uint8_t *d_flags, nn_update_report *d_report,
                                void *stream)
{
    if (!ctx || n < 0)
        return fail(NN_E_INVALID_ARGUMENT, "invalid argument");
    if (!d_T || !d_primary || !d_phase_state || !d_work_primary || !d_work_phase_state ||
        !d_status || !d_flags || !d_report)
        return fail(NN_E_INVALID_ARGUMENT, "null device pointer");
    if (d_work_primary == d_primary || d_work_phase_state == d_phase_state)
        return fail(NN_E_INVALID_ARGUMENT, "the workspace must not alias the committed vectors");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int nc = ctx->host.nc;
    const int64_t np = 3 * nc + 1 + (has_well_unknown ? 1 : 0);
    const int64_t nps = 3 * nc + 8;
    txnInitKernel<<<1, 1, 0, st>>>(d_report);
    NN_HIP(hipGetLastError());
    if (n == 0)
    {
        txnReportKernel<<<1, 1, 0, st>>>(0, d_status, d_flags, d_report);
        NN_HIP(hipGetLastError());
        return NN_OK;
    }
    NN_HIP(hipMemcpyAsync(d_work_primary, d_primary, sizeof(double) * size_t(n * np),
                          hipMemcpyDeviceToDevice, st));
    NN_HIP(hipMemcpyAsync(d_work_phase_state, d_phase_state, sizeof(double) * size_t(n * nps),
                          hipMemcpyDeviceToDevice, st));
    if (const int rc = nn_phase_update_batch(ctx, n, has_well_unknown, d_T, d_work_primary,
                                             d_work_phase_state, d_status, d_flags, stream);
        rc != NN_OK)
        return rc;
    // report fields as 64-bit scratch of the reduction: first failing index, count
    auto *first = reinterpret_cast<unsigned long long *>(&d_report->first_failure_cell);
    auto *stable = reinterpret_cast<unsigned long long *>(&d_report->stable_reduced_cells);
    txnSetKernel<<<1, 1, 0, st>>>(first, static_cast<unsigned long long>(n));
    NN_HIP(hipGetLastError());
    const unsigned grid = static_cast<unsigned>((n + 255) / 256);
    txnFirstFailureKernel<<<grid, 256, 0, st>>>(n, d_status, first);
    NN_HIP(hipGetLastError());
    txnCountKernel<<<grid, 256, 0, st>>>(n, d_flags, first, stable);
    NN_HIP(hipGetLastError());
    txnReportKernel<<<1, 1, 0, st>>>(n, d_status, d_flags, d_report);
    NN_HIP(hipGetLastError());
    const auto commitGrid = [](int64_t count) {
        return static_cast<unsigned>(std::min<int64_t>((count + 255) / 256, 4096));
    };
    txnCommitKernel<<<commitGrid(n * np), 256, 0, st>>>(n * np, d_report, d_work_primary,
                                                        d_primary);
    NN_HIP(hipGetLastError());
    txnCommitKernel<<<commitGrid(n * nps), 256, 0, st>>>(n * nps, d_report, d_work_phase_state,
                                                         d_phase_state);
    NN_HIP(hipGetLastError());
    return NN_OK;
}

// ------------------------------------------------------------------ assembly
int nn_structured_topology(int32_t nx, int32_t ny, int32_t nz, const double *dx, const double *dy,
                           const double *dz, const double *perm, int64_t *conn_offsets,
                           int64_t *conn_neighbor, double *conn_trans, double *conn_gdz,
                           int64_t *row_offsets, int64_t *col_blocks, int64_t *nconn,
                           int64_t *nnzb)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || !dx || !dy || !dz || !perm || !nconn || !nnzb)
        return fail(NN_E_INVALID_ARGUMENT, "invalid grid description");
    const int64_t ncell = int64_t(nx) * ny * nz;
    // Vertical centres accumulate dz (structured_cartesian_geometry.hpp:60-71).
    std::vector<double> zc(nz);
    double pos = 0.0;
    for (int k = 0; k < nz; ++k)
    {
        zc[k] = pos + 0.5 * dz[k];
        pos += dz[k];
    }
    auto halfTrans = [&](int i, int j, int k, int dir) {
        const int axis = dir / 2;
        const double w[3] = {dx[i], dy[j], dz[k]};
        const double kperm = perm[((int64_t(k) * ny + j) * nx + i) * 3 + axis];
        const double dist = 0.5 * w[axis];
        double normal = 0.0;
        switch (dir)
        {
        case 0: normal = -w[1] * w[2]; break;
        case 1: normal = w[1] * w[2]; break;
        case 2: normal = -w[0] * w[2]; break;
        case 3: normal = w[0] * w[2]; break;
        case 4: normal = -w[0] * w[1]; break;
        default: normal = w[0] * w[1]; break;
        }
        return kperm * std::fabs(normal) / dist;
    };
    int64_t nc_ = 0, nb_ = 0;
    const int di[6] = {-1, 1, 0, 0, 0, 0}, dj[6] = {0, 0, -1, 1, 0, 0}, dk[6] = {0, 0, 0, 0, -1, 1};
    std::vector<int64_t> cols;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
            {
                const int64_t cell = (int64_t(k) * ny + j) * nx + i;
                if (conn_offsets) conn_offsets[cell] = nc_;
                cols.clear();
                cols.push_back(cell);
                for (int d = 0; d < 6; ++d)
                {
                    const int ni = i + di[d], nj = j + dj[d], nk = k + dk[d];
                    if (ni < 0 || nj < 0 || nk < 0 || ni >= nx || nj >= ny || nk >= nz)
                        continue;
                    const int64_t nbr = (int64_t(nk) * ny + nj) * nx + ni;
                    // grid/include/structuredgrid/structuredgrid.hpp:642-667
                    const double h1 = halfTrans(i, j, k, d);
                    const double h2 = halfTrans(ni, nj, nk, d ^ 1);
                    const double t = (h1 <= 1.0e-30 || h2 <= 1.0e-30) ? 0.0 : 1.0 / (1.0 / h1 + 1.0 / h2);
                    // structuredgrid.hpp:611-624: g * (z_i - z_j) on Z faces only
                    const double g = (d >= 4) ? (zc[k] - zc[nk]) * 9.80665 : 0.0;
                    if (conn_neighbor)
                    {
                        conn_neighbor[nc_] = nbr;
                        conn_trans[nc_] = t;
                        conn_gdz[nc_] = g;
                    }
                    ++nc_;
                    cols.push_back(nbr);
                }
                std::sort(cols.begin(), cols.end());
                if (row_offsets) row_offsets[cell] = nb_;
                if (col_blocks)
                    for (size_t q = 0; q < cols.size(); ++q) col_blocks[nb_ + q] = cols[q];
                nb_ += int64_t(cols.size());
            }
    if (conn_offsets) conn_offsets[ncell] = nc_;
    if (row_offsets) row_offsets[ncell] = nb_;
    *nconn = nc_;
    *nnzb = nb_;
    return NN_OK;
}

int nn_bsr_cell_columns(int64_t ncell, int32_t np, const int64_t *conn_offsets,
                        const int64_t *conn_neighbor, const int64_t *row_offsets,
                        const int64_t *col_blocks, const double *bsr, double *values,
                        int64_t *rows, int64_t *column_offsets)
{
    if (ncell < 0 || np <= 0 || !conn_offsets || !conn_neighbor || !row_offsets || !col_blocks ||
        !bsr || !values || !rows || !column_offsets)
        return fail(NN_E_INVALID_ARGUMENT, "invalid argument");
    const size_t bs2 = size_t(np) * size_t(np);
    auto slotOf = [&](int64_t row, int64_t col) -> int64_t {
        const int64_t *b = col_blocks + row_offsets[row], *e = col_blocks + row_offsets[row + 1];
        const int64_t *it = std::lower_bound(b, e, col);
        return (it == e || *it != col) ? -1 : int64_t(it - col_blocks);
    };
    int64_t out = 0;
    for (int64_t c = 0; c < ncell; ++c)
    {
        column_offsets[c] = out;
        // natural_petsc_runtime_jacobian.inc:279-299: rows [self, neighbours in
        // connection order], one column block (the cell), blocks row-major
        const int64_t count = 1 + conn_offsets[c + 1] - conn_offsets[c];
        for (int64_t q = 0; q < count; ++q)
        {
            const int64_t row = q == 0 ? c : conn_neighbor[conn_offsets[c] + q - 1];
            if (row < 0 || row >= ncell)
                return fail(NN_E_INVALID_ARGUMENT, "connection to a cell outside the grid");
            const int64_t slot = slotOf(row, c);
            if (slot < 0)
                return fail(NN_E_INVALID_ARGUMENT, "pattern lacks a block of the cell's column");
            rows[out] = row;
            std::memcpy(values + size_t(out) * bs2, bsr + size_t(slot) * bs2, bs2 * sizeof(double));
            ++out;
        }
    }
    column_offsets[ncell] = out;
    return NN_OK;
}

int nn_pattern_add_wells(int64_t ncell, const int64_t *row_offsets, const int64_t *col_blocks,
                         int32_t nwells, const nn_well *wells, const int64_t *perforation_cells,
                         int64_t *out_row_offsets, int64_t *out_col_blocks, int64_t *out_nnzb)
{
    if (ncell < 0 || nwells < 0 || !row_offsets || !col_blocks || !out_nnzb ||
        (nwells > 0 && (!wells || !perforation_cells)) ||
        (out_col_blocks && !out_row_offsets))
        return fail(NN_E_INVALID_ARGUMENT, "invalid argument");
    // jacobian_factory.hpp:123-154: bhp row <- perforation blocks; perforation row <- bhp block
    std::vector<std::vector<int64_t>> extra(static_cast<size_t>(ncell));
    for (int w = 0; w < nwells; ++w)
    {
        const nn_well &wl = wells[w];
        if (wl.bhp_cell < 0 || wl.bhp_cell >= ncell || wl.perforation_begin < 0 ||
            wl.perforation_end < wl.perforation_begin)
            return fail(NN_E_INVALID_ARGUMENT, "invalid well");
        for (int64_t k = wl.perforation_begin; k < wl.perforation_end; ++k)
        {
            const int64_t c = perforation_cells[k];
            if (c < 0 || c >= ncell)
                return fail(NN_E_INVALID_ARGUMENT, "perforation cell outside the grid");
            extra[static_cast<size_t>(wl.bhp_cell)].push_back(c);
            extra[static_cast<size_t>(c)].push_back(wl.bhp_cell);
        }
    }
    int64_t total = 0;
    std::vector<int64_t> row;
    for (int64_t r = 0; r < ncell; ++r)
    {
        row.assign(col_blocks + row_offsets[r], col_blocks + row_offsets[r + 1]);
        row.insert(row.end(), extra[static_cast<size_t>(r)].begin(), extra[static_cast<size_t>(r)].end());
        std::sort(row.begin(), row.end());
        row.erase(std::unique(row.begin(), row.end()), row.end());
        if (out_row_offsets)
            out_row_offsets[r] = total;
        if (out_col_blocks)
            std::copy(row.begin(), row.end(), out_col_blocks + total);
        total += int64_t(row.size());
    }
    if (out_row_offsets)
        out_row_offsets[ncell] = total;
    *out_nnzb = total;
    return NN_OK;
}

int nn_state_sample(const double *d_vec, int64_t ncell, int32_t dof_per_cell,
                    const int64_t *d_cells, int64_t nsel, const int32_t *d_vars, int32_t nvar,
                    double *d_out, int32_t *d_bad, void *stream)
{
    if (!d_vec || ncell < 0 || dof_per_cell <= 0 || nsel < 0 || nvar < 0 || !d_bad ||
        (nsel > 0 && nvar > 0 && (!d_cells || !d_vars || !d_out)))
        return fail(NN_E_INVALID_ARGUMENT, "invalid argument");
    // cpgrid_output_access.hpp:78-84: the sample travels as one MPI int count
    if (nsel > 0 && nvar > 0 && nsel > int64_t(0x7FFFFFFF) / nvar)
        return fail(NN_E_INVALID_ARGUMENT, "output sample larger than INT_MAX values");
    hipStream_t st = static_cast<hipStream_t>(stream);
    // the flag is cleared by a kernel, not a memset, so the call can be captured
    clearFlagKernel<<<1, 1, 0, st>>>(d_bad);
    NN_HIP(hipGetLastError());
    const int64_t count = nsel * nvar;
    if (count == 0)
        return NN_OK;
    const unsigned grid = static_cast<unsigned>((count + 255) / 256);
    stateSampleKernel<<<grid, 256, 0, st>>>(d_vec, ncell, dof_per_cell, d_cells, nsel, d_vars,
                                            nvar, d_out, d_bad);
    NN_HIP(hipGetLastError());
    return NN_OK;
}

// std::ostream << double with setprecision(16) and the default float field
// (format.hpp:32-42) prints as printf's %.16g.
static int64_t csvAppend(char *buf, int64_t cap, int64_t at, const char *text, int64_t len)
{
    if (buf && at + len < cap)
        std::memcpy(buf + at, text, size_t(len));
    return at + len;
}

int nn_csv_sample_header(int64_t count, char *buf, int64_t cap, int64_t *length)
{
    if (count < 0 || !length || (buf && cap <= 0))
        return fail(NN_E_INVALID_ARGUMENT, "invalid argument");
    // format.hpp:19-29: value_0,...,value_{n-1}
    int64_t at = 0;
    char item[32];
    for (int64_t i = 0; i < count; ++i)
    {
        const int len = std::snprintf(item, sizeof item, "%svalue_%lld", i ? "," : "",
                                      static_cast<long long>(i));
        at = csvAppend(buf, cap, at, item, len);
    }
    at = csvAppend(buf, cap, at, "\n", 1);
    *length = at;
    if (buf)
    {
        if (at >= cap)
            return fail(NN_E_INVALID_ARGUMENT, "buffer too small");
        buf[at] = '\0';
    }
    return NN_OK;
}

int nn_csv_sample_row(const double *values, int64_t count, char *buf, int64_t cap,
                      int64_t *length)
{
    if (count < 0 || (count > 0 && !values) || !length || (buf && cap <= 0))
        return fail(NN_E_INVALID_ARGUMENT, "invalid argument");
    int64_t at = 0;
    char item[40];
    for (int64_t i = 0; i < count; ++i)
    {
        const int len = std::snprintf(item, sizeof item, "%s%.16g", i ? "," : "", values[i]);
        at = csvAppend(buf, cap, at, item, len);
    }
    at = csvAppend(buf, cap, at, "\n", 1);
    *length = at;
    if (buf)
    {
        if (at >= cap)
            return fail(NN_E_INVALID_ARGUMENT, "buffer too small");
        buf[at] = '\0';
    }
    return NN_OK;
}

int nn_csv_parse_row(const char *line, double *values, int64_t cap, int64_t *count)
{
    if (!line || !count || (cap > 0 && !values))
        return fail(NN_E_INVALID_ARGUMENT, "invalid argument");
    int64_t n = 0;
    const char *p = line;
    if (*p == '\0' || *p == '\n')
    {
        *count = 0;
        return NN_OK;
    }
    for (;;)
    {
        char *end = nullptr;
        const double v = std::strtod(p, &end);
        if (end == p)
            return fail(NN_E_INVALID_ARGUMENT, "not a number in a sample row");
        if (n < cap)
            values[n] = v;
        ++n;
        p = end;
        if (*p == ',')
        {
            ++p;
            continue;
        }
        if (*p == '\n' || *p == '\0' || (*p == '\r' && (p[1] == '\n' || p[1] == '\0')))
            break;
        return fail(NN_E_INVALID_ARGUMENT, "malformed sample row");
    }
    *count = n;
    if (n > cap)
        return fail(NN_E_INVALID_ARGUMENT, "more values than the output holds");
    return NN_OK;
}

} // extern "C"

struct nn_assembly
{
    nn_context *ctx = nullptr;
    int hasWell = 0;
    AssemblyView view{};
    int64_t nnzb = 0;
    int64_t *dSelf = nullptr;
    int64_t *dRecip = nullptr;
    double *dRecord = nullptr;
    double *dRoots = nullptr;
    const int64_t *rowOffsetsDevice = nullptr;   // the caller's pattern (kept by reference)
    const int64_t *colBlocksDevice = nullptr;
    // pattern blocks no face or diagonal term writes (the wells' couplings)
    int64_t *dZeroSlots = nullptr;
    int64_t nZero = 0;
    // wells (nn_assembly_set_wells)
    int nWells = 0;
    int64_t nWellCells = 0;
    int64_t *dWellCellOffsets = nullptr;
    int32_t *dWellCellPerfs = nullptr;
    int32_t *dPerfWell = nullptr;
    double *dPerfIndex = nullptr;
    int64_t *dPerfBhpSlot = nullptr;
    DeviceWell *dWells = nullptr;
    int64_t *dWellCells = nullptr;
};

static void releaseWells(nn_assembly *a)
{
    for (void *p : {static_cast<void *>(a->dWellCellOffsets), static_cast<void *>(a->dWellCellPerfs),
                    static_cast<void *>(a->dPerfWell), static_cast<void *>(a->dPerfIndex),
                    static_cast<void *>(a->dPerfBhpSlot), static_cast<void *>(a->dWells),
                    static_cast<void *>(a->dWellCells)})
        if (p) (void)hipFree(p);
    a->dWellCellOffsets = nullptr;
    a->dWellCellPerfs = nullptr;
    a->dPerfWell = nullptr;
    a->dPerfIndex = nullptr;
    a->dPerfBhpSlot = nullptr;
    a->dWells = nullptr;
    a->dWellCells = nullptr;
    a->nWells = 0;
    a->nWellCells = 0;
    a->view.wellCellOffsets = nullptr;
    a->view.wellCellPerfs = nullptr;
    a->view.perfWell = nullptr;
    a->view.perfIndex = nullptr;
    a->view.perfBhpSlot = nullptr;
    a->view.wells = nullptr;
}

extern "C" {

int nn_assembly_info(const nn_assembly *a, nn_context **ctx, int64_t *ncell, int64_t *nnzb,
                     int32_t *np)
{
    if (!a)
        return fail(NN_E_INVALID_ARGUMENT, "null assembly");
    if (ctx) *ctx = a->ctx;
    if (ncell) *ncell = a->view.ncell;
    if (nnzb) *nnzb = a->nnzb;
    if (np) *np = 3 * a->ctx->host.nc + 1 + (a->hasWell ? 1 : 0);
    return NN_OK;
}

int nn_assembly_destroy(nn_assembly *a)
{
    if (!a)
        return NN_OK;
    if (a->dSelf) (void)hipFree(a->dSelf);
    if (a->dRecip) (void)hipFree(a->dRecip);
    if (a->dRecord) (void)hipFree(a->dRecord);
    if (a->dRoots) (void)hipFree(a->dRoots);
    if (a->dZeroSlots) (void)hipFree(a->dZeroSlots);
    releaseWells(a);
    delete a;
    return NN_OK;
}

int nn_assembly_create(nn_context *ctx, int64_t ncell, int32_t has_well_unknown,
                       const int64_t *d_conn_offsets, const int64_t *d_conn_neighbor,
                       const double *d_conn_trans, const double *d_conn_gdz,
                       const int64_t *d_row_offsets, const int64_t *d_col_blocks, int64_t nnzb,
                       const double *d_volume, const double *d_porosity, const double *d_T,
                       nn_assembly **out)
{
    if (!ctx || !out || ncell <= 0 || !d_conn_offsets || !d_conn_neighbor || !d_conn_trans ||
        !d_conn_gdz || !d_row_offsets || !d_col_blocks || !d_volume || !d_porosity || !d_T)
        return fail(NN_E_INVALID_ARGUMENT, "invalid assembly description");
    *out = nullptr;
    if (ctx->host.water < 0)
        return fail(NN_E_INVALID_ARGUMENT, "Natural assembly needs a water component");
    if (ctx->host.nc > kAssemblyMaxNc)
        return fail(NN_E_UNSUPPORTED, "Natural assembly is built for at most NN_MAX_NC components");
    // Slots of the diagonal and reciprocal blocks in the cached pattern (setup only).
    std::vector<int64_t> co(ncell + 1), ro(ncell + 1);
    NN_HIP(hipMemcpy(co.data(), d_conn_offsets, sizeof(int64_t) * (ncell + 1), hipMemcpyDeviceToHost));
    NN_HIP(hipMemcpy(ro.data(), d_row_offsets, sizeof(int64_t) * (ncell + 1), hipMemcpyDeviceToHost));
    const int64_t nconn = co[ncell];
    if (ro[ncell] != nnzb || nconn < 0)
        return fail(NN_E_INVALID_ARGUMENT, "pattern size mismatch");
    std::vector<int64_t> nbr(nconn), cols(nnzb);
    NN_HIP(hipMemcpy(nbr.data(), d_conn_neighbor, sizeof(int64_t) * nconn, hipMemcpyDeviceToHost));
    NN_HIP(hipMemcpy(cols.data(), d_col_blocks, sizeof(int64_t) * nnzb, hipMemcpyDeviceToHost));
    auto slot = [&](int64_t row, int64_t col) -> int64_t {
        if (row < 0 || row >= ncell) return -1;
        const int64_t *b = cols.data() + ro[row], *e = cols.data() + ro[row + 1];
        const int64_t *it = std::lower_bound(b, e, col);
        return (it == e || *it != col) ? -1 : int64_t(it - cols.data());
    };
    std::vector<int64_t> self(ncell), recip(nconn);
    std::vector<uint8_t> covered(nnzb, 0);
    for (int64_t c = 0; c < ncell; ++c)
    {
        self[c] = slot(c, c);
        if (self[c] < 0)
            return fail(NN_E_INVALID_ARGUMENT, "pattern lacks a diagonal block");
        covered[self[c]]++;
        for (int64_t k = co[c]; k < co[c + 1]; ++k)
        {
            recip[k] = slot(nbr[k], c);
            if (recip[k] < 0)
                return fail(NN_E_INVALID_ARGUMENT, "pattern lacks a reciprocal block");
            covered[recip[k]]++;
        }
    }
    // a block written by two terms would be overwritten; a block no term writes is
    // a well coupling (nn_pattern_add_wells), zeroed before each assembly
    std::vector<int64_t> zeroSlots;
    for (int64_t b = 0; b < nnzb; ++b)
    {
        if (covered[b] > 1)
            return fail(NN_E_INVALID_ARGUMENT,
                        "a pattern block is both a diagonal and a face block, or two face blocks");
        if (covered[b] == 0)
            zeroSlots.push_back(b);
    }
    nn_assembly *a = new (std::nothrow) nn_assembly;
    if (!a)
        return fail(NN_E_OUT_OF_MEMORY, "host allocation failed");
    a->ctx = ctx;
    a->hasWell = has_well_unknown ? 1 : 0;
    a->nnzb = nnzb;
    a->rowOffsetsDevice = d_row_offsets;
    a->colBlocksDevice = d_col_blocks;
    const int nc = ctx->host.nc;
    const size_t recordBytes = sizeof(double) * size_t(10 + 3 * nc) * size_t(ncell);
    const bool cpa = ctx->host.model == NN_MODEL_CPA;
    if (hipMalloc(&a->dSelf, sizeof(int64_t) * ncell) != hipSuccess ||
        hipMalloc(&a->dRecip, sizeof(int64_t) * std::max<int64_t>(1, nconn)) != hipSuccess ||
        hipMalloc(&a->dRecord, recordBytes) != hipSuccess ||
        (cpa && hipMalloc(&a->dRoots, sizeof(double) * 9 * size_t(ncell)) != hipSuccess))
    {
        nn_assembly_destroy(a);
        return fail(NN_E_OUT_OF_MEMORY, "device allocation failed");
    }
    if (!zeroSlots.empty())
    {
        if (hipMalloc(&a->dZeroSlots, sizeof(int64_t) * zeroSlots.size()) != hipSuccess)
        {
            nn_assembly_destroy(a);
            return fail(NN_E_OUT_OF_MEMORY, "device allocation failed");
        }
        NN_HIP(hipMemcpy(a->dZeroSlots, zeroSlots.data(), sizeof(int64_t) * zeroSlots.size(),
                         hipMemcpyHostToDevice));
        a->nZero = int64_t(zeroSlots.size());
    }
    NN_HIP(hipMemcpy(a->dSelf, self.data(), sizeof(int64_t) * ncell, hipMemcpyHostToDevice));
    if (nconn > 0)
        NN_HIP(hipMemcpy(a->dRecip, recip.data(), sizeof(int64_t) * nconn, hipMemcpyHostToDevice));
    a->view = AssemblyView{ncell, d_conn_offsets, d_conn_neighbor, d_conn_trans, d_conn_gdz,
                           d_volume, d_porosity, d_T, a->dSelf, a->dRecip, a->dRecord,
                           a->dRoots, nullptr, {}, {}};
    for (int k = 0; k < kMaxNp; ++k)
        a->view.rowScale[k] = a->view.varScale[k] = 1.0;
    *out = a;
    return NN_OK;
}

void nn_scaling_default(nn_scaling *s)
{
    if (!s)
        return;
    s->enabled = 0;
    s->pressure_scale = 1.0e7;
    s->composition_scale = 1.0;
    s->saturation_scale = 1.0;
    s->mass_residual_scale = 1.0;
    s->fugacity_residual_scale = 1.0;
    s->closure_residual_scale = 1.0;
    s->rate_well_residual_floor = 1.0;
}

int nn_assembly_set_wells(nn_assembly *a, int32_t nwells, const nn_well *wells,
                          int64_t nperforations, const int64_t *perforation_cells,
                          const double *perforation_well_index, const double *surface_density)
{
    if (!a || nwells < 0 || nperforations < 0)
        return fail(NN_E_INVALID_ARGUMENT, "invalid argument");
    if (nwells == 0)
    {
        releaseWells(a);
        return NN_OK;
    }
    if (!wells || !surface_density || (nperforations > 0 && (!perforation_cells || !perforation_well_index)))
        return fail(NN_E_INVALID_ARGUMENT, "null argument");
    if (!a->hasWell)
        return fail(NN_E_INVALID_ARGUMENT, "wells need a well unknown per cell");
    if (nperforations > int64_t(0x7FFFFFFF))
        return fail(NN_E_INVALID_ARGUMENT, "too many perforations");
    const int64_t ncell = a->view.ncell;
    const int nc = a->ctx->host.nc;
    // well_source.hpp:69-83 and the specification's validation
    for (int p = 0; p < 3; ++p)
        if (!(surface_density[p] > 0.0) || !std::isfinite(surface_density[p]))
            return fail(NN_E_INVALID_ARGUMENT, "surface density must be positive");
    for (int64_t k = 0; k < nperforations; ++k)
    {
        if (perforation_cells[k] < 0 || perforation_cells[k] >= ncell)
            return fail(NN_E_INVALID_ARGUMENT, "perforation cell outside the grid");
        if (!(perforation_well_index[k] >= 0.0) || !std::isfinite(perforation_well_index[k]))
            return fail(NN_E_INVALID_ARGUMENT, "well index must be finite and nonnegative");
    }
    std::vector<DeviceWell> dw(static_cast<size_t>(nwells));
    std::vector<int32_t> perfWell(static_cast<size_t>(nperforations), -1);
    for (int w = 0; w < nwells; ++w)
    {
        const nn_well &wl = wells[w];
        if (wl.type != NN_WELL_INJECTOR && wl.type != NN_WELL_PRODUCER)
            return fail(NN_E_INVALID_ARGUMENT, "unknown well type");
        if (wl.control != NN_WELL_CONTROL_BHP)
            return fail(NN_E_UNSUPPORTED, "only bottom-hole-pressure control is built");
        if (wl.bhp_cell < 0 || wl.bhp_cell >= ncell || !std::isfinite(wl.target))
            return fail(NN_E_INVALID_ARGUMENT, "invalid bhp cell or target");
        if (wl.perforation_begin < 0 || wl.perforation_end < wl.perforation_begin ||
            wl.perforation_end > nperforations)
            return fail(NN_E_INVALID_ARGUMENT, "invalid perforation range");
        DeviceWell &d = dw[static_cast<size_t>(w)];
        d.type = wl.type;
        d.bhpCell = wl.bhp_cell;
        d.target = wl.target;
        for (int p = 0; p < 3; ++p) d.injPhase[p] = wl.injection_phase_fraction[p];
        for (int i = 0; i < kMaxNc; ++i)
            d.injMass[i] = i < nc ? wl.injection_component_mass_fraction[i] : 0.0;
        for (int64_t k = wl.perforation_begin; k < wl.perforation_end; ++k)
        {
            if (perfWell[static_cast<size_t>(k)] >= 0)
                return fail(NN_E_INVALID_ARGUMENT, "a perforation belongs to two wells");
            perfWell[static_cast<size_t>(k)] = w;
        }
    }
    // per-cell perforation lists in well, then perforation order (the order in which
    // natural_petsc_runtime_wells.inc:93-165 sums a cell's sources)
    std::vector<int64_t> offsets(static_cast<size_t>(ncell) + 1, 0), bhpSlot(
        static_cast<size_t>(nperforations), -1);
    std::vector<int32_t> order;
    order.reserve(static_cast<size_t>(nperforations));
    for (int w = 0; w < nwells; ++w)
        for (int64_t k = wells[w].perforation_begin; k < wells[w].perforation_end; ++k)
            order.push_back(int32_t(k));
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        return perforation_cells[x] < perforation_cells[y];
    });
    std::vector<int64_t> cellsWithWells;
    for (int32_t k : order)
    {
        const int64_t c = perforation_cells[k];
        ++offsets[static_cast<size_t>(c) + 1];
        if (cellsWithWells.empty() || cellsWithWells.back() != c)
            cellsWithWells.push_back(c);
    }
    for (int64_t c = 0; c < ncell; ++c) offsets[c + 1] += offsets[c];
    // the block (perforated cell, bhp cell) of each perforation in the pattern
    std::vector<int64_t> ro(static_cast<size_t>(ncell) + 1), cols(static_cast<size_t>(a->nnzb));
    NN_HIP(hipMemcpy(ro.data(), a->rowOffsetsDevice, sizeof(int64_t) * (ncell + 1), hipMemcpyDeviceToHost));
    NN_HIP(hipMemcpy(cols.data(), a->colBlocksDevice, sizeof(int64_t) * a->nnzb, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < nperforations; ++k)
    {
        if (perfWell[static_cast<size_t>(k)] < 0)
            return fail(NN_E_INVALID_ARGUMENT, "a perforation belongs to no well");
        const int64_t row = perforation_cells[k];
        const int64_t col = wells[perfWell[static_cast<size_t>(k)]].bhp_cell;
        const int64_t *b = cols.data() + ro[row], *e = cols.data() + ro[row + 1];
        const int64_t *it = std::lower_bound(b, e, col);
        if (it == e || *it != col)
            return fail(NN_E_INVALID_ARGUMENT,
                        "the pattern lacks a well coupling block (nn_pattern_add_wells)");
        bhpSlot[static_cast<size_t>(k)] = int64_t(it - cols.data());
    }
    releaseWells(a);
    auto upload = [](auto **dst, const auto &src) {
        using T = std::remove_reference_t<decltype(src[0])>;
        const size_t bytes = sizeof(T) * std::max<size_t>(1, src.size());
        if (hipMalloc(reinterpret_cast<void **>(dst), bytes) != hipSuccess)
            return false;
        return src.empty() ||
               hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice) ==
                   hipSuccess;
    };
    std::vector<double> wi(perforation_well_index, perforation_well_index + nperforations);
    if (!upload(&a->dWellCellOffsets, offsets) || !upload(&a->dWellCellPerfs, order) ||
        !upload(&a->dPerfWell, perfWell) || !upload(&a->dPerfIndex, wi) ||
        !upload(&a->dPerfBhpSlot, bhpSlot) || !upload(&a->dWells, dw) ||
        !upload(&a->dWellCells, cellsWithWells))
    {
        releaseWells(a);
        return fail(NN_E_OUT_OF_MEMORY, "device allocation of the wells failed");
    }
    a->nWells = nwells;
    a->nWellCells = int64_t(cellsWithWells.size());
    AssemblyView &v = a->view;
    v.wellCellOffsets = a->dWellCellOffsets;
    v.wellCellPerfs = a->dWellCellPerfs;
    v.perfWell = a->dPerfWell;
    v.perfIndex = a->dPerfIndex;
    v.perfBhpSlot = a->dPerfBhpSlot;
    v.wells = a->dWells;
    for (int p = 0; p < 3; ++p) v.surfaceDensity[p] = surface_density[p];
    return NN_OK;
}

int nn_assembly_set_scaling(nn_assembly *a, const nn_scaling *s, const double *d_well_row_scale)
{
    if (!a || !s)
        return fail(NN_E_INVALID_ARGUMENT, "null argument");
    const double f[7] = {s->pressure_scale, s->composition_scale, s->saturation_scale,
                         s->mass_residual_scale, s->fugacity_residual_scale,
                         s->closure_residual_scale, s->rate_well_residual_floor};
    for (double v : f)
        if (!(v > 0.0) || !std::isfinite(v))
            return fail(NN_E_INVALID_ARGUMENT, "scaling factors must be finite and positive");
    const int nc = a->ctx->host.nc;
    const int np = 3 * nc + 1 + (a->hasWell ? 1 : 0);
    AssemblyView &v = a->view;
    for (int k = 0; k < kMaxNp; ++k)
        v.rowScale[k] = v.varScale[k] = 1.0;
    v.wellRowScale = nullptr;
    if (!s->enabled)
        return NN_OK;
    // natural_petsc_runtime_scaling.inc:8-89 on this layout: equations
    // mass[0, nc), fugacity [nc, 2nc), water fugacity [2nc, 3nc), closure 3nc,
    // well control 3nc+1; variables p, compositions [1, 3nc-2), saturations
    // [3nc-2, 3nc+1), well pressure 3nc+1.
    for (int e = 0; e < np; ++e)
    {
        if (e < nc)
            v.rowScale[e] = 1.0 / s->mass_residual_scale;
        else if (e < 3 * nc)
            v.rowScale[e] = 1.0 / s->fugacity_residual_scale;
        else if (e == 3 * nc)
            v.rowScale[e] = 1.0 / s->closure_residual_scale;
        else
            v.rowScale[e] = 1.0 / s->pressure_scale;
    }
    for (int q = 0; q < np; ++q)
    {
        if (q == 0 || q == 3 * nc + 1)
            v.varScale[q] = s->pressure_scale;
        else if (q >= 3 * nc - 2 && q <= 3 * nc)
            v.varScale[q] = s->saturation_scale;
        else
            v.varScale[q] = s->composition_scale;
    }
    v.wellRowScale = a->hasWell ? d_well_row_scale : nullptr;
    return NN_OK;
}

int nn_newton_limit(nn_assembly *a, const double *d_primary, double *d_step,
                    double max_aqueous_unsupported, int32_t water_component,
                    int32_t co2_component, void *stream)
{
    if (!a || !d_primary || !d_step)
        return fail(NN_E_INVALID_ARGUMENT, "null argument");
    const int nc = a->ctx->host.nc;
    // natural_petsc_runtime_common.inc:43-60
    if (max_aqueous_unsupported >= 0.0 &&
        (!finite(max_aqueous_unsupported) || water_component < 0 || water_component >= nc ||
         co2_component < 0 || co2_component >= nc || water_component == co2_component))
        return fail(NN_E_INVALID_ARGUMENT, "invalid aqueous Newton composition-domain configuration");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t n = a->view.ncell;
    if (n == 0)
        return NN_OK;
    const unsigned grid = static_cast<unsigned>((n + 255) / 256);
    return dispatchNc(nc, [&](auto ncTag) {
        constexpr int NC = decltype(ncTag)::value;
        if (a->hasWell)
            newtonLimitKernel<NC, true><<<grid, 256, 0, st>>>(n, d_primary, d_step, a->view,
                                                              max_aqueous_unsupported,
                                                              water_component, co2_component);
        else
            newtonLimitKernel<NC, false><<<grid, 256, 0, st>>>(n, d_primary, d_step, a->view,
                                                               max_aqueous_unsupported,
                                                               water_component, co2_component);
        NN_HIP(hipGetLastError());
        return NN_OK;
    });
}

int nn_assembly_accumulation(nn_assembly *a, const double *d_primary, const double *d_phase_state,
                             double *d_accum, void *stream)
{
    if (!a || !d_primary || !d_phase_state || !d_accum)
        return fail(NN_E_INVALID_ARGUMENT, "null argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    AssemblyView v = a->view;
    v.record = nullptr;
    v.roots = nullptr;
    return dispatchNc(a->ctx->host.nc, [&](auto ncTag) {
        constexpr int NC = decltype(ncTag)::value;
        NN_HIP((Launch<NC>::record(a->hasWell != 0, st, a->ctx->dFluid, v, d_primary, d_phase_state,
                                   d_accum, nullptr)));
        return NN_OK;
    });
}

int nn_assembly_evaluate(nn_assembly *a, double dt, double fugacity_scaling,
                         const double *d_primary, const double *d_phase_state,
                         const double *d_accum_prev, double *d_residual, double *d_bsr,
                         uint8_t *d_decisions, int32_t *d_status, void *stream)
{
    if (!a || !d_primary || !d_phase_state || !d_accum_prev)
        return fail(NN_E_INVALID_ARGUMENT, "null argument");
    if (!(dt > 0.0) || !(fugacity_scaling > 0.0))
        return fail(NN_E_INVALID_ARGUMENT, "time step and fugacity scaling must be positive");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const AssemblyView v = a->view;
    return dispatchNc(a->ctx->host.nc, [&](auto ncTag) {
        constexpr int NC = decltype(ncTag)::value;
        LaunchTimer tr(a->ctx, NN_KERNEL_CELL_RECORD, st);
        NN_HIP((Launch<NC>::record(a->hasWell != 0, st, a->ctx->dFluid, v, d_primary, d_phase_state,
                                   nullptr, nullptr)));
        tr.done();
        if (d_bsr && a->nZero > 0)
        {
            const int64_t np = a->hasWell ? 3 * NC + 2 : 3 * NC + 1;
            const int64_t count = a->nZero * np * np;
            zeroBlocksKernel<<<static_cast<unsigned>((count + 255) / 256), 256, 0, st>>>(
                d_bsr, a->dZeroSlots, a->nZero, int(np));
            NN_HIP(hipGetLastError());
        }
        LaunchTimer ta(a->ctx, NN_KERNEL_ASSEMBLE, st);
        NN_HIP((Launch<NC>::assemble(a->hasWell != 0, st, a->ctx->dFluid, v, dt, fugacity_scaling,
                                     d_primary, d_phase_state, d_accum_prev, d_residual, d_bsr,
                                     d_decisions, d_status)));
        ta.done();
        if (a->nWells > 0)
            NN_HIP((Launch<NC>::wells(st, a->ctx->dFluid, v, a->nWells, a->dWellCells,
                                      a->nWellCells, d_primary, d_phase_state, d_residual, d_bsr,
                                      d_status)));
        return NN_OK;
    });
}

} // extern "C"
