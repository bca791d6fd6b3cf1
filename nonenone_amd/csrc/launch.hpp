// launch.hpp — host-side launch interface of the per-component-count kernels.
//
// Each component count is compiled in its own translation unit
// (nn_nc3.hip, nn_nc4.hip, nn_nc6.hip) that instantiates Launch<NC>; the C ABI
// in nn_api.hip dispatches on the run-time nc.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nonenone_amd.h"
#include "common.hpp"
#include "wells.hpp"

namespace nn
{

constexpr int kBlock = 64;  // one wavefront per block: divergent per-state control flow

// Minimum wavefronts per SIMD that every per-state kernel is compiled for.  The
// flash, stability and phase-update kernels share out-of-line solver functions
// (Flash::flashTwo, flashThree, thermoForRole, ...); a callee's register budget
// follows the launch bounds only when every kernel that calls it asks for the
// same occupancy, so all kernels of this file carry the same value.  At 1 the
// shared callees take the full 512 registers (one wavefront per SIMD).  At 2 the
// PR-CPA phase update of a 500,000-cell grid ran 7 % faster, but scratch grows
// with occupancy (private segment x resident waves per queue), and a second
// hardware queue (another stream, e.g. a graph-capture stream) could then not
// be given its scratch: HSA_STATUS_ERROR_OUT_OF_RESOURCES.  Build parameter
// (nonenone_amd/build.py, NN_WAVES_PER_EU); 1 by default.
#ifndef NN_WAVES_PER_EU
#define NN_WAVES_PER_EU 1
#endif
constexpr int kWavesPerEu = NN_WAVES_PER_EU;

// Derivative columns per chunk of the assembly kernel's forward-mode AD: each
// (cell, chunk of primary variables) pair is one work-item evaluating the cell
// with Dual<kAssemblyChunk>.
// Build parameter (nonenone_amd/build.py, NN_ASSEMBLY_CHUNK).
#ifndef NN_ASSEMBLY_CHUNK
#define NN_ASSEMBLY_CHUNK 2
#endif
constexpr int kAssemblyChunk = NN_ASSEMBLY_CHUNK;

// Largest component count the Natural assembly is built for.
constexpr int kAssemblyMaxNc = NN_MAX_NC;

constexpr int kMaxNp = 3 * NN_MAX_NC + 2;  // primary variables with a well unknown

struct AssemblyView
{
    int64_t ncell;
    const int64_t *connOffsets;
    const int64_t *connNeighbor;
    const double *connTrans;
    const double *connGdz;
    const double *volume;
    const double *porosity;
    const double *T;
    const int64_t *selfSlot;   // BSR slot of block (c, c)
    const int64_t *recipSlot;  // per connection: BSR slot of block (neighbour, c)
    double *record;            // neighbour property records, field-major (SoA)
    // PR-CPA (else null): per cell and phase, the density root and its two
    // correction slopes from the record pass, field-major (9 fields), which the
    // assembly pass reuses instead of repeating the root search.
    double *roots;
    // Equation/variable scaling of the Natural runtime (all 1 when disabled):
    // residual[r] *= rowScale[r]; J[r][q] = (dR_r/dx_q * rowScale[r]) * varScale[q].
    // wellRowScale (per cell, may be null) replaces rowScale of the well-control row.
    const double *wellRowScale;
    double rowScale[kMaxNp];
    double varScale[kMaxNp];
    // Wells (nn_assembly_set_wells; wellCellOffsets null when none): per cell the
    // perforations [wellCellOffsets[c], wellCellOffsets[c + 1]) of wellCellPerfs, in
    // well then perforation order; per perforation its well, well index and the
    // BSR slot of block (cell, bhp cell).
    const int64_t *wellCellOffsets;
    const int32_t *wellCellPerfs;
    const int32_t *perfWell;
    const double *perfIndex;
    const int64_t *perfBhpSlot;
    const DeviceWell *wells;
    double surfaceDensity[3];
};

// Counters of a two-pass (lane-per-item, then wavefront-per-deferred-item) launch.
struct WorkQueue
{
    unsigned int head;      // next item of the lane-per-item pass
    unsigned int deferred;  // items handed to the cooperative pass (= list length)
    unsigned int coopHead;  // next list entry of the cooperative pass
    unsigned int deferAll;  // nonzero: every item goes to the cooperative pass (testing)
    unsigned int claim;     // 0: lanes claim single items; 1: wavefronts claim 64-item tiles
    unsigned int pending;   // items recorded for the lockstep stability trials (list length)
    unsigned int trialHead; // next pending item of the lockstep trial pass
    unsigned int replayHead;// next pending item of the replay pass
    unsigned int reflash;   // items whose re-flash the replay pass held back (list length)
    unsigned int reflashHead; // next item of the re-flash pass
};

// Processing order of the re-flash pass: the key kernel writes n keys and cell
// indices (held cells first, in list order; every other position keyed last)
// into keys/idx, and sort(arg, stream, &order) returns them sorted by key, so
// neighbouring positions hold cells of similar conditions.
struct ReflashOrder
{
    unsigned int *keys;
    unsigned int *idx;
    hipError_t (*sort)(void *arg, hipStream_t st, const unsigned int **order);
    void *arg;
};

// Per-state hints of the learned initialiser (both may be null): hint[s] is
// NN_HINT_*, logK is n x NC.
struct FlashHint
{
    const uint8_t *hint;
    const double *logK;
};

template <int NC>
struct Launch
{
    static hipError_t eos(unsigned grid, hipStream_t st, const FluidParams *fp, int64_t n,
                          const double *P, const double *T, const double *x, const int32_t *role,
                          double *Z, double *lnphi, double *fug, int32_t *status);
    // Lane-per-state pass, then one wavefront per deferred state.  `splitA`
    // and `splitB` (may be null) are recorded between the two launches.
    // `plain`: one work-item per state in index order instead of the persistent
    // queue (the baseline form; `grid` is then ceil(n / kBlock)).
    static hipError_t flash(unsigned grid, unsigned coopGrid, hipStream_t st,
                            const FluidParams *fp, const FlashOptions *op, int64_t n,
                            const double *P, const double *T, const double *z,
                            const uint8_t *allowed, const double *seed, FlashHint hint,
                            nn_flash_out out, WorkQueue *q, unsigned int *list,
                            const unsigned int *order, bool plain, hipEvent_t splitA,
                            hipEvent_t splitB);
    static hipError_t stability(unsigned grid, hipStream_t st, const FluidParams *fp,
                                const FlashOptions *op, int64_t n, const double *P,
                                const double *T, const double *z, const uint8_t *active,
                                const double *x, uint8_t *stable, uint8_t *valid, uint8_t *mask,
                                double *trial, double *incipient, int32_t *status);
    // `pendingList` / `records` (may be null: no lockstep trials): cells whose
    // direct stability test is recorded by the lane-per-cell pass, run by the
    // lockstep trial pass and finished by the replay pass — whose cells that
    // must re-flash go to a last pass of their own (`pendingList + n`) — all
    // before the cooperative pass.  `pendingList` holds 2n entries, `records`
    // n StabRecord<NC>.
    // `reorder` (may be null): sorts the re-flash pass's cells by condition first.
    static hipError_t phaseUpdate(bool well, unsigned grid, unsigned coopGrid, hipStream_t st,
                                  const FluidParams *fp, const FlashOptions *op, int64_t n,
                                  const double *T, double *primary, double *phaseState,
                                  int32_t *status, uint8_t *flags, WorkQueue *q,
                                  unsigned int *list, const unsigned int *order,
                                  unsigned int *pendingList, void *records,
                                  const ReflashOrder *reorder, hipEvent_t splitA,
                                  hipEvent_t splitB);
    // Sort keys (and identity indices) for the processing order of a batch.
    static hipError_t flashKeys(hipStream_t st, int64_t n, const double *P, const double *T,
                                const double *z, const uint8_t *hint, unsigned int *keys,
                                unsigned int *idx);
    static hipError_t phaseUpdateKeys(bool well, hipStream_t st, int64_t n, const double *T,
                                      const double *primary, const double *phaseState,
                                      unsigned int *keys, unsigned int *idx);
    static hipError_t record(bool well, hipStream_t st, const FluidParams *fp, AssemblyView v,
                             const double *primary, const double *phaseState, double *accum,
                             int32_t *status);
    static hipError_t assemble(bool well, hipStream_t st, const FluidParams *fp, AssemblyView v,
                               double dt, double fugScale, const double *primary,
                               const double *phaseState, const double *accumPrev,
                               double *residual, double *bsr, uint8_t *decisions,
                               int32_t *status);
    // The wells' Jacobian rows (one thread per perforated cell of `cells`) and
    // well-control residuals, after assemble.
    static hipError_t wells(hipStream_t st, const FluidParams *fp, AssemblyView v, int nwells,
                            const int64_t *cells, int64_t ncells, const double *primary,
                            const double *phaseState, double *residual, double *bsr,
                            int32_t *status);
};

} // namespace nn
