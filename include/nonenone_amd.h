/*
 * nonenone_amd.h — C ABI of the MI355X per-cell flash + Natural assembly library.
 *
 * Every entry point takes plain pointers and sizes.  Array arguments named
 * `d_*` are DEVICE pointers (HBM, allocated by the caller); `stream` is a
 * hipStream_t passed as void* (NULL = default stream).  All calls are
 * asynchronous on `stream` unless stated otherwise.  Return value: NN_OK or a
 * negative NN_E* code; per-item outcomes are reported in status arrays.
 *
 * Units: pressure Pa, temperature K, molar mass kg/mol, critical volume m3/mol.
 * Phase slots follow the reference: 0 = oil-rich liquid, 1 = gas, 2 = water-rich.
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   fluid/EOS configuration   models/include/natural/compositional_mixture.hpp:24-157,
 *                             models/include/natural/thermo/cubic_eos.hpp:178-311
 *   flash options             models/include/natural/thermo/three_phase_flash_types.hpp:21-41
 *   flash / stability         models/include/natural/thermo/three_phase_flash.hpp:83-489
 *   phase-state update        models/include/natural/state/three_phase_equilibrium.hpp:119-496
 *   residual / Jacobian       models/include/natural/petsc/detail/natural_petsc_runtime_residual.inc:19-133,
 *                             models/include/natural/petsc/detail/natural_petsc_runtime_jacobian.inc:8-299
 *   cached BSR topology       models/include/natural/petsc/jacobian_factory.hpp:27-238
 *   TPFA connections          grid/include/structuredgrid/structuredgrid.hpp:541-667,877-907
 */
#ifndef NONENONE_AMD_H
#define NONENONE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NN_MAX_NC 8

/* ---- return codes ---------------------------------------------------- */
#define NN_OK 0
#define NN_E_INVALID_ARGUMENT (-1)
#define NN_E_UNSUPPORTED (-2)
#define NN_E_HIP (-3)
#define NN_E_NO_DEVICE (-4)
#define NN_E_OUT_OF_MEMORY (-5)

/* ---- per-item status (uint8/int32 status arrays) ---------------------- */
#define NN_STATUS_OK 0             /* converged / evaluated                  */
#define NN_STATUS_NOT_CONVERGED 1  /* flash returned converged=false          */
#define NN_STATUS_THERMO_ERROR 2   /* reference would throw (no admissible
                                      root, non-positive volume, ...)         */
#define NN_STATUS_INVALID_INPUT 3  /* non-finite / non-positive P, T, z       */
#define NN_STATUS_RECOVERABLE_FAILURE 4 /* phase update restored its inputs   */

/* thermodynamic model, models/include/natural/thermo/thermodynamic_model.hpp:10-15 */
#define NN_MODEL_PR 0
#define NN_MODEL_SW 1
#define NN_MODEL_CPA 2

/*
 * Fluid description.  Mirrors CompositionalMixture (Tc, Pc, Vc, omega, M, k_ij),
 * the CubicEquationOfState constructor (Omega_a, Omega_b, eosType, m1, m2),
 * SoreideWhitsonOptions (water index, salinity, aqueous H2O-solute k_ij) and
 * CubicPlusAssociationOptions, plus the benchmark Corey relative permeability
 * (case/h2o_co2_nc10_2d_benchmark/benchmark_common.hpp:97-126).
 */
typedef struct nn_fluid_desc {
    int32_t nc;                 /* number of components, 2..NN_MAX_NC          */
    int32_t thermo_model;       /* NN_MODEL_*                                  */
    int32_t eos_type;           /* 1 = PR76 m(omega); 5 = PR78 for omega>0.49  */
    int32_t water_component;    /* H2O index (three-phase flash needs >= 0)    */
    double omega_a, omega_b, m1, m2;
    double tc[NN_MAX_NC], pc[NN_MAX_NC], vc[NN_MAX_NC];
    double acentric[NN_MAX_NC], molar_mass[NN_MAX_NC];
    double kij[NN_MAX_NC * NN_MAX_NC];          /* row-major, symmetric       */
    /* Soreide-Whitson (thermo_model == NN_MODEL_SW) */
    double sw_salinity;                         /* NaCl molality               */
    double sw_aqueous_kij[NN_MAX_NC];           /* k(H2O,i) in the Water role  */
    int32_t sw_aqueous_kij_set[NN_MAX_NC];      /* 0: fall back to kij         */
    /* CPA (thermo_model == NN_MODEL_CPA) */
    int32_t cpa_physical_term;                  /* 0 SRK, 1 PR                 */
    int32_t cpa_max_iterations;                 /* association iterations      */
    double cpa_tolerance, cpa_damping;
    double cpa_a0[NN_MAX_NC], cpa_b[NN_MAX_NC], cpa_c1[NN_MAX_NC];
    double cpa_epsilon[NN_MAX_NC], cpa_beta[NN_MAX_NC];
    int32_t cpa_donors[NN_MAX_NC], cpa_acceptors[NN_MAX_NC];
    double cpa_cross_epsilon[NN_MAX_NC * NN_MAX_NC];
    double cpa_cross_beta[NN_MAX_NC * NN_MAX_NC];
    /* Corey relative permeability: swc, sorw, sgc, sorg, krw_max, krg_max, kro_max */
    double relperm[7];
    /* Property closures of the Water role (all zero = off; case_config.hpp:70-74
     * switches both on in the reference's benchmark case):
     *   aqueous_volume 1: IAPWS-IF97 water density + Garcia (2001) CO2 apparent
     *     volume for the Water-role molar volume (CubicEquationOfState::
     *     configureAqueousVolume, cubic_eos.hpp:405-412; aqueous_volume.hpp),
     *     with its composition domain: at most aqueous_volume_max_unsupported of
     *     components other than H2O and CO2;
     *   aqueous_viscosity NN_AQ_VISCOSITY_MCBRIDE_WRIGHT: McBride-Wright (2015)
     *     H2O-CO2 viscosity (FluidSystem::configureMcBrideWrightAqueousViscosity),
     *     NN_AQ_VISCOSITY_IAPWS2008: IAPWS-2008 industrial pure-water viscosity
     *     (configureIapws2008AqueousViscosity together with the EOS composition
     *     domain configureAqueousCompositionDomain); aqueous_viscosity_max_unsupported
     *     is the former's non-H2O/CO2 and the latter's non-H2O limit
     *     (aqueous_viscosity.hpp). */
    int32_t aqueous_volume;
    int32_t aqueous_viscosity;
    int32_t co2_component;
    int32_t aqueous_reserved;
    double water_molar_mass;                  /* kg/mol, e.g. 0.018015268 */
    double aqueous_volume_max_unsupported;    /* e.g. 1e-4 */
    double aqueous_viscosity_max_unsupported; /* e.g. 1e-4 */
} nn_fluid_desc;

#define NN_AQ_VISCOSITY_NONE 0
#define NN_AQ_VISCOSITY_MCBRIDE_WRIGHT 1
#define NN_AQ_VISCOSITY_IAPWS2008 2

/* ThreePhaseFlashOptions (three_phase_flash_types.hpp:21-41); waterComponent
 * is taken from nn_fluid_desc.water_component. */
typedef struct nn_flash_options {
    int32_t maximum_iterations;               /* 120 */
    int32_t maximum_rachford_rice_iterations; /* 80  */
    int32_t maximum_stability_iterations;     /* 80  */
    int32_t stability_ssi_before_newton;      /* 8   */
    int32_t maximum_stability_newton;         /* 12  */
    int32_t two_phase_ssi_before_newton;      /* 8   */
    int32_t maximum_two_phase_newton;         /* 12  */
    int32_t three_phase_ssi_before_newton;    /* 10  */
    int32_t maximum_three_phase_newton;       /* 12  */
    double fugacity_tolerance;                /* 1e-9  */
    double rachford_rice_tolerance;           /* 1e-12 */
    double stability_tolerance;               /* 1e-9  */
    double phase_fraction_tolerance;          /* 1e-10 */
    double coincident_composition_tolerance;  /* 1e-8  */
    double composition_floor;                 /* 1e-30 */
    double log_k_step_limit;                  /* 2     */
    double numerical_jacobian_step;           /* 1e-6  */
} nn_flash_options;

/* Fills `o` with the reference defaults listed above. */
void nn_flash_options_default(nn_flash_options *o);

typedef struct nn_context nn_context;

/* Creates a context on the current HIP device: uploads the fluid to device
 * constant storage.  Replaces constructing CubicEquationOfState +
 * CubicThreePhaseFlash (three_phase_flash.hpp:47-81).  Synchronous.
 * Returns NN_E_UNSUPPORTED for nc = 5 and nc = 8: those device builds are
 * switched off (a GPU memory fault of the five-component phase update whose
 * cause is not identified). */
int nn_context_create(const nn_fluid_desc *fluid, const nn_flash_options *opts,
                      nn_context **out);
int nn_context_destroy(nn_context *ctx);
/* Last error text of this context (static storage, never NULL). */
const char *nn_last_error(void);

/*
 * Kernel timing (no counterpart in the reference: instrumentation for
 * throughput and roofline figures).  While enabled, every launch made through
 * this context is bracketed by HIP events recorded on the launch's own stream;
 * nn_context_kernel_time waits for those events and returns the summed
 * duration and launch count of one kernel kind since timing was last enabled.
 * Enabling again clears the record.
 */
#define NN_KERNEL_EOS 0
#define NN_KERNEL_FLASH 1
#define NN_KERNEL_STABILITY 2
#define NN_KERNEL_PHASE_UPDATE 3
#define NN_KERNEL_CELL_RECORD 4
#define NN_KERNEL_ASSEMBLE 5
#define NN_KERNEL_FLASH_COOP 6        /* wavefront-per-state pass of nn_flash_*       */
#define NN_KERNEL_PHASE_UPDATE_COOP 7 /* wavefront-per-cell pass of nn_phase_update  */
#define NN_KERNEL_KINDS 8
int nn_context_set_kernel_timing(nn_context *ctx, int32_t enable);
int nn_context_kernel_time(nn_context *ctx, int32_t kind, double *total_ms,
                           int64_t *launches);

/* Work counters of the context's last flash or phase-update launch (no
 * counterpart in the reference: instrumentation; synchronises the device):
 * items handed to the wavefront-per-item pass, cells recorded for the lockstep
 * stability trials, and cells held for the re-flash pass. */
int nn_context_queue_counts(nn_context *ctx, int64_t *deferred, int64_t *recorded,
                            int64_t *reflashed);

/*
 * E1-E5: batched single-phase EOS evaluation, CubicEquationOfState::phaseResult
 * with a phase role (cubic_eos.hpp:834-856): mixingParameters -> phaseMixing ->
 * liquid/vapour root -> fugacityAtRoot.  role[i] in {0 oil,1 gas,2 water}.
 * Layout: d_x, d_lnphi, d_fug are n x nc row-major.
 */
int nn_eos_phase_batch(nn_context *ctx, int64_t n,
                       const double *d_P, const double *d_T, const double *d_x,
                       const int32_t *d_role,
                       double *d_Z, double *d_lnphi, double *d_fug,
                       int32_t *d_status, void *stream);

/* Flash outputs (ThreePhaseFlashResult, three_phase_flash_types.hpp:43-58),
 * structure of arrays, all device pointers, n items. */
typedef struct nn_flash_out {
    uint8_t *presence;   /* n     : bit0 oil, bit1 gas, bit2 water            */
    uint8_t *converged;  /* n                                                  */
    int32_t *status;     /* n     : NN_STATUS_*                               */
    int32_t *iterations; /* n                                                  */
    double *beta;        /* n x 3 : phase mole fractions                      */
    double *saturation;  /* n x 3                                             */
    double *x;           /* n x 3 x nc : phase compositions                   */
    double *Z;           /* n x 3 : compressibility factors                   */
    double *rho_m;       /* n x 3 : molar densities                           */
    double *k_go;        /* n x nc: vaporOilK                                 */
    double *k_wo;        /* n x nc: waterOilK                                 */
} nn_flash_out;

/* F2: CubicThreePhaseFlash::flash(P,T,z) (three_phase_flash.hpp:83-227) on n
 * independent states.  d_z is n x nc.  Persistent work-stealing launch. */
int nn_flash_batch(nn_context *ctx, int64_t n,
                   const double *d_P, const double *d_T, const double *d_z,
                   const nn_flash_out *out, void *stream);

/*
 * F-seed: nn_flash_batch with a per-state hint from a learned initialiser (no
 * counterpart in the reference).  d_hint[i]: NN_HINT_NONE runs the reference
 * path unchanged; NN_HINT_SINGLE_PHASE skips the first gas-oil two-phase start
 * of three_phase_flash.hpp:93-101 and continues from the single-phase candidate;
 * NN_HINT_LOG_K replaces that start's Wilson K by exp(d_log_k[i][:]) (n x nc,
 * may be NULL when no state carries NN_HINT_LOG_K).  Every later stage (active-set
 * equilibration with its stability tests, three-phase flash, recovery chain) is
 * the reference's, so the stability tests accept or overrule the hint.  Ordered
 * batches sort by hint first, so states on the same solver path share wavefronts.
 */
#define NN_HINT_NONE 0
#define NN_HINT_SINGLE_PHASE 1
#define NN_HINT_LOG_K 2
/* Bits 2-3 of a hint: predicted phase count - 1, used only to order the batch
 * (the solver path is the reference's); nn_mlp_hints sets them with modes bit
 * NN_HINT_ORDER.  A batch whose hints carry only these bits gives results
 * bit-identical to nn_flash_batch. */
#define NN_HINT_ORDER 4
int nn_flash_seeded_batch(nn_context *ctx, int64_t n,
                          const double *d_P, const double *d_T, const double *d_z,
                          const uint8_t *d_hint, const double *d_log_k,
                          const nn_flash_out *out, void *stream);

/*
 * nn_flash_batch without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The persistent
 * kernels' work queue and deferred-state list live in a file-scope device array
 * of the library (256 B + 4 B x NN_STATIC_MAX_STATES = 4 MiB per GPU, held until
 * the process ends; its address is resolved per device by nn_context_create).
 * Same kernels and results as nn_flash_batch in index order; the schedule's
 * sorted order is not applied (it needs sort buffers).  n > NN_STATIC_MAX_STATES
 * returns NN_E_INVALID_ARGUMENT.  Every call on one GPU shares the array: calls
 * that overlap on one GPU (other streams, other threads) race and must not be
 * issued.  Calls no hipMalloc/hipFree and no allocating library routine.
 */
#define NN_STATIC_MAX_STATES (1 << 20)
int nn_flash_batch_static(nn_context *ctx, int64_t n,
                          const double *d_P, const double *d_T, const double *d_z,
                          const nn_flash_out *out, void *stream);

/*
 * Launch shapes of the flash-like batches of a context (no counterpart in the
 * reference: how the per-state work maps onto the GPU).  Defaults {16, 4, 1, 0, 0}.
 *   persistent_blocks_per_cu  grid of the lane-per-state pass (64-thread blocks
 *                             per CU, capped by the batch), 1..64
 *   coop_blocks_per_cu        grid of the wavefront-per-deferred-state pass, 1..64
 *   ordered                   process batches of >= 4096 states in radix-sorted
 *                             condition-key order
 *   wave_claim                0: lanes claim single states from the queue;
 *                             1: wavefronts claim tiles of 64 consecutive states
 *   plain                     1: the baseline form for flash batches and phase
 *                             updates: one work-item per state or cell in index
 *                             order (no persistent queue, no sort, no lockstep
 *                             stability trials; other fields ignored); states
 *                             past the lane budget still go to the cooperative
 *                             pass
 */
typedef struct nn_schedule {
    int32_t persistent_blocks_per_cu;
    int32_t coop_blocks_per_cu;
    int32_t ordered;
    int32_t wave_claim;
    int32_t plain;
} nn_schedule;
int nn_context_get_schedule(nn_context *ctx, nn_schedule *s);
int nn_context_set_schedule(nn_context *ctx, const nn_schedule *s);

/* Autotuner: times nn_flash_batch on the given sample (device arrays, outputs
 * overwritten) under each candidate schedule (ordered x wave_claim x
 * persistent_blocks_per_cu in {4, 8, 16, 32}), best of `repeats` runs each,
 * keeps the fastest in the context and returns it with its time.  Synchronous. */
int nn_flash_autotune(nn_context *ctx, int64_t n,
                      const double *d_P, const double *d_T, const double *d_z,
                      const nn_flash_out *out, int32_t repeats,
                      nn_schedule *best, double *best_ms, void *stream);

/* F-restricted: CubicThreePhaseFlash::flashRestricted(P,T,z,allowed[,seed])
 * (three_phase_flash.hpp:235-292).  d_allowed: n presence masks; d_seed may be
 * NULL (no composition seed) or n x 3 x nc. */
int nn_flash_restricted_batch(nn_context *ctx, int64_t n,
                              const double *d_P, const double *d_T,
                              const double *d_z, const uint8_t *d_allowed,
                              const double *d_seed,
                              const nn_flash_out *out, void *stream);

/* F7: CubicThreePhaseFlash::stabilityTest(P,T,z,active,x) (three_phase_flash.hpp:300-489).
 * d_x: n x 3 x nc active compositions.  Outputs: stable/valid flags,
 * missing-phase-unstable mask (bit per slot), trial sums n x 3 and incipient
 * compositions n x 3 x nc. */
int nn_stability_batch(nn_context *ctx, int64_t n,
                       const double *d_P, const double *d_T, const double *d_z,
                       const uint8_t *d_active, const double *d_x,
                       uint8_t *d_stable, uint8_t *d_valid,
                       uint8_t *d_unstable_mask, double *d_trial_sum,
                       double *d_incipient, int32_t *d_status, void *stream);

/*
 * S1: FullyCompositionalThreePhaseEquilibrium::updatePhaseState
 * (three_phase_equilibrium.hpp:282-496) on n cells, in place.
 * d_primary: n x np (np = 3nc+1, +1 with a well unknown), reference Primary layout
 *            p, x_o[0..nc-2], x_g[..], x_w[..], S_o, S_g, S_w [, p_bhp].
 * d_phase_state: n x (3nc+8), reference PhaseState Vec layout
 *            flag, K_go[nc], K_wo[nc], z[nc], beta[3], Z_o, Z_g, Z_w, suppression.
 * d_T: per-cell temperature.  d_status: NN_STATUS_* (RECOVERABLE_FAILURE =
 * inputs restored).  d_flags: bit0 phaseRemoved, bit1 stabilityInvalid,
 * bit2 missingPhaseUnstable, bit3 restrictedFlashFailed,
 * bit4 unrestrictedFlashFailed, bit5 stableReducedSet (may be NULL); 0 for a
 * cell whose status is THERMO_ERROR or INVALID_INPUT (the reference throws and
 * returns no report).
 */
int nn_phase_update_batch(nn_context *ctx, int64_t n, int32_t has_well_unknown,
                          const double *d_T, double *d_primary,
                          double *d_phase_state, int32_t *d_status,
                          uint8_t *d_flags, void *stream);

/*
 * All-or-none phase update: NaturalPetscRuntime::updateStateWithReport
 * (models/include/natural/petsc/detail/natural_petsc_runtime_state.inc:166-334).
 * Every cell is updated on copies in the caller's workspace (d_work_primary
 * n x np, d_work_phase_state n x nps); only when no cell fails are the copies
 * written back over d_primary / d_phase_state, otherwise both stay untouched
 * (the reference's transaction: a recoverable thermodynamic failure -- and here
 * any non-OK cell status -- leaves no half-updated state).  The report, written
 * to device memory (d_report), follows StateUpdateReport: the first failing cell
 * in index order with its status and flag bits, the cells updated before it
 * (all n on success) and how many of those were StableReducedSet updates
 * (flag bit 5).  d_status / d_flags receive every cell's outcome.  Asynchronous
 * on `stream`, no host synchronisation and no memset (capturable in a graph).
 */
typedef struct nn_update_report {
    int64_t first_failure_cell;    /* -1: every cell updated and committed       */
    int64_t updated_cells;         /* cells before the first failure (n on success) */
    int64_t stable_reduced_cells;  /* of those, StableReducedSet updates          */
    int32_t succeeded;             /* 1: committed                                */
    int32_t first_failure_status;  /* NN_STATUS_* of the first failing cell       */
    int32_t first_failure_flags;   /* its flag bits (as d_flags)                  */
    int32_t reserved;
} nn_update_report;
int nn_phase_update_transaction(nn_context *ctx, int64_t n, int32_t has_well_unknown,
                                const double *d_T, double *d_primary, double *d_phase_state,
                                double *d_work_primary, double *d_work_phase_state,
                                int32_t *d_status, uint8_t *d_flags,
                                nn_update_report *d_report, void *stream);

/* ---- Natural assembly -------------------------------------------------- */

/* Host-side setup (no device work): 7-point structured TPFA connections in the
 * reference face order X-,X+,Y-,Y+,Z-,Z+ with harmonic transmissibility and
 * g*dz on Z faces (structuredgrid.hpp:541-667,877-907), plus the cached BSR
 * pattern: per row block the sorted unique column blocks {self, neighbours}
 * (jacobian_factory.hpp:187-201).  perm is ncell x 3 (kx,ky,kz), dx/dy/dz are
 * per-axis widths.  Output sizes: conn_offsets ncell+1, conn_* nconn
 * (<= 6*ncell), row_offsets ncell+1, col_blocks nnzb.  Pass NULL outputs to
 * query counts via *nconn / *nnzb.  Cell order: i fastest, then j, then k. */
int nn_structured_topology(int32_t nx, int32_t ny, int32_t nz,
                           const double *dx, const double *dy, const double *dz,
                           const double *perm,
                           int64_t *conn_offsets, int64_t *conn_neighbor,
                           double *conn_trans, double *conn_gdz,
                           int64_t *row_offsets, int64_t *col_blocks,
                           int64_t *nconn, int64_t *nnzb);

/* Host routine (no device work): the BSR values of nn_assembly_evaluate (copied
 * to the host) regrouped by column, in the order of the reference's one
 * MatSetValuesBlocked(J, 1 + n_conn, rows, 1, &cell, values, ADD_VALUES) call per
 * cell (natural_petsc_runtime_jacobian.inc:279-299): for cell c the rows are
 * [c, its face neighbours in connection order] (rows[column_offsets[c] ..
 * column_offsets[c+1])) and the values the blocks (row, c), each np x np
 * row-major, in the same order.  values: nnzb x np x np, rows: nnzb,
 * column_offsets: ncell + 1.  (A caller that creates the BAIJ matrix from the
 * cached pattern can hand the unregrouped values to
 * MatSeqBAIJSetPreallocationCSR with MAT_ROW_ORIENTED: they are already
 * nnzb x np x np row-major blocks in rowOffsets/columnBlocks order.) */
int nn_bsr_cell_columns(int64_t ncell, int32_t np, const int64_t *conn_offsets,
                        const int64_t *conn_neighbor, const int64_t *row_offsets,
                        const int64_t *col_blocks, const double *bsr, double *values,
                        int64_t *rows, int64_t *column_offsets);

/* ---- wells ------------------------------------------------------------------
 * A well of the Natural runtime (well/specification.hpp WellSpecification,
 * petsc/well_runtime.hpp): Peaceman-type perforation sources in the mass
 * conservation rows of the perforated cells (physics/well_source.hpp:19-329),
 * with the bottom-hole pressure the well-pressure unknown of its
 * representative cell `bhp_cell`, whose well-control row holds the control
 * equation.  perforation_begin/end index the perforation arrays passed with it
 * (cell, well index WI).  Only bottom-hole-pressure control is built
 * (control = NN_WELL_CONTROL_BHP, R_w = p_bhp - target); rate controls are
 * refused.  Injection fractions are used when a phase flows into the reservoir. */
#define NN_WELL_INJECTOR 0          /* well/types.hpp:16-20 order */
#define NN_WELL_PRODUCER 1
#define NN_WELL_CONTROL_BHP 0
typedef struct nn_well {
    int32_t type;
    int32_t control;
    int64_t bhp_cell;
    double target;                                      /* Pa */
    double injection_phase_fraction[3];                 /* oil, gas, water */
    double injection_component_mass_fraction[NN_MAX_NC];
    int64_t perforation_begin;
    int64_t perforation_end;
} nn_well;

/* Host routine: the cached BSR pattern with the wells' couplings added
 * (jacobian_factory.hpp:65-155): the row of each well's bhp_cell gains the
 * column blocks of its perforated cells, each perforated cell's row the
 * bhp_cell's column block; rows stay sorted and unique.  Pass NULL outputs to
 * query the new block count in *out_nnzb. */
int nn_pattern_add_wells(int64_t ncell, const int64_t *row_offsets, const int64_t *col_blocks,
                         int32_t nwells, const nn_well *wells, const int64_t *perforation_cells,
                         int64_t *out_row_offsets, int64_t *out_col_blocks, int64_t *out_nnzb);

/* ---- state samples and their CSV text ------------------------------------ */

/* Output sample of a cell-state array on the device, as the reference's
 * CpGridOutputAccess::sample (output/include/output/petsc/cpgrid_output_access.hpp:
 * 42-140) takes it from a PETSc Vec: d_out[i * nvar + k] = d_vec[d_cells[i] *
 * dof_per_cell + d_vars[k]].  d_vec: ncell x dof_per_cell (a primary or
 * phase-state array); d_cells (nsel, int64) and d_vars (nvar, int32) index it.
 * An index outside the array writes NaN to its value and sets *d_bad (the
 * reference throws std::out_of_range); at most INT_MAX values per sample.
 * Asynchronous on `stream`. */
int nn_state_sample(const double *d_vec, int64_t ncell, int32_t dof_per_cell,
                    const int64_t *d_cells, int64_t nsel, const int32_t *d_vars, int32_t nvar,
                    double *d_out, int32_t *d_bad, void *stream);

/* Host routines (no device work): the text of the reference's sample CSV files
 * (solution_samples.csv, phase_state_samples.csv and the mass sample files of
 * OutputWriter, output/include/output/core/writer.hpp:58-93).  The header is
 * "value_0,...,value_{count-1}\n" (format.hpp:19-29), a row the values printed
 * with 16 significant digits, comma-separated, "\n"-terminated (format.hpp:
 * 32-42).  *length receives the text length without the terminating NUL; with
 * buf NULL only the length is computed, else cap must exceed it.
 * nn_csv_parse_row reads one such row back (count values into values[cap]). */
int nn_csv_sample_header(int64_t count, char *buf, int64_t cap, int64_t *length);
int nn_csv_sample_row(const double *values, int64_t count, char *buf, int64_t cap,
                      int64_t *length);
int nn_csv_parse_row(const char *line, double *values, int64_t cap, int64_t *count);

typedef struct nn_assembly nn_assembly;

/* Uploads topology (device pointers, kept by reference, must outlive the
 * handle) and precomputes, per connection, the BSR slot of its reciprocal block
 * and, per cell, the slot of its diagonal block.  Synchronous.  Any component
 * count a context accepts (2, 3, 4 and 6). */
int nn_assembly_create(nn_context *ctx, int64_t ncell, int32_t has_well_unknown,
                       const int64_t *d_conn_offsets, const int64_t *d_conn_neighbor,
                       const double *d_conn_trans, const double *d_conn_gdz,
                       const int64_t *d_row_offsets, const int64_t *d_col_blocks,
                       int64_t nnzb, const double *d_volume,
                       const double *d_porosity, const double *d_T,
                       nn_assembly **out);
int nn_assembly_destroy(nn_assembly *a);

/* R6: accepted-step accumulation M^n (per cell nc doubles) from a state
 * (natural_petsc_runtime_cache_history.inc:177-260). */
int nn_assembly_accumulation(nn_assembly *a, const double *d_primary,
                             const double *d_phase_state, double *d_accum,
                             void *stream);

/*
 * R2 + R3: residual F (ncell x np) and BSR Jacobian values (nnzb x np x np,
 * row-major blocks, in the cached pattern's rowOffsets/columnBlocks order),
 * exactly the values the reference writes with MatSetValuesBlocked
 * (with the handle's scaling, nn_assembly_set_scaling; none by default).
 * d_residual or d_bsr may be NULL to skip that part.  d_decisions: nconn x 3
 * uint8 upwind flags (0 = interior upwind), may be NULL.
 */
int nn_assembly_evaluate(nn_assembly *a, double dt, double fugacity_scaling,
                         const double *d_primary, const double *d_phase_state,
                         const double *d_accum_prev, double *d_residual,
                         double *d_bsr, uint8_t *d_decisions,
                         int32_t *d_status, void *stream);

/*
 * Equation and variable scaling of the Natural system (NaturalScalingOptions,
 * models/include/natural/petsc/natural_scaling.hpp:20-47; factors
 * models/include/natural/petsc/detail/natural_petsc_runtime_scaling.inc:8-89,
 * applied as natural_petsc_runtime_jacobian.inc:196-213, 262-268 and
 * natural_petsc_runtime_common.inc:345-356).  With `enabled`, every later
 * nn_assembly_evaluate on the handle writes
 *   residual[c][r]  = R_r * rowScale(r, c)
 *   J[r][q]         = (dR_r/dx_q * rowScale(r, c)) * varScale(q)
 * where rowScale is 1/mass_residual_scale on conservation rows,
 * 1/fugacity_residual_scale on both fugacity blocks, 1/closure_residual_scale on
 * the closure row and, on the well-control row, d_well_row_scale[c] when given
 * (the reference's representative-well rule: 1/max(|target|,
 * rate_well_residual_floor) for a cell of an active rate-controlled well,
 * 1/pressure_scale otherwise), else 1/pressure_scale; varScale is
 * pressure_scale on p and the well pressure, saturation_scale on saturations,
 * composition_scale on mole fractions.  d_well_row_scale (ncell doubles, device)
 * is kept by reference.  Disabled (the default) = every factor 1.
 */
typedef struct nn_scaling {
    int32_t enabled;                  /* 0 */
    double pressure_scale;            /* 1e7 Pa */
    double composition_scale;         /* 1 */
    double saturation_scale;          /* 1 */
    double mass_residual_scale;       /* 1 */
    double fugacity_residual_scale;   /* 1 */
    double closure_residual_scale;    /* 1 */
    double rate_well_residual_floor;  /* 1 (used by the caller's well-row rule) */
} nn_scaling;
void nn_scaling_default(nn_scaling *s);
int nn_assembly_set_scaling(nn_assembly *a, const nn_scaling *s, const double *d_well_row_scale);

/* Newton-step limiter of the Natural runtime, in place on the device
 * (replaces NaturalPetscRuntime::limitNewtonStep, natural_petsc_runtime_state.inc:
 * 452-545, with state/newton_limiter.hpp:124-245).  d_step holds PETSc's
 * correction Y for every cell (ncell x np, the primary layout), d_primary the
 * current solution.  Per cell, on delta = -Y: with the handle's scaling
 * enabled (nn_assembly_set_scaling) the scaled increment is recovered
 * (delta[q] *= variable scale q); the three saturation increments are limited
 * together to |dS| <= 0.1, each phase's composition block (its independent
 * fractions and the dependent last component) to 0.1, the pressure to 25 % of
 * |p|; with max_aqueous_unsupported >= 0 the Water composition block is then
 * scaled so that the Water role's mole fraction of components other than
 * water_component and co2_component stays below that ceiling
 * (limitAqueousCompositionNewtonIncrement, maximumAqueousUnsupportedNewton
 * MoleFraction of natural_runtime_options.hpp:49).  d_step receives -delta.
 * Asynchronous on `stream`. */
int nn_newton_limit(nn_assembly *a, const double *d_primary, double *d_step,
                    double max_aqueous_unsupported, int32_t water_component,
                    int32_t co2_component, void *stream);

/* Wells of the handle's later evaluations (NaturalPetscRuntime::setWells and
 * the well parts of formFunction / formJacobian: natural_petsc_runtime_wells.inc,
 * residual.inc:44-121, jacobian.inc:301-457).  Host arrays, copied; nwells = 0
 * removes them.  Needs a well unknown per cell (has_well_unknown) and a
 * pattern holding the wells' couplings (nn_pattern_add_wells).  In every later
 * nn_assembly_evaluate: each perforated cell's mass rows get
 * R_i += outward flux_i - sum over its perforations of q_i (in well, then
 * perforation order), each bhp_cell's well-control row R = p_bhp - target, and
 * the Jacobian gains the perforations' derivatives: columns of the perforated
 * cell, except the well-pressure column, which goes to the block (cell,
 * bhp_cell); pattern blocks no face or diagonal term covers are written as
 * zeros first.  surface_density: oil, gas, water (kg/m3), > 0. */
int nn_assembly_set_wells(nn_assembly *a, int32_t nwells, const nn_well *wells,
                          int64_t nperforations, const int64_t *perforation_cells,
                          const double *perforation_well_index, const double *surface_density);

/* Sizes of an assembly handle: its context, cells, BSR blocks and primary
 * variables per cell (no counterpart in the reference). */
int nn_assembly_info(const nn_assembly *a, nn_context **ctx, int64_t *ncell,
                     int64_t *nnzb, int32_t *np);

/* ---- serving: continuous batching of Newton-step requests (no counterpart in
 * the reference, which assembles one simulation per process) ----------------
 *
 * Ensemble members keep their state on the device (primary + phase-state
 * vectors: each cell's K-values and phase hysteresis) and submit steps: the
 * phase update (nn_phase_update_batch) then residual + BSR Jacobian
 * (nn_assembly_evaluate) of the member's grid.  A worker thread forms batches
 * continuously: whenever fewer than two batches are in flight it takes pending
 * steps in submission order (at most max_batch_steps, members with no step in
 * flight, output buffers within the memory budget), merges the cells of steps
 * with the same fluid_group into ONE phase-update launch (gather into a staging
 * buffer, update, scatter back), then runs each step's assembly.  Steps that
 * arrive while a batch runs join the next one.  Everything is issued on one
 * stream (the one given at creation, NULL = default): kernels whose lanes need
 * tens of KB of scratch could not be given scratch on additional hardware
 * queues, and merged launches fill the GPU without them.  The phase update's
 * arithmetic is per cell; results are bit-identical alone and in the mix in the
 * tested cases (members of up to a few thousand cells), but in large merged
 * launches some five- and six-component steps have come out differently from
 * the one-step form (the device build's results have depended on which cells
 * share a wavefront, cause not identified), so bitwise order independence is
 * not promised there.  Steps of one member
 * run in submission order.  Output buffers belong to the step until
 * nn_server_release.  max_batch_steps = 1 gives the plain one-request-at-a-time
 * form.  fluid_group: steps whose members share fluid and flash options may
 * carry the same group id >= 0 (the first step's context runs the merged
 * update); -1 never merges.
 */
typedef struct nn_server nn_server;
typedef struct nn_step_request {
    nn_assembly *assembly;       /* member's grid; its context runs the phase update */
    int64_t member;              /* member id (serialises that member's steps)       */
    int64_t fluid_group;         /* merge key of the phase update, -1 = alone        */
    int32_t has_well_unknown;
    int32_t want_bsr;            /* 0: residual only                                 */
    double dt, fugacity_scaling;
    const double *d_T;           /* per-cell temperature                             */
    double *d_primary;           /* member state, updated in place by the phase update */
    double *d_phase_state;
    const double *d_accum_prev;  /* accepted-step accumulation                       */
} nn_step_request;
typedef struct nn_step_result {
    int32_t rc;                  /* NN_OK or the failing call's code                 */
    int32_t np;
    int64_t ncell, nnzb;
    double submit_s, start_s, end_s;  /* server clock: submission, launch, completion seen */
    int64_t batch;               /* id of the batch the step ran in                  */
    double *d_residual;          /* ncell x np                                       */
    double *d_bsr;               /* nnzb x np x np (NULL without want_bsr)           */
    int32_t *d_update_status;    /* ncell NN_STATUS_* of the phase update            */
    int32_t *d_assembly_status;  /* ncell                                            */
    uint8_t *d_update_flags;     /* ncell                                            */
} nn_step_result;
typedef struct nn_server_stats {
    int64_t completed, pending, batches;
    int64_t inflight, peak_inflight;   /* steps in flight                          */
    int64_t peak_batch_steps, peak_merged_cells;
    int64_t memory_in_use, peak_memory_in_use, cached_bytes;
} nn_server_stats;
int nn_server_create(int32_t max_batch_steps, int64_t memory_budget_bytes, void *stream,
                     nn_server **out);
int nn_server_destroy(nn_server *s);   /* finishes every submitted step first */
int nn_server_submit_step(nn_server *s, const nn_step_request *req, int64_t *ticket);
int nn_server_wait(nn_server *s, int64_t ticket, nn_step_result *res);  /* blocks */
int nn_server_release(nn_server *s, int64_t ticket);
int nn_server_get_stats(nn_server *s, nn_server_stats *out);
/* hold != 0: issue no new batch (batches already issued run to completion) until
 * called again with hold = 0; steps submitted meanwhile wait and then form batches
 * together.  No counterpart in the reference. */
int nn_server_hold(nn_server *s, int32_t hold);
const char *nn_server_last_error(void);

/* ---- learned phase-count / log-K initialiser (no counterpart in the reference) ----
 *
 * MLP: features (ln P - 15.5)/1.5, (T - 450)/150, z[0..nc-1] -> three hidden
 * layers of 64 SiLU units -> 3 phase-count logits (1, 2, 3 phases) + nc values of
 * ln K_go.  FP32 training with the library's own forward/backward kernels and
 * Adam (beta 0.9/0.999, eps 1e-8); BF16 inference (weights and activations
 * rounded to BF16, FP32 accumulation).  Parameters: per layer W (out x in,
 * row-major) then b (out), layers in order, FP32 on the device.
 * Loss per state, averaged over the batch:
 *   -ln softmax(o[0..2])[label] + w * mask * (1/nc) sum_k (o[3+k] - logk[k])^2.
 */
#define NN_MLP_MAX_BATCH (1 << 20)
typedef struct nn_mlp nn_mlp;
int nn_mlp_create(int32_t nc, uint64_t seed, nn_mlp **out);   /* nc in {3, 4, 6} */
int nn_mlp_destroy(nn_mlp *m);
int64_t nn_mlp_param_count(const nn_mlp *m);
float *nn_mlp_params(nn_mlp *m);                 /* device pointer, param_count floats */
int nn_mlp_reset_optimizer(nn_mlp *m);            /* zero Adam moments and step count */
/* copy the parameters out to / in from a device array of param_count floats */
int nn_mlp_get_params(nn_mlp *m, float *d_out, void *stream);
int nn_mlp_set_params(nn_mlp *m, const float *d_in, void *stream);
const char *nn_mlp_last_error(void);
/* d_X: n x (nc+2) FP32 features of flash inputs */
int nn_mlp_features(nn_mlp *m, int64_t n, const double *d_P, const double *d_T,
                    const double *d_z, float *d_X, void *stream);
/* Loss (d_loss[0]) and its gradient w.r.t. the parameters (d_grad), no update.
 * d_label: class 0..2 (= phase count - 1); d_logk n x nc targets; d_mask: 1 where
 * the log-K target applies.  n <= NN_MLP_MAX_BATCH. */
int nn_mlp_loss_grad(nn_mlp *m, int64_t n, const float *d_X, const int32_t *d_label,
                     const float *d_logk, const uint8_t *d_mask, float logk_weight,
                     float *d_loss, float *d_grad, void *stream);
/* One Adam step on a mini-batch (loss before the step in d_loss[0]). */
int nn_mlp_train_step(nn_mlp *m, int64_t n, const float *d_X, const int32_t *d_label,
                      const float *d_logk, const uint8_t *d_mask, float logk_weight,
                      float lr, float *d_loss, void *stream);
/* BF16 inference: d_prob n x 3 class probabilities, d_logk n x nc. */
int nn_mlp_predict(nn_mlp *m, int64_t n, const float *d_X, float *d_prob, float *d_logk,
                   void *stream);
/* Flash hints from predictions: NN_HINT_SINGLE_PHASE where p(1 phase) >= threshold
 * and modes has that bit, else NN_HINT_LOG_K where p(2) or p(3) >= threshold and
 * modes has that bit, else NN_HINT_NONE; with NN_HINT_ORDER in modes, bits 2-3
 * carry the most probable phase count - 1; d_seed_log_k (n x nc FP64) =
 * predicted log-K. */
int nn_mlp_hints(nn_mlp *m, int64_t n, const float *d_prob, const float *d_logk,
                 float threshold, int32_t modes, uint8_t *d_hint, double *d_seed_log_k,
                 void *stream);

#ifdef __cplusplus
}
#endif

#endif /* NONENONE_AMD_H */
