"""ctypes mirror of include/nonenone_amd.h and the loader of the HIP library.

The product library is ``nonenone_amd/libnonenone_amd.so`` (built in-tree by
``nonenone_amd/build.py``).  There is no CPU fallback: :func:`lib` raises when
the library is missing, and every compute entry point launches HIP kernels.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

NN_MAX_NC = 8

NN_OK = 0
NN_E_INVALID_ARGUMENT = -1
NN_E_UNSUPPORTED = -2
NN_E_HIP = -3
NN_E_NO_DEVICE = -4
NN_E_OUT_OF_MEMORY = -5
NN_STATUS_OK = 0
NN_STATUS_NOT_CONVERGED = 1
NN_STATUS_THERMO_ERROR = 2
NN_STATUS_INVALID_INPUT = 3
NN_STATUS_RECOVERABLE_FAILURE = 4

NN_MODEL_PR = 0
NN_MODEL_SW = 1
NN_MODEL_CPA = 2

_D = ctypes.c_double
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_P = ctypes.c_void_p


class Well(ctypes.Structure):
    """nn_well."""

    _fields_ = [
        ("type", _I32),
        ("control", _I32),
        ("bhp_cell", _I64),
        ("target", _D),
        ("injection_phase_fraction", _D * 3),
        ("injection_component_mass_fraction", _D * NN_MAX_NC),
        ("perforation_begin", _I64),
        ("perforation_end", _I64),
    ]


NN_WELL_INJECTOR = 0
NN_WELL_PRODUCER = 1
NN_WELL_CONTROL_BHP = 0


class FluidDesc(ctypes.Structure):
    """nn_fluid_desc."""

    _fields_ = [
        ("nc", _I32),
        ("thermo_model", _I32),
        ("eos_type", _I32),
        ("water_component", _I32),
        ("omega_a", _D),
        ("omega_b", _D),
        ("m1", _D),
        ("m2", _D),
        ("tc", _D * NN_MAX_NC),
        ("pc", _D * NN_MAX_NC),
        ("vc", _D * NN_MAX_NC),
        ("acentric", _D * NN_MAX_NC),
        ("molar_mass", _D * NN_MAX_NC),
        ("kij", _D * (NN_MAX_NC * NN_MAX_NC)),
        ("sw_salinity", _D),
        ("sw_aqueous_kij", _D * NN_MAX_NC),
        ("sw_aqueous_kij_set", _I32 * NN_MAX_NC),
        ("cpa_physical_term", _I32),
        ("cpa_max_iterations", _I32),
        ("cpa_tolerance", _D),
        ("cpa_damping", _D),
        ("cpa_a0", _D * NN_MAX_NC),
        ("cpa_b", _D * NN_MAX_NC),
        ("cpa_c1", _D * NN_MAX_NC),
        ("cpa_epsilon", _D * NN_MAX_NC),
        ("cpa_beta", _D * NN_MAX_NC),
        ("cpa_donors", _I32 * NN_MAX_NC),
        ("cpa_acceptors", _I32 * NN_MAX_NC),
        ("cpa_cross_epsilon", _D * (NN_MAX_NC * NN_MAX_NC)),
        ("cpa_cross_beta", _D * (NN_MAX_NC * NN_MAX_NC)),
        ("relperm", _D * 7),
        ("aqueous_volume", _I32),
        ("aqueous_viscosity", _I32),
        ("co2_component", _I32),
        ("aqueous_reserved", _I32),
        ("water_molar_mass", _D),
        ("aqueous_volume_max_unsupported", _D),
        ("aqueous_viscosity_max_unsupported", _D),
    ]


class FlashOptions(ctypes.Structure):
    """nn_flash_options (ThreePhaseFlashOptions)."""

    _fields_ = [
        ("maximum_iterations", _I32),
        ("maximum_rachford_rice_iterations", _I32),
        ("maximum_stability_iterations", _I32),
        ("stability_ssi_before_newton", _I32),
        ("maximum_stability_newton", _I32),
        ("two_phase_ssi_before_newton", _I32),
        ("maximum_two_phase_newton", _I32),
        ("three_phase_ssi_before_newton", _I32),
        ("maximum_three_phase_newton", _I32),
        ("fugacity_tolerance", _D),
        ("rachford_rice_tolerance", _D),
        ("stability_tolerance", _D),
        ("phase_fraction_tolerance", _D),
        ("coincident_composition_tolerance", _D),
        ("composition_floor", _D),
        ("log_k_step_limit", _D),
        ("numerical_jacobian_step", _D),
    ]

    @classmethod
    def defaults(cls) -> "FlashOptions":
        """Reference defaults, three_phase_flash_types.hpp:21-41."""
        return cls(120, 80, 80, 8, 12, 8, 12, 10, 12,
                   1e-9, 1e-12, 1e-9, 1e-10, 1e-8, 1e-30, 2.0, 1e-6)


class FlashOut(ctypes.Structure):
    """nn_flash_out: device pointers."""

    _fields_ = [
        ("presence", _P),
        ("converged", _P),
        ("status", _P),
        ("iterations", _P),
        ("beta", _P),
        ("saturation", _P),
        ("x", _P),
        ("Z", _P),
        ("rho_m", _P),
        ("k_go", _P),
        ("k_wo", _P),
    ]


class UpdateReport(ctypes.Structure):
    """nn_update_report: StateUpdateReport of the all-or-none phase update
    (natural_petsc_runtime_state.inc:166-334)."""

    _fields_ = [
        ("first_failure_cell", _I64),
        ("updated_cells", _I64),
        ("stable_reduced_cells", _I64),
        ("succeeded", _I32),
        ("first_failure_status", _I32),
        ("first_failure_flags", _I32),
        ("reserved", _I32),
    ]


class Scaling(ctypes.Structure):
    """nn_scaling: the Natural runtime's equation/variable scaling
    (NaturalScalingOptions, models/include/natural/petsc/natural_scaling.hpp:20-47)."""

    _fields_ = [
        ("enabled", _I32),
        ("pressure_scale", _D),
        ("composition_scale", _D),
        ("saturation_scale", _D),
        ("mass_residual_scale", _D),
        ("fugacity_residual_scale", _D),
        ("closure_residual_scale", _D),
        ("rate_well_residual_floor", _D),
    ]

    @classmethod
    def defaults(cls) -> "Scaling":
        return cls(0, 1.0e7, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)


class Schedule(ctypes.Structure):
    """nn_schedule: launch shapes of the flash-like batches of a context."""

    _fields_ = [
        ("persistent_blocks_per_cu", _I32),
        ("coop_blocks_per_cu", _I32),
        ("ordered", _I32),
        ("wave_claim", _I32),
        ("plain", _I32),
    ]

    def as_dict(self) -> dict:
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


NN_HINT_NONE = 0
NN_HINT_SINGLE_PHASE = 1
NN_HINT_LOG_K = 2
NN_HINT_ORDER = 4
NN_MLP_MAX_BATCH = 1 << 20

_PRODUCT_LIB = None


def library_path() -> Path:
    """The in-tree build; NN_LIBRARY names another build of the same sources (a
    variant compiled with different build parameters, for timing)."""
    override = os.environ.get("NN_LIBRARY")
    if override:
        return Path(override).resolve()
    return Path(__file__).resolve().parent / "libnonenone_amd.so"


def lib() -> ctypes.CDLL:
    """Load the HIP library.  Raises when it is not built: no CPU fallback."""
    global _PRODUCT_LIB
    if _PRODUCT_LIB is not None:
        return _PRODUCT_LIB
    path = library_path()
    if not path.exists():
        raise RuntimeError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " (hipcc --offload-arch=gfx950); there is no CPU fallback")
    L = ctypes.CDLL(str(path), mode=os.RTLD_GLOBAL if hasattr(os, "RTLD_GLOBAL") else 0)
    _declare(L)
    _PRODUCT_LIB = L
    return L


def _declare(L: ctypes.CDLL) -> None:
    L.nn_flash_options_default.argtypes = [ctypes.POINTER(FlashOptions)]
    L.nn_flash_options_default.restype = None
    L.nn_context_create.argtypes = [ctypes.POINTER(FluidDesc), ctypes.POINTER(FlashOptions),
                                    ctypes.POINTER(_P)]
    L.nn_context_destroy.argtypes = [_P]
    L.nn_last_error.restype = ctypes.c_char_p
    L.nn_eos_phase_batch.argtypes = [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    L.nn_flash_batch.argtypes = [_P, _I64, _P, _P, _P, ctypes.POINTER(FlashOut), _P]
    L.nn_flash_batch_static.argtypes = [_P, _I64, _P, _P, _P, ctypes.POINTER(FlashOut), _P]
    L.nn_flash_restricted_batch.argtypes = [_P, _I64, _P, _P, _P, _P, _P,
                                            ctypes.POINTER(FlashOut), _P]
    L.nn_stability_batch.argtypes = [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    L.nn_phase_update_batch.argtypes = [_P, _I64, _I32, _P, _P, _P, _P, _P, _P]
    L.nn_structured_topology.argtypes = [_I32, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P,
                                         _P, _P, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]
    L.nn_assembly_create.argtypes = [_P, _I64, _I32, _P, _P, _P, _P, _P, _P, _I64, _P, _P, _P,
                                     ctypes.POINTER(_P)]
    L.nn_assembly_destroy.argtypes = [_P]
    L.nn_assembly_accumulation.argtypes = [_P, _P, _P, _P, _P]
    L.nn_assembly_evaluate.argtypes = [_P, _D, _D, _P, _P, _P, _P, _P, _P, _P, _P]
    L.nn_context_set_kernel_timing.argtypes = [_P, _I32]
    L.nn_context_kernel_time.argtypes = [_P, _I32, ctypes.POINTER(_D), ctypes.POINTER(_I64)]
    L.nn_context_queue_counts.argtypes = [_P] + [ctypes.POINTER(_I64)] * 3
    L.nn_flash_seeded_batch.argtypes = [_P, _I64, _P, _P, _P, _P, _P, ctypes.POINTER(FlashOut),
                                        _P]
    L.nn_context_get_schedule.argtypes = [_P, ctypes.POINTER(Schedule)]
    L.nn_context_set_schedule.argtypes = [_P, ctypes.POINTER(Schedule)]
    L.nn_flash_autotune.argtypes = [_P, _I64, _P, _P, _P, ctypes.POINTER(FlashOut), _I32,
                                    ctypes.POINTER(Schedule), ctypes.POINTER(_D), _P]
    L.nn_mlp_create.argtypes = [_I32, ctypes.c_uint64, ctypes.POINTER(_P)]
    L.nn_mlp_destroy.argtypes = [_P]
    L.nn_mlp_param_count.argtypes = [_P]
    L.nn_mlp_param_count.restype = _I64
    L.nn_mlp_params.argtypes = [_P]
    L.nn_mlp_params.restype = _P
    L.nn_mlp_reset_optimizer.argtypes = [_P]
    L.nn_mlp_get_params.argtypes = [_P, _P, _P]
    L.nn_mlp_set_params.argtypes = [_P, _P, _P]
    L.nn_mlp_last_error.restype = ctypes.c_char_p
    L.nn_mlp_features.argtypes = [_P, _I64, _P, _P, _P, _P, _P]
    L.nn_mlp_loss_grad.argtypes = [_P, _I64, _P, _P, _P, _P, ctypes.c_float, _P, _P, _P]
    L.nn_mlp_train_step.argtypes = [_P, _I64, _P, _P, _P, _P, ctypes.c_float, ctypes.c_float, _P,
                                    _P]
    L.nn_mlp_predict.argtypes = [_P, _I64, _P, _P, _P, _P]
    L.nn_mlp_hints.argtypes = [_P, _I64, _P, _P, ctypes.c_float, _I32, _P, _P, _P]
    L.nn_phase_update_transaction.argtypes = [_P, _I64, _I32, _P, _P, _P, _P, _P, _P, _P, _P,
                                              _P]
    L.nn_bsr_cell_columns.argtypes = [_I64, _I32, _P, _P, _P, _P, _P, _P, _P, _P]
    L.nn_scaling_default.argtypes = [ctypes.POINTER(Scaling)]
    L.nn_scaling_default.restype = None
    L.nn_assembly_set_scaling.argtypes = [_P, ctypes.POINTER(Scaling), _P]
    L.nn_assembly_info.argtypes = [_P, ctypes.POINTER(_P), ctypes.POINTER(_I64),
                                   ctypes.POINTER(_I64), ctypes.POINTER(_I32)]
    L.nn_assembly_info.restype = ctypes.c_int
    L.nn_newton_limit.argtypes = [_P, _P, _P, _D, _I32, _I32, _P]
    L.nn_newton_limit.restype = ctypes.c_int
    L.nn_pattern_add_wells.argtypes = [_I64, _P, _P, _I32, _P, _P, _P, _P, ctypes.POINTER(_I64)]
    L.nn_pattern_add_wells.restype = ctypes.c_int
    L.nn_assembly_set_wells.argtypes = [_P, _I32, _P, _I64, _P, _P, _P]
    L.nn_assembly_set_wells.restype = ctypes.c_int
    L.nn_state_sample.argtypes = [_P, _I64, _I32, _P, _I64, _P, _I32, _P, _P, _P]
    L.nn_state_sample.restype = ctypes.c_int
    L.nn_csv_sample_header.argtypes = [_I64, ctypes.c_char_p, _I64, ctypes.POINTER(_I64)]
    L.nn_csv_sample_header.restype = ctypes.c_int
    L.nn_csv_sample_row.argtypes = [_P, _I64, ctypes.c_char_p, _I64, ctypes.POINTER(_I64)]
    L.nn_csv_sample_row.restype = ctypes.c_int
    L.nn_csv_parse_row.argtypes = [ctypes.c_char_p, _P, _I64, ctypes.POINTER(_I64)]
    L.nn_csv_parse_row.restype = ctypes.c_int
    for name in ("nn_context_set_kernel_timing", "nn_context_kernel_time",
                 "nn_context_queue_counts", "nn_context_create", "nn_context_destroy", "nn_eos_phase_batch",
                 "nn_flash_batch", "nn_flash_restricted_batch", "nn_stability_batch",
                 "nn_phase_update_batch", "nn_structured_topology", "nn_assembly_create",
                 "nn_assembly_destroy", "nn_assembly_accumulation", "nn_assembly_evaluate",
                 "nn_flash_seeded_batch", "nn_context_get_schedule", "nn_context_set_schedule",
                 "nn_flash_autotune", "nn_mlp_create", "nn_mlp_destroy",
                 "nn_mlp_reset_optimizer", "nn_mlp_features", "nn_mlp_loss_grad",
                 "nn_mlp_train_step", "nn_mlp_predict", "nn_mlp_hints", "nn_mlp_get_params",
                 "nn_mlp_set_params"):
        getattr(L, name).restype = ctypes.c_int


# Every symbol include/nonenone_amd.h declares.
EXPORTED_SYMBOLS = (
    "nn_flash_options_default", "nn_context_create", "nn_context_destroy", "nn_last_error",
    "nn_eos_phase_batch", "nn_flash_batch", "nn_flash_restricted_batch",
    "nn_stability_batch", "nn_phase_update_batch", "nn_structured_topology",
    "nn_assembly_create", "nn_assembly_destroy", "nn_assembly_accumulation",
    "nn_assembly_evaluate", "nn_context_set_kernel_timing", "nn_context_kernel_time",
    "nn_flash_seeded_batch", "nn_context_get_schedule", "nn_context_set_schedule",
    "nn_flash_autotune", "nn_mlp_create", "nn_mlp_destroy", "nn_mlp_param_count",
    "nn_mlp_params", "nn_mlp_reset_optimizer", "nn_mlp_last_error", "nn_mlp_features",
    "nn_mlp_loss_grad", "nn_mlp_train_step", "nn_mlp_predict", "nn_mlp_hints",
    "nn_mlp_get_params", "nn_mlp_set_params", "nn_assembly_info",
    "nn_server_create", "nn_server_destroy", "nn_server_submit_step", "nn_server_wait",
    "nn_server_release", "nn_server_get_stats", "nn_server_last_error", "nn_server_hold",
    "nn_flash_batch_static", "nn_scaling_default", "nn_assembly_set_scaling",
    "nn_phase_update_transaction", "nn_bsr_cell_columns", "nn_context_queue_counts",
    "nn_newton_limit", "nn_state_sample", "nn_csv_sample_header", "nn_csv_sample_row",
    "nn_csv_parse_row", "nn_pattern_add_wells", "nn_assembly_set_wells",
)

# Water-role viscosity closures (nn_fluid_desc.aqueous_viscosity).
NN_AQ_VISCOSITY_NONE = 0
NN_AQ_VISCOSITY_MCBRIDE_WRIGHT = 1
NN_AQ_VISCOSITY_IAPWS2008 = 2

# Kernel kinds of nn_context_kernel_time.
KERNEL_KINDS = {"eos": 0, "flash": 1, "stability": 2, "phase_update": 3, "cell_record": 4,
                "assemble": 5, "flash_coop": 6, "phase_update_coop": 7}


def check(rc: int, what: str) -> None:
    if rc != NN_OK:
        getter = lib().nn_mlp_last_error if what.startswith("nn_mlp") else lib().nn_last_error
        msg = getter()
        raise RuntimeError(f"{what} failed with code {rc}: {msg.decode() if msg else ''}")
