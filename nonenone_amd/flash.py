"""Batched three-phase flash on the GPU: the host mirror of CubicThreePhaseFlash.

Reference surface (models/include/natural/thermo/three_phase_flash.hpp):
    CubicThreePhaseFlash(eos, options)      -> ThreePhaseFlash(fluid, options)
    flash(P, T, z)                          -> flash(P, T, z)            [batched]
    flashRestricted(P, T, z, allowed[, x])  -> flash_restricted(...)     [batched]
    stabilityTest(P, T, z, active, x)       -> stability_test(...)       [batched]
Inputs are torch tensors (any device; moved to cuda) of shape (n,), (n,), (n, nc).
Results are device tensors in the reference's slot order (0 oil, 1 gas, 2 water).
Every call runs HIP kernels from libnonenone_amd.so; nothing is computed on the CPU.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _abi
from ._abi import FlashOptions, FlashOut, FluidDesc


def _dev(t, dtype, device="cuda"):
    return torch.as_tensor(t, dtype=dtype).to(device).contiguous()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@dataclass
class FlashBatchResult:
    """ThreePhaseFlashResult as structure-of-arrays device tensors."""

    presence: torch.Tensor   # (n,) uint8, bit0 oil, bit1 gas, bit2 water
    converged: torch.Tensor  # (n,) uint8
    status: torch.Tensor     # (n,) int32 NN_STATUS_*
    iterations: torch.Tensor  # (n,) int32
    beta: torch.Tensor       # (n, 3)
    saturation: torch.Tensor  # (n, 3)
    x: torch.Tensor          # (n, 3, nc)
    Z: torch.Tensor          # (n, 3)
    rho_m: torch.Tensor      # (n, 3)
    k_go: torch.Tensor       # (n, nc)
    k_wo: torch.Tensor       # (n, nc)

    @classmethod
    def empty(cls, n: int, nc: int, device="cuda") -> "FlashBatchResult":
        f64 = dict(dtype=torch.float64, device=device)
        return cls(torch.zeros(n, dtype=torch.uint8, device=device),
                   torch.zeros(n, dtype=torch.uint8, device=device),
                   torch.zeros(n, dtype=torch.int32, device=device),
                   torch.zeros(n, dtype=torch.int32, device=device),
                   torch.zeros(n, 3, **f64), torch.zeros(n, 3, **f64),
                   torch.zeros(n, 3, nc, **f64), torch.zeros(n, 3, **f64),
                   torch.zeros(n, 3, **f64), torch.zeros(n, nc, **f64),
                   torch.zeros(n, nc, **f64))

    def abi(self) -> FlashOut:
        return FlashOut(*(self._p(t) for t in (
            self.presence, self.converged, self.status, self.iterations, self.beta,
            self.saturation, self.x, self.Z, self.rho_m, self.k_go, self.k_wo)))

    @staticmethod
    def _p(t):
        return t.data_ptr()


class Context:
    """Owns an nn_context (fluid parameters resident on the current HIP device)."""

    def __init__(self, fluid: FluidDesc, options: FlashOptions | None = None):
        self.lib = _abi.lib()
        if not torch.cuda.is_available():
            raise RuntimeError("no HIP device: this library runs on the GPU only")
        self.fluid = fluid
        self.options = options or FlashOptions.defaults()
        self.nc = fluid.nc
        h = ctypes.c_void_p()
        _abi.check(self.lib.nn_context_create(ctypes.byref(fluid), ctypes.byref(self.options),
                                              ctypes.byref(h)), "nn_context_create")
        self.handle = h

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.nn_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_kernel_timing(self, enable: bool) -> None:
        """Bracket every launch with HIP events on its stream (clears earlier records)."""
        _abi.check(self.lib.nn_context_set_kernel_timing(self.handle, int(bool(enable))),
                   "nn_context_set_kernel_timing")

    def schedule(self) -> dict:
        """Launch shapes of this context's flash-like batches (nn_schedule)."""
        s = _abi.Schedule()
        _abi.check(self.lib.nn_context_get_schedule(self.handle, ctypes.byref(s)),
                   "nn_context_get_schedule")
        return s.as_dict()

    def set_schedule(self, **fields) -> dict:
        """Change some nn_schedule fields (persistent_blocks_per_cu, coop_blocks_per_cu,
        ordered, wave_claim, plain); returns the schedule now in force."""
        cur = self.schedule()
        unknown = set(fields) - set(cur)
        if unknown:
            raise KeyError(f"unknown schedule fields {sorted(unknown)}")
        cur.update({k: int(v) for k, v in fields.items()})
        s = _abi.Schedule(**cur)
        _abi.check(self.lib.nn_context_set_schedule(self.handle, ctypes.byref(s)),
                   "nn_context_set_schedule")
        return self.schedule()

    def queue_counts(self) -> dict:
        """Work counters of the last flash / phase-update launch: items handed to the
        wavefront-per-item pass, recorded for the lockstep trials, held for re-flash."""
        v = [ctypes.c_int64() for _ in range(3)]
        _abi.check(self.lib.nn_context_queue_counts(self.handle, *(ctypes.byref(x) for x in v)),
                   "nn_context_queue_counts")
        return {"deferred": v[0].value, "recorded": v[1].value, "reflashed": v[2].value}

    def kernel_time(self, kind: str) -> tuple[float, int]:
        """(summed milliseconds, launches) of one kernel kind since timing was enabled."""
        ms = ctypes.c_double()
        count = ctypes.c_int64()
        _abi.check(self.lib.nn_context_kernel_time(self.handle, _abi.KERNEL_KINDS[kind],
                                                   ctypes.byref(ms), ctypes.byref(count)),
                   "nn_context_kernel_time")
        return ms.value, count.value


class ThreePhaseFlash(Context):
    """Batched CubicThreePhaseFlash (PR / PR78 / Soreide-Whitson)."""

    def eos_phase(self, P, T, x, role):
        """CubicEquationOfState::phaseResult(P, T, x, role) -> (Z, lnphi, fugacity, status)."""
        P, T = _dev(P, torch.float64), _dev(T, torch.float64)
        x = _dev(x, torch.float64)
        role = _dev(role, torch.int32)
        n = P.shape[0]
        Z = torch.empty(n, dtype=torch.float64, device=P.device)
        lnphi = torch.empty(n, self.nc, dtype=torch.float64, device=P.device)
        fug = torch.empty_like(lnphi)
        status = torch.empty(n, dtype=torch.int32, device=P.device)
        _abi.check(self.lib.nn_eos_phase_batch(self.handle, n, _ptr(P), _ptr(T), _ptr(x),
                                               _ptr(role), _ptr(Z), _ptr(lnphi), _ptr(fug),
                                               _ptr(status), _stream()), "nn_eos_phase_batch")
        return Z, lnphi, fug, status

    def flash(self, P, T, z, out: FlashBatchResult | None = None) -> FlashBatchResult:
        """CubicThreePhaseFlash::flash on every state."""
        P, T, z = _dev(P, torch.float64), _dev(T, torch.float64), _dev(z, torch.float64)
        n = P.shape[0]
        out = out or FlashBatchResult.empty(n, self.nc, P.device)
        abi = out.abi()
        _abi.check(self.lib.nn_flash_batch(self.handle, n, _ptr(P), _ptr(T), _ptr(z),
                                           ctypes.byref(abi), _stream()), "nn_flash_batch")
        return out

    def flash_seeded(self, P, T, z, hint, log_k=None,
                     out: FlashBatchResult | None = None) -> FlashBatchResult:
        """flash with learned-initialiser hints (NN_HINT_* per state, log-K seeds n x nc)."""
        P, T, z = _dev(P, torch.float64), _dev(T, torch.float64), _dev(z, torch.float64)
        hint = _dev(hint, torch.uint8)
        log_k = _dev(log_k, torch.float64) if log_k is not None else None
        n = P.shape[0]
        out = out or FlashBatchResult.empty(n, self.nc, P.device)
        abi = out.abi()
        _abi.check(self.lib.nn_flash_seeded_batch(self.handle, n, _ptr(P), _ptr(T), _ptr(z),
                                                  _ptr(hint), _ptr(log_k), ctypes.byref(abi),
                                                  _stream()), "nn_flash_seeded_batch")
        return out

    def autotune(self, P, T, z, repeats: int = 2) -> tuple[dict, float]:
        """nn_flash_autotune on a sample batch: keeps and returns the fastest schedule
        and its time in ms."""
        P, T, z = _dev(P, torch.float64), _dev(T, torch.float64), _dev(z, torch.float64)
        n = P.shape[0]
        out = FlashBatchResult.empty(n, self.nc, P.device)
        abi = out.abi()
        best = _abi.Schedule()
        ms = ctypes.c_double()
        _abi.check(self.lib.nn_flash_autotune(self.handle, n, _ptr(P), _ptr(T), _ptr(z),
                                              ctypes.byref(abi), int(repeats), ctypes.byref(best),
                                              ctypes.byref(ms), _stream()), "nn_flash_autotune")
        return best.as_dict(), ms.value

    def flash_restricted(self, P, T, z, allowed, seed=None) -> FlashBatchResult:
        """CubicThreePhaseFlash::flashRestricted (optionally with composition seeds)."""
        P, T, z = _dev(P, torch.float64), _dev(T, torch.float64), _dev(z, torch.float64)
        allowed = _dev(allowed, torch.uint8)
        seed = _dev(seed, torch.float64) if seed is not None else None
        n = P.shape[0]
        out = FlashBatchResult.empty(n, self.nc, P.device)
        abi = out.abi()
        _abi.check(self.lib.nn_flash_restricted_batch(self.handle, n, _ptr(P), _ptr(T), _ptr(z),
                                                      _ptr(allowed), _ptr(seed), ctypes.byref(abi),
                                                      _stream()), "nn_flash_restricted_batch")
        return out

    def stability_test(self, P, T, z, active, x):
        """CubicThreePhaseFlash::stabilityTest -> dict of device tensors."""
        P, T, z = _dev(P, torch.float64), _dev(T, torch.float64), _dev(z, torch.float64)
        active = _dev(active, torch.uint8)
        x = _dev(x, torch.float64)
        n = P.shape[0]
        dev = P.device
        r = dict(stable=torch.empty(n, dtype=torch.uint8, device=dev),
                 valid=torch.empty(n, dtype=torch.uint8, device=dev),
                 unstable_mask=torch.empty(n, dtype=torch.uint8, device=dev),
                 trial_sum=torch.empty(n, 3, dtype=torch.float64, device=dev),
                 incipient=torch.empty(n, 3, self.nc, dtype=torch.float64, device=dev),
                 status=torch.empty(n, dtype=torch.int32, device=dev))
        _abi.check(self.lib.nn_stability_batch(
            self.handle, n, _ptr(P), _ptr(T), _ptr(z), _ptr(active), _ptr(x),
            _ptr(r["stable"]), _ptr(r["valid"]), _ptr(r["unstable_mask"]), _ptr(r["trial_sum"]),
            _ptr(r["incipient"]), _ptr(r["status"]), _stream()), "nn_stability_batch")
        return r

    def update_phase_state(self, T, primary, phase_state, has_well_unknown=False):
        """FullyCompositionalThreePhaseEquilibrium::updatePhaseState on every cell, in place."""
        T = _dev(T, torch.float64)
        n = T.shape[0]
        status = torch.empty(n, dtype=torch.int32, device=T.device)
        flags = torch.empty(n, dtype=torch.uint8, device=T.device)
        assert primary.is_contiguous() and phase_state.is_contiguous()
        _abi.check(self.lib.nn_phase_update_batch(self.handle, n, int(bool(has_well_unknown)),
                                                  _ptr(T), _ptr(primary), _ptr(phase_state),
                                                  _ptr(status), _ptr(flags), _stream()),
                   "nn_phase_update_batch")
        return status, flags

    def update_state_with_report(self, T, primary, phase_state, has_well_unknown=False):
        """NaturalPetscRuntime::updateStateWithReport (natural_petsc_runtime_state.inc:166-334):
        every cell updated on copies, committed over primary / phase_state only when
        no cell fails.  Returns (status, flags, report dict)."""
        T = _dev(T, torch.float64)
        n = T.shape[0]
        assert primary.is_contiguous() and phase_state.is_contiguous()
        status = torch.empty(n, dtype=torch.int32, device=T.device)
        flags = torch.empty(n, dtype=torch.uint8, device=T.device)
        work_p = torch.empty_like(primary)
        work_s = torch.empty_like(phase_state)
        report = torch.empty(ctypes.sizeof(_abi.UpdateReport), dtype=torch.uint8, device=T.device)
        _abi.check(self.lib.nn_phase_update_transaction(
            self.handle, n, int(bool(has_well_unknown)), _ptr(T), _ptr(primary),
            _ptr(phase_state), _ptr(work_p), _ptr(work_s), _ptr(status), _ptr(flags),
            _ptr(report), _stream()), "nn_phase_update_transaction")
        r = _abi.UpdateReport.from_buffer_copy(report.cpu().numpy().tobytes())
        return status, flags, {name: int(getattr(r, name)) for name, _ in r._fields_
                               if name != "reserved"}
