"""Serving: many ensemble members submit Newton-step requests to one GPU (no
counterpart in the reference, whose runtime assembles one simulation per process).

Host mirror of the nn_server C ABI (continuous batching in the library's native
worker thread: steps that arrive while a batch runs join the next batch, and the
cells of steps whose members share a fluid go through one merged phase-update
launch) plus the pieces a driver needs: ``Member`` (a device-resident ensemble
member: its own context, grid, assembly and state — the phase-state vector holds
each cell's K-values and hysteresis between requests) and ``run_trace`` (submit
steps at given arrival times, wait, collect latencies and bitwise checksums of
every result).
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi, fluids, workloads
from .assembly import NaturalAssembly, assign_flash_result, perturb_newton_state
from .flash import ThreePhaseFlash, _ptr

_D = ctypes.c_double
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_P = ctypes.c_void_p


class StepRequest(ctypes.Structure):
    _fields_ = [("assembly", _P), ("member", _I64), ("fluid_group", _I64), ("has_well_unknown", _I32),
                ("want_bsr", _I32), ("dt", _D), ("fugacity_scaling", _D), ("d_T", _P),
                ("d_primary", _P), ("d_phase_state", _P), ("d_accum_prev", _P)]


class StepResult(ctypes.Structure):
    _fields_ = [("rc", _I32), ("np", _I32), ("ncell", _I64), ("nnzb", _I64),
                ("submit_s", _D), ("start_s", _D), ("end_s", _D), ("batch", _I64), ("d_residual", _P),
                ("d_bsr", _P), ("d_update_status", _P), ("d_assembly_status", _P),
                ("d_update_flags", _P)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", _I64), ("pending", _I64), ("batches", _I64), ("inflight", _I64),
                ("peak_inflight", _I64), ("peak_batch_steps", _I64), ("peak_merged_cells", _I64),
                ("memory_in_use", _I64), ("peak_memory_in_use", _I64), ("cached_bytes", _I64)]


def declare(L) -> None:
    L.nn_server_create.argtypes = [_I32, _I64, _P, ctypes.POINTER(_P)]
    L.nn_server_destroy.argtypes = [_P]
    L.nn_server_submit_step.argtypes = [_P, ctypes.POINTER(StepRequest), ctypes.POINTER(_I64)]
    L.nn_server_wait.argtypes = [_P, _I64, ctypes.POINTER(StepResult)]
    L.nn_server_release.argtypes = [_P, _I64]
    L.nn_server_get_stats.argtypes = [_P, ctypes.POINTER(ServerStats)]
    L.nn_server_hold.argtypes = [_P, _I32]
    L.nn_server_last_error.restype = ctypes.c_char_p
    for name in ("nn_server_create", "nn_server_destroy", "nn_server_submit_step",
                 "nn_server_wait", "nn_server_release", "nn_server_get_stats", "nn_server_hold"):
        getattr(L, name).restype = ctypes.c_int


class _DeviceArray:
    """A view of server-owned device memory for torch (via __cuda_array_interface__)."""

    def __init__(self, ptr: int, shape, typestr: str):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr,
                                         "data": (int(ptr), False), "version": 2,
                                         "strides": None}


def device_view(ptr: int, shape, dtype: torch.dtype) -> torch.Tensor:
    typestr = {torch.float64: "<f8", torch.int32: "<i4", torch.uint8: "|u1"}[dtype]
    return torch.as_tensor(_DeviceArray(ptr, shape, typestr), device="cuda")


def checksum(t: torch.Tensor) -> int:
    """Bitwise checksum of a device tensor (sum of its 64-bit words, wrapping)."""
    b = t.contiguous().view(torch.uint8).flatten()
    pad = (-b.numel()) % 8
    if pad:
        b = torch.cat([b, torch.zeros(pad, dtype=torch.uint8, device=b.device)])
    w = b.view(torch.int64)
    # two half-sums keep every partial sum exact in int64 regardless of order
    lo = (w & 0xFFFFFFFF).sum()
    hi = (w >> 32).sum()
    return int(lo.item()) ^ (int(hi.item()) << 1)


class Server:
    """nn_server: continuous batching of member steps, at most `max_batch_steps` per
    batch, issued on torch's current stream."""

    def __init__(self, max_batch_steps: int, memory_budget_bytes: int):
        self.lib = _abi.lib()
        declare(self.lib)
        h = _P()
        stream = _P(torch.cuda.current_stream().cuda_stream)
        rc = self.lib.nn_server_create(int(max_batch_steps), int(memory_budget_bytes), stream,
                                       ctypes.byref(h))
        self._check(rc, "nn_server_create")
        self.handle = h

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.nn_server_last_error() or self.lib.nn_last_error()
            raise RuntimeError(f"{what} failed with code {rc}: {msg.decode() if msg else ''}")

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.nn_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, member_id: int, m: "Member", dt: float = workloads.DAY,
               want_bsr: bool = True) -> int:
        r = StepRequest(m.asm.handle, int(member_id), int(m.group), int(m.has_well),
                        int(bool(want_bsr)), float(dt), 1.0, _ptr(m.T), _ptr(m.prim), _ptr(m.ps),
                        _ptr(m.acc))
        t = _I64()
        self._check(self.lib.nn_server_submit_step(self.handle, ctypes.byref(r), ctypes.byref(t)),
                    "nn_server_submit_step")
        return t.value

    def wait(self, ticket: int) -> StepResult:
        res = StepResult()
        rc = self.lib.nn_server_wait(self.handle, int(ticket), ctypes.byref(res))
        self._check(rc, "nn_server_wait")
        return res

    def release(self, ticket: int) -> None:
        self._check(self.lib.nn_server_release(self.handle, int(ticket)), "nn_server_release")

    def hold(self, on: bool) -> None:
        """While held, the server issues no new batch; steps submitted meanwhile
        form batches together once released."""
        self._check(self.lib.nn_server_hold(self.handle, int(bool(on))), "nn_server_hold")

    def stats(self) -> dict:
        s = ServerStats()
        self._check(self.lib.nn_server_get_stats(self.handle, ctypes.byref(s)),
                    "nn_server_get_stats")
        return {name: int(getattr(s, name)) for name, _ in s._fields_}


MEMBER_FLUIDS = {
    "nc3-PR": lambda: fluids.h2o_co2_nc10("PR"),
    "nc3-SW": lambda: fluids.h2o_co2_nc10("SW"),
    "nc3-CPA": fluids.cpa_h2o_co2_nc10,
    "nc5-PR": lambda: fluids.five_component("PR"),
    "nc5-SW": lambda: fluids.five_component("SW"),
    "nc6-PR": lambda: fluids.six_component("PR"),
    "nc6-SW": lambda: fluids.six_component("SW"),
    "nc8-PR": lambda: fluids.eight_component("PR"),
    "nc8-SW": lambda: fluids.eight_component("SW"),
}


class Member:
    """One ensemble member resident on the device: its own context (work queue),
    structured nx x ny x 1 grid, assembly, initial state (history flash + seeded
    Newton perturbation) and accepted-step accumulation."""

    def __init__(self, kind: str, nx: int, ny: int, seed: int, has_well: bool = False):
        self.kind = kind
        # members of one kind share fluid and flash options: their phase updates may merge
        self.group = list(MEMBER_FLUIDS).index(kind)
        self.fluid = MEMBER_FLUIDS[kind]()
        self.nc = self.fluid.nc
        self.has_well = has_well
        self.fl = ThreePhaseFlash(self.fluid)
        w = workloads.grid_workload(nx, ny, 1, 10.0, 10.0, 5.0, self.fluid, seed=seed)
        self.ncell = w["ncell"]
        res = self.fl.flash(w["P"], w["T"], w["z"])
        flash = {k: getattr(res, k).cpu().numpy() for k in
                 ("presence", "status", "beta", "saturation", "x", "Z", "k_go", "k_wo")}
        primary, ps = assign_flash_result(flash, w["P"], has_well=has_well)
        pert = perturb_newton_state(primary, ps, self.nc, seed=seed + 1)
        self.asm = NaturalAssembly(self.fl, w["topo"], w["volume"], w["porosity"], w["T"],
                                   has_well_unknown=has_well)
        dev = torch.device("cuda")
        self.T = torch.as_tensor(w["T"], device=dev)
        self.acc = self.asm.accumulation(torch.as_tensor(primary, device=dev),
                                         torch.as_tensor(ps, device=dev))
        self.prim0 = torch.as_tensor(pert, device=dev)
        self.ps0 = torch.as_tensor(ps, device=dev)
        self.prim = self.prim0.clone()
        self.ps = self.ps0.clone()

    def reset(self) -> None:
        """Back to the initial state (the server updates the state in place)."""
        self.prim.copy_(self.prim0)
        self.ps.copy_(self.ps0)

    def output_bytes(self) -> int:
        np_ = self.asm.np_
        return self.ncell * np_ * 8 + self.asm.nnzb * np_ * np_ * 8 + self.ncell * 9


@dataclass
class TraceResult:
    latencies_ms: np.ndarray   # per request, completion - submission
    makespan_s: float
    cells: int
    checksums: list            # per request (member, step) -> (residual, bsr, statuses)
    stats: dict
    nonok: int


def run_trace(server: Server, members: list, arrivals, want_bsr: bool = True,
              hold_start: bool = False) -> TraceResult:
    """Submit one step of members[m] at each (t, m) of `arrivals` (seconds from the
    start, increasing), wait for all, and collect latencies and checksums.
    `hold_start`: the steps arriving at t = 0 are all queued before the server
    issues anything (so they batch together however the threads are scheduled)."""
    for m in members:
        m.reset()
    torch.cuda.synchronize()
    tickets = []
    if hold_start:
        server.hold(True)
    held = hold_start
    t0 = time.perf_counter()
    for t, mi in arrivals:
        if held and t > 0.0:
            server.hold(False)
            held = False
        delay = t - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        tickets.append((mi, server.submit(mi, members[mi], want_bsr=want_bsr)))
    if held:
        server.hold(False)
    lat, sums, cells, nonok = [], [], 0, 0
    step_of = {}
    end, first = 0.0, float("inf")
    for mi, tk in tickets:
        r = server.wait(tk)
        if r.rc != 0:
            raise RuntimeError(f"step of member {mi} failed with code {r.rc}")
        m = members[mi]
        lat.append((r.end_s - r.submit_s) * 1e3)
        end = max(end, r.end_s)
        first = min(first, r.submit_s)
        k = step_of.get(mi, 0)
        step_of[mi] = k + 1
        res = device_view(r.d_residual, (r.ncell, r.np), torch.float64)
        st = device_view(r.d_update_status, (r.ncell,), torch.int32)
        ast = device_view(r.d_assembly_status, (r.ncell,), torch.int32)
        entry = [checksum(res), checksum(st), checksum(ast)]
        if want_bsr:
            entry.append(checksum(device_view(r.d_bsr, (r.nnzb, r.np, r.np), torch.float64)))
        nonok += int((st != 0).sum().item())
        sums.append(((mi, k), tuple(entry)))
        cells += m.ncell
        server.release(tk)
    makespan = end - first
    return TraceResult(np.asarray(lat), makespan, cells, sums, server.stats(), nonok)


def poisson_arrivals(n_requests: int, n_members: int, rate: float, seed: int):
    """Poisson process of `rate` requests/s; each request picks a member uniformly."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    who = rng.integers(0, n_members, n_requests)
    return list(zip(t.tolist(), who.tolist()))
