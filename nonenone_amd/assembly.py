"""Natural residual / Jacobian assembly on the GPU: the host mirror of the
reference runtime's formFunction / formJacobian for one Newton state.

Reference surface (models/include/natural/petsc/natural_petsc_runtime.hpp):
    runtime.formFunction(snes, X, F)   -> NaturalAssembly.evaluate(..., want_bsr=False)
    runtime.formJacobian(X, J, P)      -> NaturalAssembly.evaluate(...) BSR values in the
                                          cached pattern order (jacobian_factory.hpp:187-201),
                                          ready for one bulk MatSetValuesBlocked
    updateStateWithReport per cell     -> ThreePhaseFlash.update_phase_state
    refreshPreviousPropertiesCache_    -> NaturalAssembly.accumulation
Layouts: primary (ncell, 3nc+1[+1]) and phase state (ncell, 3nc+8) exactly as the
reference's PETSc Vecs (indices/include/indices/indices.hpp:46-292).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _abi
from .flash import Context, _ptr, _stream


class NaturalAssembly:
    def __init__(self, ctx: Context, topo: dict, volume, porosity, T, has_well_unknown=False,
                 device="cuda"):
        self.ctx = ctx
        self.lib = ctx.lib
        self.nc = ctx.nc
        self.has_well = bool(has_well_unknown)
        self.np_ = 3 * self.nc + 1 + (1 if self.has_well else 0)
        self.nps = 3 * self.nc + 8
        self.ncell = len(topo["conn_offsets"]) - 1
        self.nnzb = int(topo["row_offsets"][-1])
        self.nconn = int(topo["conn_offsets"][-1])
        t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(device)  # noqa: E731
        self.d = dict(conn_offsets=t(topo["conn_offsets"], torch.int64),
                      conn_neighbor=t(topo["conn_neighbor"], torch.int64),
                      conn_trans=t(topo["conn_trans"], torch.float64),
                      conn_gdz=t(topo["conn_gdz"], torch.float64),
                      row_offsets=t(topo["row_offsets"], torch.int64),
                      col_blocks=t(topo["col_blocks"], torch.int64),
                      volume=t(volume, torch.float64), porosity=t(porosity, torch.float64),
                      T=t(T, torch.float64))
        h = ctypes.c_void_p()
        d = self.d
        _abi.check(self.lib.nn_assembly_create(
            ctx.handle, self.ncell, int(self.has_well), _ptr(d["conn_offsets"]),
            _ptr(d["conn_neighbor"]), _ptr(d["conn_trans"]), _ptr(d["conn_gdz"]),
            _ptr(d["row_offsets"]), _ptr(d["col_blocks"]), self.nnzb, _ptr(d["volume"]),
            _ptr(d["porosity"]), _ptr(d["T"]), ctypes.byref(h)), "nn_assembly_create")
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self.lib.nn_assembly_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scaling(self, enabled=True, well_row_scale=None, **factors) -> None:
        """The runtime's equation/variable scaling for later evaluations
        (natural_petsc_runtime_scaling.inc:8-89): factors named as nn_scaling's
        fields (reference defaults otherwise); well_row_scale: per-cell factor of
        the well-control row (host or device array), None = 1/pressure_scale."""
        sc = _abi.Scaling.defaults()
        sc.enabled = 1 if enabled else 0
        for name, value in factors.items():
            setattr(sc, name, float(value))
        self._well_row_scale = (None if well_row_scale is None else
                                torch.as_tensor(np.ascontiguousarray(well_row_scale),
                                                dtype=torch.float64).to(self.d["T"].device))
        _abi.check(self.lib.nn_assembly_set_scaling(self.handle, ctypes.byref(sc),
                                                    _ptr(self._well_row_scale)),
                   "nn_assembly_set_scaling")

    def set_wells(self, wells, perforation_cells, perforation_well_index, surface_density):
        """NaturalPetscRuntime::setWells for the later evaluations (nn_assembly_set_wells):
        `wells` an nn_well ctypes array (workloads.make_wells), perforation arrays on the
        host; an empty list removes the wells."""
        pc = np.ascontiguousarray(perforation_cells, np.int64)
        wi = np.ascontiguousarray(perforation_well_index, np.float64)
        sd = np.ascontiguousarray(surface_density, np.float64)
        p = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
        _abi.check(self.lib.nn_assembly_set_wells(
            self.handle, len(wells), ctypes.cast(wells, ctypes.c_void_p) if len(wells) else None,
            len(pc), p(pc), p(wi), p(sd)), "nn_assembly_set_wells")

    def limit_newton_step(self, primary: torch.Tensor, step: torch.Tensor,
                          max_aqueous_unsupported: float | None = None, water: int = 0,
                          co2: int = 1) -> torch.Tensor:
        """NaturalPetscRuntime::limitNewtonStep on the device, in place on `step`
        (PETSc's correction Y, ncell x np): scaling recovery when the assembly's
        scaling is on, the phase-aware increment limiter, and with
        `max_aqueous_unsupported` the Water composition block kept inside the
        aqueous closures' domain (natural_petsc_runtime_state.inc:452-545)."""
        limit = -1.0 if max_aqueous_unsupported is None else float(max_aqueous_unsupported)
        _abi.check(self.lib.nn_newton_limit(self.handle, _ptr(primary), _ptr(step), limit,
                                            int(water), int(co2), _stream()),
                   "nn_newton_limit")
        return step

    def accumulation(self, primary: torch.Tensor, phase_state: torch.Tensor) -> torch.Tensor:
        acc = torch.empty(self.ncell, self.nc, dtype=torch.float64, device=primary.device)
        _abi.check(self.lib.nn_assembly_accumulation(self.handle, _ptr(primary), _ptr(phase_state),
                                                     _ptr(acc), _stream()),
                   "nn_assembly_accumulation")
        return acc

    def allocate(self, device="cuda"):
        return dict(residual=torch.empty(self.ncell, self.np_, dtype=torch.float64, device=device),
                    bsr=torch.empty(self.nnzb, self.np_, self.np_, dtype=torch.float64, device=device),
                    decisions=torch.empty(self.nconn, 3, dtype=torch.uint8, device=device),
                    status=torch.empty(self.ncell, dtype=torch.int32, device=device))

    def evaluate(self, dt, primary, phase_state, accum_prev, out=None, fugacity_scaling=1.0,
                 want_residual=True, want_bsr=True):
        out = out or self.allocate(primary.device)
        _abi.check(self.lib.nn_assembly_evaluate(
            self.handle, float(dt), float(fugacity_scaling), _ptr(primary), _ptr(phase_state),
            _ptr(accum_prev), _ptr(out["residual"] if want_residual else None),
            _ptr(out["bsr"] if want_bsr else None), _ptr(out["decisions"]), _ptr(out["status"]),
            _stream()), "nn_assembly_evaluate")
        return out


# ------------------------------------------------------------------ state builders
def assign_flash_result(flash: dict, P, has_well=False, floor=1e-30):
    """FullyCompositionalThreePhaseEquilibrium::assignFlashResult
    (state/three_phase_equilibrium.hpp:128-156) for every cell, on host arrays."""
    P = np.asarray(P, np.float64)
    n = len(P)
    nc = flash["x"].shape[2]
    np_ = 3 * nc + 1 + (1 if has_well else 0)
    primary = np.zeros((n, np_))
    primary[:, 0] = P
    for p in range(3):
        primary[:, 1 + p * (nc - 1):1 + (p + 1) * (nc - 1)] = flash["x"][:, p, :nc - 1]
    primary[:, 3 * nc - 2:3 * nc + 1] = flash["saturation"]
    if has_well:
        primary[:, 3 * nc + 1] = P
    ps = np.zeros((n, 3 * nc + 8))
    ps[:, 0] = flash["presence"].astype(np.float64)
    ps[:, 1:1 + nc] = flash["k_go"]
    ps[:, 1 + nc:1 + 2 * nc] = flash["k_wo"]
    beta, x = flash["beta"], flash["x"]
    z = 0.0 + beta[:, 0:1] * x[:, 0]
    z = z + beta[:, 1:2] * x[:, 1]
    z = z + beta[:, 2:3] * x[:, 2]
    z = np.maximum(z, floor)
    z = z / z.sum(axis=1, keepdims=True)
    ps[:, 1 + 2 * nc:1 + 3 * nc] = z
    ps[:, 1 + 3 * nc:4 + 3 * nc] = beta
    ps[:, 4 + 3 * nc:7 + 3 * nc] = flash["Z"]
    ps[:, 7 + 3 * nc] = 0.0
    return primary, ps


def perturb_newton_state(primary, phase_state, nc, seed, push_fraction=0.05):
    """Newton-like perturbation of an equilibrium state: dp ~ U(-1e5, 1e5) Pa,
    dx ~ U(-1e-3, 1e-3) on independent compositions and dS ~ U(-1e-3, 1e-3) on
    saturations of active phases; a fraction of cells gets one active saturation
    pushed to -1e-6 (forcing the phase-removal / stability path)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    prim = primary.copy()
    n = len(prim)
    presence = phase_state[:, 0].round().astype(np.int64)
    prim[:, 0] += rng.uniform(-1e5, 1e5, n)
    for p in range(3):
        on = (presence >> p) & 1 == 1
        sl = slice(1 + p * (nc - 1), 1 + (p + 1) * (nc - 1))
        prim[on, sl] += rng.uniform(-1e-3, 1e-3, (int(on.sum()), nc - 1))
        prim[on, 3 * nc - 2 + p] += rng.uniform(-1e-3, 1e-3, int(on.sum()))
    pushed = rng.random(n) < push_fraction
    for c in np.nonzero(pushed)[0]:
        act = [p for p in range(3) if (presence[c] >> p) & 1]
        if act:
            prim[c, 3 * nc - 2 + act[rng.integers(len(act))]] = -1e-6
    return prim
