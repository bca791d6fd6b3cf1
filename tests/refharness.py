"""ctypes binding of the oracle's reference build (oracle/_ref/libnn_ref.so).

Test infrastructure only: the shared object is the reference implementation
compiled from its own sources by oracle/Makefile.  It is absent on machines
without the reference checkout unless it was built beforehand; callers skip
(or fall back to committed golden fixtures) when :func:`available` is False.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
REF_SO = ROOT / "oracle" / "_ref" / "libnn_ref.so"

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_lib = None


def available() -> bool:
    return REF_SO.exists()


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(str(REF_SO))
        L.ref_create.restype = _P
        L.ref_create.argtypes = [_P, _P]
        L.ref_destroy.argtypes = [_P]
        L.ref_eos_phase.argtypes = [_P, _I64] + [_P] * 8
        L.ref_flash.argtypes = [_P, _I64] + [_P] * 16 + [ctypes.c_int]
        L.ref_stability.argtypes = [_P, _I64] + [_P] * 11
        L.ref_phase_update.argtypes = [_P, _I64, ctypes.c_int] + [_P] * 5 + [ctypes.c_int]
        L.ref_assembly.argtypes = ([_P, _I64, ctypes.c_int] + [_P] * 9 +
                                   [ctypes.c_double, ctypes.c_double] + [_P] * 7 + [ctypes.c_int])
        L.ref_assembly_scaled.argtypes = ([_P, _I64, ctypes.c_int] + [_P] * 9 +
                                          [ctypes.c_double, ctypes.c_double] + [_P] * 9 +
                                          [ctypes.c_int])
        L.ref_water_closures.argtypes = [_I64] + [_P] * 7
        L.ref_newton_limit.argtypes = [_P, _I64, ctypes.c_int, _P, _P, _P, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int]
        L.ref_assembly_wells.argtypes = [_P, _I64, ctypes.c_int, _P, _P, _P, _P, _P, _P, _P, _P,
                                         _P, ctypes.c_double, ctypes.c_double, _P, _P,
                                         ctypes.c_int32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                         ctypes.c_int]
        L.ref_last_error.restype = ctypes.c_char_p
        L.ref_csv_samples.argtypes = [_I64, _P, ctypes.c_char_p, _I64]
        L.ref_csv_samples.restype = _I64
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


def csv_samples(values) -> str:
    """The reference's sample CSV text (writeIndexedCsvHeader + writeCsvValues) for
    one row of `values`."""
    values = np.ascontiguousarray(values, np.float64)
    n = lib().ref_csv_samples(len(values), _p(values), None, 0)
    buf = ctypes.create_string_buffer(int(n) + 1)
    lib().ref_csv_samples(len(values), _p(values), buf, n + 1)
    return buf.value.decode()


def water_closures(P, T, xco2):
    """The reference's IF97 density, Garcia CO2 volume, IAPWS-2008 and McBride-Wright
    viscosities (NaN where it throws)."""
    P, T, xco2 = (np.ascontiguousarray(a, np.float64) for a in (P, T, xco2))
    out = [np.zeros(len(P)) for _ in range(4)]
    lib().ref_water_closures(len(P), _p(P), _p(T), _p(xco2), *(_p(o) for o in out))
    return out


class Reference:
    def __init__(self, fluid, options):
        self.L = lib()
        self.fluid = fluid
        self.options = options
        self.nc = fluid.nc
        self.h = self.L.ref_create(ctypes.byref(fluid), ctypes.byref(options))

    def __del__(self):
        try:
            self.L.ref_destroy(self.h)
        except Exception:
            pass

    def eos_phase(self, P, T, x, role):
        n = len(P)
        P, T = np.ascontiguousarray(P, np.float64), np.ascontiguousarray(T, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        role = np.ascontiguousarray(role, np.int32)
        Z = np.zeros(n)
        lnphi = np.zeros((n, self.nc))
        fug = np.zeros((n, self.nc))
        status = np.zeros(n, np.int32)
        self.L.ref_eos_phase(self.h, n, _p(P), _p(T), _p(x), _p(role), _p(Z), _p(lnphi),
                             _p(fug), _p(status))
        return Z, lnphi, fug, status

    def flash(self, P, T, z, allowed=None, seed=None, threads=1):
        n = len(P)
        nc = self.nc
        P, T = np.ascontiguousarray(P, np.float64), np.ascontiguousarray(T, np.float64)
        z = np.ascontiguousarray(z, np.float64)
        if allowed is not None:
            allowed = np.ascontiguousarray(allowed, np.uint8)
        if seed is not None:
            seed = np.ascontiguousarray(seed, np.float64)
        out = dict(presence=np.zeros(n, np.uint8), converged=np.zeros(n, np.uint8),
                   status=np.zeros(n, np.int32), iterations=np.zeros(n, np.int32),
                   beta=np.zeros((n, 3)), saturation=np.zeros((n, 3)), x=np.zeros((n, 3, nc)),
                   Z=np.zeros((n, 3)), rho_m=np.zeros((n, 3)), k_go=np.zeros((n, nc)),
                   k_wo=np.zeros((n, nc)))
        self.L.ref_flash(self.h, n, _p(P), _p(T), _p(z), _p(allowed), _p(seed),
                         _p(out["presence"]), _p(out["converged"]), _p(out["status"]),
                         _p(out["iterations"]), _p(out["beta"]), _p(out["saturation"]),
                         _p(out["x"]), _p(out["Z"]), _p(out["rho_m"]), _p(out["k_go"]),
                         _p(out["k_wo"]), int(threads))
        return out

    def stability(self, P, T, z, active, x):
        n = len(P)
        nc = self.nc
        P, T = np.ascontiguousarray(P, np.float64), np.ascontiguousarray(T, np.float64)
        z = np.ascontiguousarray(z, np.float64)
        active = np.ascontiguousarray(active, np.uint8)
        x = np.ascontiguousarray(x, np.float64)
        out = dict(stable=np.zeros(n, np.uint8), valid=np.zeros(n, np.uint8),
                   unstable_mask=np.zeros(n, np.uint8), trial_sum=np.zeros((n, 3)),
                   incipient=np.zeros((n, 3, nc)), status=np.zeros(n, np.int32))
        self.L.ref_stability(self.h, n, _p(P), _p(T), _p(z), _p(active), _p(x),
                             _p(out["stable"]), _p(out["valid"]), _p(out["unstable_mask"]),
                             _p(out["trial_sum"]), _p(out["incipient"]), _p(out["status"]))
        return out

    def newton_limit(self, primary, step, has_well=False, scaling=None, max_aqueous=None,
                     water=0, co2=1):
        """limitNewtonStep (restated driver around the reference's limiters) on host
        arrays; returns the limited step.  scaling: the six factors or None."""
        primary = np.ascontiguousarray(primary, np.float64)
        out = np.array(step, np.float64, copy=True, order="C")
        sc = None if scaling is None else np.ascontiguousarray(scaling, np.float64)
        rc = self.L.ref_newton_limit(self.h, len(primary), int(has_well), _p(primary), _p(out),
                                     _p(sc), -1.0 if max_aqueous is None else float(max_aqueous),
                                     int(water), int(co2))
        if rc != 0:
            raise RuntimeError(f"ref_newton_limit failed: {rc}")
        return out

    def phase_update(self, T, primary, phase_state, has_well=False, threads=1):
        n = len(T)
        T = np.ascontiguousarray(T, np.float64)
        primary = np.array(primary, np.float64, copy=True, order="C")
        phase_state = np.array(phase_state, np.float64, copy=True, order="C")
        status = np.zeros(n, np.int32)
        flags = np.zeros(n, np.uint8)
        self.L.ref_phase_update(self.h, n, int(has_well), _p(T), _p(primary), _p(phase_state),
                                _p(status), _p(flags), int(threads))
        return primary, phase_state, status, flags

    def assembly_wells(self, topo, volume, porosity, T, primary, phase_state, accum_prev, dt,
                       wells, perforation_cells, perforation_well_index, surface_density,
                       fug_scale=1.0, threads=1, scaling=None, well_row_scale=None):
        """The restated runtime loops with wells (formFunction's perforation sources and
        control rows, insertWellJacobian_), around the reference's own
        computePerforationWellSourceInto.  wells: an nn_well ctypes array; the
        pattern in `topo` must hold the wells' couplings.  Returns (residual, bsr)."""
        n = len(T)
        np_ = primary.shape[1]
        arr = {k: np.ascontiguousarray(v) for k, v in topo.items()}
        vol = np.ascontiguousarray(volume, np.float64)
        por = np.ascontiguousarray(porosity, np.float64)
        T = np.ascontiguousarray(T, np.float64)
        primary = np.ascontiguousarray(primary, np.float64)
        phase_state = np.ascontiguousarray(phase_state, np.float64)
        accum_prev = np.ascontiguousarray(accum_prev, np.float64)
        pc = np.ascontiguousarray(perforation_cells, np.int64)
        wi = np.ascontiguousarray(perforation_well_index, np.float64)
        sd = np.ascontiguousarray(surface_density, np.float64)
        nnzb = int(arr["row_offsets"][-1])
        res = np.zeros((n, np_))
        bsr = np.zeros((nnzb, np_, np_))
        dec = np.zeros((int(arr["conn_offsets"][-1]), 3), np.uint8)
        sc = None if scaling is None else np.ascontiguousarray(scaling, np.float64)
        wr = None if well_row_scale is None else np.ascontiguousarray(well_row_scale, np.float64)
        rc = self.L.ref_assembly_wells(
            self.h, n, 1, _p(arr["conn_offsets"]), _p(arr["conn_neighbor"]),
            _p(arr["conn_trans"]), _p(arr["conn_gdz"]), _p(arr["row_offsets"]),
            _p(arr["col_blocks"]), _p(vol), _p(por), _p(T), float(dt), float(fug_scale), _p(sc),
            _p(wr), len(wells), ctypes.cast(wells, ctypes.c_void_p), _p(pc), _p(wi), _p(sd),
            _p(primary), _p(phase_state), _p(accum_prev), _p(res), _p(bsr), _p(dec),
            int(threads))
        if rc != 0:
            raise RuntimeError(f"ref_assembly_wells failed: {rc}: {self.L.ref_last_error()}")
        return res, bsr

    def assembly(self, topo, volume, porosity, T, primary, phase_state, accum_prev, dt,
                 fug_scale=1.0, has_well=False, want_residual=True, want_bsr=True, threads=1,
                 scaling=None, well_row_scale=None):
        """Restated runtime loops around the reference's per-cell kernels.  scaling:
        None (disabled) or the six factors (pressure, composition, saturation, mass
        residual, fugacity residual, closure residual); well_row_scale: per-cell
        factor of the well-control row or None."""
        n = len(T)
        np_ = primary.shape[1]
        arr = {k: np.ascontiguousarray(v) for k, v in topo.items()}
        arr["volume"] = np.ascontiguousarray(volume, np.float64)
        arr["porosity"] = np.ascontiguousarray(porosity, np.float64)
        T = np.ascontiguousarray(T, np.float64)
        primary = np.ascontiguousarray(primary, np.float64)
        phase_state = np.ascontiguousarray(phase_state, np.float64)
        accum_prev = np.ascontiguousarray(accum_prev, np.float64) if accum_prev is not None else None
        nnzb = int(arr["row_offsets"][-1])
        res = np.zeros((n, np_)) if want_residual else None
        bsr = np.zeros((nnzb, np_, np_)) if want_bsr else None
        nconn = int(arr["conn_offsets"][-1])
        dec = np.zeros((nconn, 3), np.uint8)
        acc = np.zeros((n, self.nc))
        sc = None if scaling is None else np.ascontiguousarray(scaling, np.float64)
        assert sc is None or sc.shape == (6,)
        wr = None if well_row_scale is None else np.ascontiguousarray(well_row_scale, np.float64)
        rc = self.L.ref_assembly_scaled(self.h, n, int(has_well), _p(arr["conn_offsets"]),
                                        _p(arr["conn_neighbor"]), _p(arr["conn_trans"]),
                                        _p(arr["conn_gdz"]), _p(arr["row_offsets"]),
                                        _p(arr["col_blocks"]), _p(arr["volume"]),
                                        _p(arr["porosity"]), _p(T), float(dt), float(fug_scale),
                                        _p(sc), _p(wr), _p(primary), _p(phase_state),
                                        _p(accum_prev), _p(res), _p(bsr), _p(dec), _p(acc),
                                        int(threads))
        if rc != 0:
            raise RuntimeError(f"ref_assembly failed: {rc}")
        return res, bsr, dec, acc
