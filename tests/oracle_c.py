"""ctypes binding of the oracle's plain-C restatement (oracle/liboracle.so)."""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
SO = ROOT / "oracle" / "liboracle.so"
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not SO.exists():
            subprocess.run(["make", "-C", str(ROOT / "oracle"), "-s", "liboracle.so"], check=True)
        L = ctypes.CDLL(str(SO))
        P = ctypes.c_void_p
        L.oracle_phase.argtypes = [P, ctypes.c_double, ctypes.c_double, P, ctypes.c_int, P, P]
        L.oracle_phase.restype = ctypes.c_int
        L.oracle_eos_batch.argtypes = [P, ctypes.c_int64, P, P, P, P, P, P, P]
        L.oracle_wilson_k.argtypes = [P, ctypes.c_double, ctypes.c_double, P]
        L.oracle_rr2.argtypes = [ctypes.c_int, P, P, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int, P]
        L.oracle_rr2.restype = ctypes.c_int
        L.oracle_plant.argtypes = [ctypes.c_int]
        L.oracle_plant.restype = None
        L.oracle_lbc_viscosity.argtypes = [P, P, ctypes.c_double, ctypes.c_double]
        L.oracle_lbc_viscosity.restype = ctypes.c_double
        D = ctypes.c_double
        for name, args in (("oracle_if97_density", [D, D]),
                           ("oracle_if97_region3_pressure", [D, D]),
                           ("oracle_if97_volume", [ctypes.c_int, D, D]),
                           ("oracle_iapws2008_viscosity", [D, D]),
                           ("oracle_mcbride_wright_viscosity", [D, D, D]),
                           ("oracle_garcia_co2_volume", [D])):
            getattr(L, name).argtypes = args
            getattr(L, name).restype = D
        L.oracle_newton_limit.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, P, P, P,
                                          D, ctypes.c_int, ctypes.c_int]
        L.oracle_newton_limit.restype = None
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def eos_batch(fluid, P, T, x, role):
    n = len(P)
    P = np.ascontiguousarray(P, np.float64)
    T = np.ascontiguousarray(T, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    role = np.ascontiguousarray(role, np.int32)
    Z = np.zeros(n)
    lnphi = np.zeros((n, fluid.nc))
    st = np.zeros(n, np.int32)
    lib().oracle_eos_batch(ctypes.byref(fluid), n, _p(P), _p(T), _p(x), _p(role), _p(Z),
                           _p(lnphi), _p(st))
    return Z, lnphi, st


def rr2(z, K, beta0=0.5, tol=1e-12, maxit=80):
    z = np.ascontiguousarray(z, np.float64)
    K = np.ascontiguousarray(K, np.float64)
    out = np.zeros(1)
    ok = lib().oracle_rr2(len(z), _p(z), _p(K), beta0, tol, maxit, _p(out))
    return bool(ok), float(out[0])


def lbc(fluid, x, rho, T):
    x = np.ascontiguousarray(x, np.float64)
    return lib().oracle_lbc_viscosity(ctypes.byref(fluid), _p(x), float(rho), float(T))


def newton_limit(nc, has_well, primary, step, var_scale=None, max_aqueous=None, water=0, co2=1):
    """The oracle's restatement of limitNewtonStep; returns the limited step."""
    primary = np.ascontiguousarray(primary, np.float64)
    out = np.array(step, np.float64, copy=True, order="C")
    vs = None if var_scale is None else np.ascontiguousarray(var_scale, np.float64)
    lib().oracle_newton_limit(nc, int(has_well), len(primary), _p(primary), _p(out),
                              ctypes.c_void_p(0) if vs is None else _p(vs),
                              -1.0 if max_aqueous is None else float(max_aqueous), water, co2)
    return out
