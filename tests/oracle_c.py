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
        L.oracle_lbc_viscosity.argtypes = [P, P, ctypes.c_double, ctypes.c_double]
        L.oracle_lbc_viscosity.restype = ctypes.c_double
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
