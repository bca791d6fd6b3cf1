"""Seeded synthetic workloads (numpy PCG64) used by tests and bench.py.

batched_flash_states: the batched-flash workload — P ~ U(1, 40) MPa,
T ~ U(300, 600) K, z ~ Dirichlet(1, ..., 1).  A fraction of the states can be
drawn near the mixture's vapour-liquid critical region (see
``near_critical_fraction``): those states take z from the same Dirichlet and
put (P, T) within 2% of a pseudo-critical point estimated with Kay's rule
(T_pc = sum z_i Tc_i, P_pc = sum z_i Pc_i), the classical cheap estimate that
the reference's tools/include/tools/phase_diagram.hpp refines with a full
critical-locus solve.
"""
from __future__ import annotations

import ctypes

import numpy as np

MILLIDARCY = 9.869233e-16   # case/h2o_co2_nc10_2d_benchmark/benchmark_common.hpp:18
DAY = 86400.0


def structured_topology(nx: int, ny: int, nz: int, dx: float, dy: float, dz: float, perm):
    """Connections and cached BSR pattern of a structured grid (host setup through
    nn_structured_topology).  perm: (ncell,) isotropic or (ncell, 3)."""
    from ._abi import lib
    L = lib()
    ncell = nx * ny * nz
    perm = np.asarray(perm, np.float64)
    if perm.ndim == 1:
        perm = np.repeat(perm[:, None], 3, axis=1)
    perm = np.ascontiguousarray(perm)
    dxa = np.full(nx, float(dx))
    dya = np.full(ny, float(dy))
    dza = np.full(nz, float(dz))
    nconn = ctypes.c_int64()
    nnzb = ctypes.c_int64()
    p = lambda a: ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)  # noqa: E731
    rc = L.nn_structured_topology(nx, ny, nz, p(dxa), p(dya), p(dza), p(perm), None, None, None,
                                  None, None, None, ctypes.byref(nconn), ctypes.byref(nnzb))
    if rc != 0:
        raise RuntimeError("nn_structured_topology failed")
    t = dict(conn_offsets=np.zeros(ncell + 1, np.int64),
             conn_neighbor=np.zeros(nconn.value, np.int64),
             conn_trans=np.zeros(nconn.value), conn_gdz=np.zeros(nconn.value),
             row_offsets=np.zeros(ncell + 1, np.int64), col_blocks=np.zeros(nnzb.value, np.int64))
    rc = L.nn_structured_topology(nx, ny, nz, p(dxa), p(dya), p(dza), p(perm),
                                  p(t["conn_offsets"]), p(t["conn_neighbor"]), p(t["conn_trans"]),
                                  p(t["conn_gdz"]), p(t["row_offsets"]), p(t["col_blocks"]),
                                  ctypes.byref(nconn), ctypes.byref(nnzb))
    if rc != 0:
        raise RuntimeError("nn_structured_topology failed")
    return t


def grid_workload(nx: int, ny: int, nz: int, dx: float, dy: float, dz: float, fluid,
                  seed: int = 0, hydrostatic: bool = False):
    """Seeded cell fields of the assembly workload: P ~ U(8, 30) MPa,
    T ~ U(320, 450) K, z ~ Dirichlet(4, 1, 1) (water first), porosity ~ U(0.1, 0.3),
    isotropic log-normal permeability (median 100 mD, sigma_ln = 1).  With
    `hydrostatic`, P gains rho_w g depth (1000 kg/m3) down the k index."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ncell = nx * ny * nz
    nc = fluid.nc
    P = rng.uniform(8.0e6, 30.0e6, ncell)
    if hydrostatic:
        k = np.arange(ncell) // (nx * ny)
        P = P + 1000.0 * 9.80665 * (k + 0.5) * dz
    T = rng.uniform(320.0, 450.0, ncell)
    alpha = np.ones(nc)
    if fluid.water_component >= 0:
        alpha[fluid.water_component] = 4.0
    z = rng.dirichlet(alpha, ncell)
    porosity = rng.uniform(0.1, 0.3, ncell)
    perm = 100.0 * MILLIDARCY * np.exp(rng.normal(0.0, 1.0, ncell))
    volume = np.full(ncell, float(dx) * dy * dz)
    topo = structured_topology(nx, ny, nz, dx, dy, dz, perm)
    return dict(P=P, T=T, z=z, porosity=porosity, perm=perm, volume=volume, topo=topo,
                ncell=ncell, seed=seed)


def batched_flash_states(n: int, fluid, seed: int = 0, near_critical_fraction: float = 0.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    nc = fluid.nc
    P = rng.uniform(1.0e6, 40.0e6, n)
    T = rng.uniform(300.0, 600.0, n)
    z = rng.dirichlet(np.ones(nc), n)
    m = int(round(near_critical_fraction * n))
    if m > 0:
        idx = rng.choice(n, m, replace=False)
        tc = np.array([fluid.tc[i] for i in range(nc)])
        pc = np.array([fluid.pc[i] for i in range(nc)])
        tpc = z[idx] @ tc
        ppc = z[idx] @ pc
        T[idx] = tpc * rng.uniform(0.98, 1.02, m)
        P[idx] = np.clip(ppc * rng.uniform(0.98, 1.02, m), 1.0e5, None)
    return P, T, z
