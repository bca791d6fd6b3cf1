"""Seeded synthetic workloads (numpy PCG64) used by tests and bench.py.

batched_flash_states: the batched-flash workload — P ~ U(1, 40) MPa,
T ~ U(300, 600) K, z ~ Dirichlet(1, ..., 1).  A fraction of the states can be
drawn near the mixture's vapour-liquid critical region (see
``near_critical_fraction``): with a `locus` table (critical_points) they take a
composition of the table and (P, T) within 2 % of its estimated critical point;
without one, z from the same Dirichlet and (P, T) within 2 % of Kay's-rule
pseudo-critical point (T_pc = sum z_i Tc_i, P_pc = sum z_i Pc_i).

critical_points: the reference's envelope-coalescence estimate of a mixture's
critical point (tools/include/tools/phase_diagram.hpp:633-704, 860-908), with
any batched flash.
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


def batched_flash_states(n: int, fluid, seed: int = 0, near_critical_fraction: float = 0.0,
                         locus: dict | None = None):
    rng = np.random.Generator(np.random.PCG64(seed))
    nc = fluid.nc
    P = rng.uniform(1.0e6, 40.0e6, n)
    T = rng.uniform(300.0, 600.0, n)
    z = rng.dirichlet(np.ones(nc), n)
    m = int(round(near_critical_fraction * n))
    if m > 0:
        idx = rng.choice(n, m, replace=False)
        if locus is not None:
            ok = np.nonzero(locus["valid"])[0]
            if len(ok) == 0:
                raise ValueError("the critical-point table has no valid point")
            pick = ok[rng.integers(0, len(ok), m)]
            z[idx] = locus["z"][pick]
            T[idx] = locus["Tc"][pick] * rng.uniform(0.98, 1.02, m)
            P[idx] = np.clip(locus["Pc"][pick] * rng.uniform(0.98, 1.02, m), 1.0e5, None)
        else:
            tc = np.array([fluid.tc[i] for i in range(nc)])
            pc = np.array([fluid.pc[i] for i in range(nc)])
            tpc = z[idx] @ tc
            ppc = z[idx] @ pc
            T[idx] = tpc * rng.uniform(0.98, 1.02, m)
            P[idx] = np.clip(ppc * rng.uniform(0.98, 1.02, m), 1.0e5, None)
    return P, T, z


def critical_points(flash, z, temperatures, pressures, refine_iterations=40,
                    relative_tolerance=1.0e-8, score_tolerance=0.20):
    """Estimate of each composition's vapour-liquid critical point from batched
    flashes.

    flash(P, T, z) -> dict with "presence" (n), "status" (n) and "x" (n, 3, nc)
    as host arrays, for a batch.  The scan and the boundary refinement restate
    the reference's phase-diagram tool (tools/include/tools/phase_diagram.hpp):
    for each composition and each temperature of the `temperatures` axis the
    ascending `pressures` axis is flashed (pressureTemperature), and the last
    upper boundary of the two-phase region along it, two-phase (code 3) below
    and liquid (1, a bubble point) or vapour (2, an upper dew point above the
    critical temperature) above, is refined by bisection on the phase code until
    the interval is within `relative_tolerance` or `refine_iterations` pass
    (refineTransitionPressure, :949-993).  Phase codes are the oil/gas presence
    bits (the water bit is ignored, so water-bearing mixtures have an oil/gas
    envelope).

    The critical point is where the coexisting phases become identical, so each
    boundary point is scored by the L1 distance of the oil and gas compositions
    (oilGasCompositionL1, :453-466) flashed at the interval's two-phase end, and
    the lowest score over the temperatures gives (Tc, Pc), valid when it is at
    most `score_tolerance` (the reference's quality gate, :145-147).  (The
    reference scores the midpoint between a dew and a bubble point plus the
    envelope's relative width, :672-704; for mixtures whose lower dew branch is
    still far below the bubble point at the critical temperature, such as CO2
    with a heavy alkane, that score does not approach zero, and the boundary's
    composition distance does.)  All flashes of one stage run as one batch.
    Returns dict(z, Tc, Pc, score, valid)."""
    z = np.atleast_2d(np.asarray(z, np.float64))
    Ts = np.asarray(temperatures, np.float64)
    Ps = np.asarray(pressures, np.float64)
    nz, nt, npr = len(z), len(Ts), len(Ps)

    def codes(res):
        code = np.asarray(res["presence"]).astype(np.int64) & 3
        code[np.asarray(res["status"]) != 0] = -1
        return code

    iz, it, ip = np.meshgrid(np.arange(nz), np.arange(nt), np.arange(npr), indexing="ij")
    iz, it, ip = iz.ravel(), it.ravel(), ip.ravel()
    scan = codes(flash(Ps[ip], Ts[it], z[iz])).reshape(nz, nt, npr)
    # the highest two-phase -> single-phase step along each pressure line
    up = (scan[:, :, :-1] == 3) & ((scan[:, :, 1:] == 1) | (scan[:, :, 1:] == 2))
    last = np.where(up.any(axis=2), npr - 2 - np.argmax(up[:, :, ::-1], axis=2), -1)
    sel = np.argwhere(last >= 0)
    out = dict(z=z, Tc=np.full(nz, np.nan), Pc=np.full(nz, np.nan), score=np.full(nz, np.inf),
               valid=np.zeros(nz, bool))
    if len(sel) == 0:
        return out
    k0 = last[sel[:, 0], sel[:, 1]]
    single_code = scan[sel[:, 0], sel[:, 1], k0 + 1]
    p_two, p_single = Ps[k0].copy(), Ps[k0 + 1].copy()
    active = np.ones(len(sel), bool)
    for _ in range(refine_iterations):
        width = np.abs(p_single - p_two)
        scale = np.maximum(np.maximum(1.0, np.abs(p_single)), np.abs(p_two))
        active &= ~(width <= relative_tolerance * scale)
        w = np.nonzero(active)[0]
        if len(w) == 0:
            break
        mid = 0.5 * (p_single[w] + p_two[w])
        c = codes(flash(mid, Ts[sel[w, 1]], z[sel[w, 0]]))
        two = c == 3
        single = c == single_code[w]
        p_two[w] = np.where(two, mid, p_two[w])
        p_single[w] = np.where(single, mid, p_single[w])
        active[w[~two & ~single]] = False   # a failed or other code stops the refinement
    res = flash(p_two, Ts[sel[:, 1]], z[sel[:, 0]])
    x = np.asarray(res["x"])
    l1 = np.abs(x[:, 0, :] - x[:, 1, :]).sum(axis=1)
    score = np.where(codes(res) == 3, l1, np.inf)
    for r in range(len(sel)):   # lowest score per composition, first temperature on ties
        k = sel[r, 0]
        if score[r] < out["score"][k]:
            out["score"][k] = score[r]
            out["Tc"][k] = Ts[sel[r, 1]]
            out["Pc"][k] = p_two[r]
    out["valid"] = out["score"] <= score_tolerance
    return out


# Surface densities (oil, gas, water; kg/m3) of case/h2o_co2_nc10_2d_benchmark/case_config.hpp:65.
SURFACE_DENSITY = (730.0, 1.8, 985.4040020947351)


def make_wells(specs, nc):
    """An nn_well ctypes array from dicts with keys type ("injector" / "producer"),
    bhp_cell, target (Pa), perforations (range of the perforation arrays) and, for
    injectors, injection_phase_fraction (3) and injection_component_mass_fraction (nc)."""
    from ._abi import NN_WELL_CONTROL_BHP, NN_WELL_INJECTOR, NN_WELL_PRODUCER, Well
    arr = (Well * len(specs))()
    for w, sp in zip(arr, specs):
        w.type = NN_WELL_INJECTOR if sp["type"] == "injector" else NN_WELL_PRODUCER
        w.control = NN_WELL_CONTROL_BHP
        w.bhp_cell = int(sp["bhp_cell"])
        w.target = float(sp["target"])
        for p, v in enumerate(sp.get("injection_phase_fraction", (0.0, 0.0, 0.0))):
            w.injection_phase_fraction[p] = float(v)
        for i, v in enumerate(sp.get("injection_component_mass_fraction", (0.0,) * nc)):
            w.injection_component_mass_fraction[i] = float(v)
        w.perforation_begin, w.perforation_end = int(sp["perforations"][0]), int(sp["perforations"][1])
    return arr


def vertical_bhp_wells(nx, ny, nz, dx, dy, dz, perm, P, co2=1, nc=3, wells=None):
    """Bottom-hole-pressure wells completed through every layer of a column: by
    default two CO2 injectors (gas phase, pure CO2, BHP 2 MPa above the column's top
    pressure) at (nx/4, ny/4) and (3nx/4, 3ny/4) and two producers (BHP 0.5 MPa
    below it) at (3nx/4, ny/4) and (nx/4, 3ny/4), the offsets of the reference's
    benchmark wells (case/h2o_co2_nc10_2d_benchmark/benchmark_common.hpp:53-54).
    Well index: Peaceman's isotropic form 2 pi k dz / ln(r_eq / r_w),
    r_eq = 0.14 sqrt(dx^2 + dy^2), r_w = 0.1 m.  The bhp cell is the top
    perforation.  Returns (specs, perforation_cells, perforation_well_index)."""
    perm = np.asarray(perm, np.float64)
    kcell = perm if perm.ndim == 1 else perm[:, 0]
    if wells is None:
        wells = [("injector", nx // 4, ny // 4), ("producer", (3 * nx) // 4, ny // 4),
                 ("injector", (3 * nx) // 4, (3 * ny) // 4), ("producer", nx // 4, (3 * ny) // 4)]
    req = 0.14 * np.hypot(dx, dy)
    specs, cells, wis = [], [], []
    for kind, i, j in wells:
        col = [i + nx * (j + ny * k) for k in range(nz)]
        top = col[0]
        begin = len(cells)
        cells += col
        wis += [2.0 * np.pi * kcell[c] * dz / np.log(req / 0.1) for c in col]
        sp = dict(type=kind, bhp_cell=top, perforations=(begin, len(cells)))
        if kind == "injector":
            sp["target"] = float(P[top]) + 2.0e6
            sp["injection_phase_fraction"] = (0.0, 1.0, 0.0)
            mf = [0.0] * nc
            mf[co2] = 1.0
            sp["injection_component_mass_fraction"] = mf
        else:
            sp["target"] = float(P[top]) - 0.5e6
        specs.append(sp)
    return specs, np.array(cells, np.int64), np.array(wis)


def pattern_add_wells(topo, wells, perforation_cells):
    """The topology with the wells' couplings added to its cached BSR pattern
    (nn_pattern_add_wells)."""
    from ._abi import lib
    L = lib()
    n = len(topo["row_offsets"]) - 1
    pc = np.ascontiguousarray(perforation_cells, np.int64)
    p = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    ro = np.ascontiguousarray(topo["row_offsets"], np.int64)
    cb = np.ascontiguousarray(topo["col_blocks"], np.int64)
    nnzb = ctypes.c_int64()
    w = ctypes.cast(wells, ctypes.c_void_p)
    if L.nn_pattern_add_wells(n, p(ro), p(cb), len(wells), w, p(pc), None, None,
                              ctypes.byref(nnzb)) != 0:
        raise RuntimeError("nn_pattern_add_wells failed")
    out_ro = np.zeros(n + 1, np.int64)
    out_cb = np.zeros(nnzb.value, np.int64)
    if L.nn_pattern_add_wells(n, p(ro), p(cb), len(wells), w, p(pc), p(out_ro), p(out_cb),
                              ctypes.byref(nnzb)) != 0:
        raise RuntimeError("nn_pattern_add_wells failed")
    t = dict(topo)
    t["row_offsets"], t["col_blocks"] = out_ro, out_cb
    return t
