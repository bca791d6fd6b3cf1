"""Central-difference check of a BSR Jacobian against its own residual.

Independent of how the Jacobian is assembled (self block, sign of the reciprocal
block, slot placement in the cached pattern): block (row i, column j)[r][q] must
equal dR_i[r]/dx_j[q], estimated by (R_i(x + h e_jq) - R_i(x - h e_jq)) / 2h with
the phase state held fixed.

Columns are perturbed several at a time: with a 7-point stencil, cells of equal
colour (i + 2j + 3k) mod 7 share no row of the Jacobian (two cells whose index
offset is a face neighbour or a sum of two face offsets have different colours),
so one residual pair per (colour, variable) gives every block of those columns.

Steps: h = 1e-7 |p| on pressures (~1-3 Pa at 10-30 MPa), h = 1e-6 on mole
fractions and saturations (chosen by a sweep over 1e-9..1e-5 on the reference
build: the error is smallest there for every fluid).  The truncation error is
h^2/6 |R'''|; roundoff is ~eps |R| / h.  A column whose perturbation changes an
upwind decision, a phase-state decision, or the status of any cell is left out
(the residual is not differentiable across those switches); entries are compared
by err = |J_fd - J| / max(s, 1e4 eps |R|_r / 2h): s = the largest |dR_i[r]/dx[q]|
over every block of matrix row i (one equation, one variable: one unit),
|R|_r = the largest |R| of equation r over the grid (the roundoff floor of the
difference quotient).  Callers bound the returned worst err (1e-3 in the tests;
truncation at these steps measured <= 9e-5 on PR, PR-CPA and SW grids).

Gravity faces: the reference's AD face flux divides the saturation-weighted
density sum by the VALUE of the mean saturation (face_flux.hpp:115-123,
`std::max(scalarValue(saturationAverage), 1.0e-8)`), so its Jacobian leaves out
-g dz * rho_face / (S_in + S_ex) in d(flux)/dS on every face with g dz != 0.  That
is part of the linearisation being reproduced, not a defect of either build: on
such faces the saturation columns of the blocks that carry the face term (the
cell's own block and the reciprocal block of a Z face) are not compared here
(`gravity_saturation_entries` counts them); the bit-level parity with the
reference build covers them.  On grids with nz = 1 every entry is compared.
"""
from __future__ import annotations

import numpy as np


def colours(nx: int, ny: int, nz: int) -> np.ndarray:
    i = np.arange(nx)[None, None, :]
    j = np.arange(ny)[None, :, None]
    k = np.arange(nz)[:, None, None]
    return ((i + 2 * j + 3 * k) % 7).reshape(-1)


def steps(primary: np.ndarray) -> np.ndarray:
    """Per-entry central-difference step: relative on pressures (column 0 and, with
    a well unknown, the last column when it holds a pressure), absolute elsewhere."""
    h = np.full(primary.shape, 1.0e-6)
    h[:, 0] = 1.0e-7 * np.abs(primary[:, 0])
    if primary.shape[1] % 3 == 2:           # 3nc + 2 columns: well (BHP) unknown last
        h[:, -1] = 1.0e-7 * np.maximum(np.abs(primary[:, -1]), 1.0e5)
    return h


def block_slot(topo: dict, row: int, col: int) -> int:
    ro, cb = topo["row_offsets"], topo["col_blocks"]
    lo, hi = int(ro[row]), int(ro[row + 1])
    k = lo + int(np.searchsorted(cb[lo:hi], col))
    assert k < hi and cb[k] == col
    return k


def identity_rows(flag: int, nc: int, npv: int) -> dict:
    """Rows of an inactive phase's block (cell_residual.hpp:257-319): the first
    nc - 1 rows hold x - value(x) for that phase's independent compositions, zero
    in value with derivative one, which a difference quotient cannot see; so does
    the well-control row of a cell without completion (cell_residual.hpp:414-422)
    when the primary vector carries a well unknown.  Maps equation -> the primary
    column whose identity it is."""
    oil, gas, water = bool(flag & 1), bool(flag & 2), bool(flag & 4)
    xcol = lambda phase: 1 + phase * (nc - 1)  # noqa: E731
    blocks = []                                # (first equation, inactive phase)
    if oil:
        if not gas:
            blocks.append((nc, 1))
        if not water:
            blocks.append((2 * nc, 2))
    else:
        blocks.append((nc, 0))
        if gas and not water:
            blocks.append((2 * nc, 2))
        elif not gas:
            blocks.append((2 * nc, 1))
    rows = {npv - 1: npv - 1} if npv == 3 * nc + 2 else {}
    for eq0, phase in blocks:
        for a in range(nc - 1):
            rows[eq0 + a] = xcol(phase) + a
    return rows


ROUNDOFF = 1.0e4 * np.finfo(np.float64).eps
TRACE_MARGIN = 1.0e-12   # above the residual's trace threshold (1e-14) and decode snap (1e-13)


def near_trace(primary: np.ndarray, flags: np.ndarray, nc: int, h: np.ndarray) -> np.ndarray:
    """(ncell, np) mask of composition columns whose +-h perturbation brings a
    composition of an active phase (including the dependent last one,
    1 - sum of the others) to the trace threshold: the residual switches its
    equation there (cell_residual.hpp:209-245), so no derivative exists."""
    mask = np.zeros(primary.shape, bool)
    for p in range(3):
        active = (flags.astype(np.int64) >> p) & 1 == 1
        cols = list(range(1 + p * (nc - 1), 1 + (p + 1) * (nc - 1)))
        x = primary[:, cols]
        last = 1.0 - x.sum(axis=1)
        for a, q in enumerate(cols):
            lo_q = x[:, a] - h[:, q]
            lo_last = last - h[:, q]
            lo_other = np.delete(x, a, axis=1).min(axis=1) if nc > 2 else np.full(len(x), np.inf)
            mask[:, q] = active & (np.minimum(np.minimum(lo_q, lo_last), lo_other) <= TRACE_MARGIN)
    return mask


def check(residual, primary: np.ndarray, bsr: np.ndarray, topo: dict, colour: np.ndarray,
          flags: np.ndarray, nc: int):
    """residual(primary) -> (R (ncell, np), decisions (nconn, 3), status (ncell,)).
    flags: per-cell phase-presence bits (bit0 oil, bit1 gas, bit2 water).

    Returns dict(worst relative error, compared entries, columns left out at a
    switch, gravity-face saturation entries not compared, details of the worst)."""
    ncell, npv = primary.shape
    sat_cols = set(range(1 + 3 * (nc - 1), 1 + 3 * (nc - 1) + 3))
    has_gravity = np.zeros(ncell, bool)
    gravity_face = set()
    for i in range(ncell):
        for k in range(int(topo["conn_offsets"][i]), int(topo["conn_offsets"][i + 1])):
            if topo["conn_gdz"][k] != 0.0:
                has_gravity[i] = True
                gravity_face.add((i, int(topo["conn_neighbor"][k])))
    R0, dec0, st0 = residual(primary)
    assert np.all(st0 == 0)
    # magnitude of each equation over the grid: the central difference's roundoff
    # is ~eps times the size of the terms that make up the residual, divided by 2h
    eq_scale = np.abs(R0).max(axis=0)
    h = steps(primary)
    trace = near_trace(primary, flags, nc, h)
    offs, nbr = topo["conn_offsets"], topo["conn_neighbor"]
    worst, compared, skipped, where = 0.0, 0, 0, None
    gravity_saturation_entries = 0
    for c in range(int(colour.max()) + 1):
        cells = np.nonzero(colour == c)[0]
        if len(cells) == 0:
            continue
        for q in range(npv):
            xp, xm = primary.copy(), primary.copy()
            xp[cells, q] += h[cells, q]
            xm[cells, q] -= h[cells, q]
            Rp, decp, stp = residual(xp)
            Rm, decm, stm = residual(xm)
            for j in cells:
                rows = [int(j)] + [int(v) for v in nbr[offs[j]:offs[j + 1]]]
                conns = []
                for i in rows:
                    conns.extend(range(int(offs[i]), int(offs[i + 1])))
                if (trace[j, q] or np.any(stp[rows] != 0) or np.any(stm[rows] != 0) or
                        np.any(decp[conns] != dec0[conns]) or np.any(decm[conns] != dec0[conns])):
                    skipped += 1
                    continue
                for i in rows:
                    if q in sat_cols and ((i == j and has_gravity[i]) or (i, int(j)) in gravity_face):
                        gravity_saturation_entries += npv
                        continue
                    fd = (Rp[i] - Rm[i]) / (2.0 * h[j, q])
                    blk = bsr[block_slot(topo, i, int(j))]
                    # identity rows: exactly e_col in the own block, zero elsewhere
                    for eq, col in identity_rows(int(flags[i]), nc, npv).items():
                        fd[eq] = 1.0 if (i == j and q == col) else 0.0
                    # same equation, same variable, every block of the matrix row:
                    # one unit, so a term missing from one block cannot hide
                    # behind a larger entry of another variable
                    lo, hi = int(topo["row_offsets"][i]), int(topo["row_offsets"][i + 1])
                    scale = np.abs(bsr[lo:hi, :, q]).max(axis=0)
                    scale = np.maximum(scale, ROUNDOFF * eq_scale / (2.0 * h[j, q]))
                    scale = np.where(scale > 0.0, scale, 1.0)
                    err = np.abs(fd - blk[:, q]) / scale
                    compared += len(err)
                    r = int(np.argmax(err))
                    if err[r] > worst:
                        worst = float(err[r])
                        where = dict(row_cell=i, col_cell=int(j), var=q, eq=r, fd=float(fd[r]),
                                     jac=float(blk[r, q]), scale=float(scale[r]))
    return dict(worst=worst, compared=compared, skipped=skipped,
                gravity_saturation_entries=gravity_saturation_entries, where=where)
