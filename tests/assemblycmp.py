"""Comparison rules for residuals and BSR Jacobians.

A value a (device) matches b (oracle) when |a - b| <= 1e-10 * max(|a|, |b|, s),
where s is the row scale: for residual entry (cell, r) the max |b| of equation r
over all cells; for Jacobian entry (block, r, col) the max |b| of that block row.
Exact zeros in the oracle must be exact zeros on the device (sparsity parity).
"""
from __future__ import annotations

import numpy as np

REL_TOL = 1.0e-10


def residual_error(mine, ref):
    s = np.max(np.abs(ref), axis=0, keepdims=True)
    scale = np.maximum(np.maximum(np.abs(mine), np.abs(ref)), s)
    scale = np.where(scale > 0, scale, 1.0)
    return np.abs(mine - ref) / scale


def bsr_error(mine, ref):
    s = np.max(np.abs(ref), axis=2, keepdims=True)
    scale = np.maximum(np.maximum(np.abs(mine), np.abs(ref)), s)
    scale = np.where(scale > 0, scale, 1.0)
    return np.abs(mine - ref) / scale


def structural_zero_mismatch(mine, ref, rel=1e-12):
    """Entries zero on one side only, not counting cancellation: a value on the
    other side below `rel` times the largest magnitude of its block row is a
    difference a - b whose operands agree to the last bits on one side (the two
    builds' transcendentals differ by an ulp), not a missing or extra term."""
    mine = np.asarray(mine)
    ref = np.asarray(ref)
    differ = (ref == 0.0) != (mine == 0.0)
    if not differ.any():
        return 0
    scale = np.maximum(np.abs(ref).max(axis=-1, keepdims=True),
                       np.abs(mine).max(axis=-1, keepdims=True))
    value = np.maximum(np.abs(ref), np.abs(mine))
    cancel = differ & (value <= rel * scale)
    for idx in zip(*np.nonzero(differ)):
        print(f"zero on one side at {tuple(int(i) for i in idx)}: mine {mine[idx]:.3e}, "
              f"reference {ref[idx]:.3e}, row scale {scale[idx[:-1]][0]:.3e}")
    return int(np.sum(differ & ~cancel))


def reference_grid_states(fluid, nx, ny, seed, ref, has_well=False, nz=1, hydrostatic=False):
    """Seeded nx x ny x nz grid (dx = dy = 10 m, dz = 5 m; with `hydrostatic`, P gains
    rho_w g depth down the k index), history state from the reference's flash, then
    the seeded Newton-like perturbation.  Host arrays: (workload dict, history
    primary, history phase state, perturbed primary).

    With the Water-role closures configured (fluid.aqueous_volume), the Water
    composition block of the perturbation goes through the reference's
    aqueous-domain Newton limiter (limitAqueousCompositionNewtonIncrement,
    newton_limiter.hpp:181-245) at the closures' ceiling, as the runtime limits
    its Newton steps, so the perturbed states stay inside the closures' domain."""
    from nonenone_amd.assembly import assign_flash_result, perturb_newton_state
    from nonenone_amd.workloads import grid_workload

    w = grid_workload(nx, ny, nz, 10.0, 10.0, 5.0, fluid, seed=seed, hydrostatic=hydrostatic)
    fl = ref.flash(w["P"], w["T"], w["z"], threads=8)
    keep = fl["converged"] == 1
    assert keep.mean() > 0.95
    # non-converged history cells take a converged neighbour's state (workload construction)
    donor = np.nonzero(keep)[0][0]
    for c in np.nonzero(~keep)[0]:
        for k in fl:
            fl[k][c] = fl[k][donor]
        w["P"][c] = w["P"][donor]
    hist_p, hist_ps = assign_flash_result(fl, w["P"], has_well=has_well)
    cur = perturb_newton_state(hist_p, hist_ps, fluid.nc, seed + 1)
    if fluid.aqueous_volume:
        nc = fluid.nc
        limited = ref.newton_limit(hist_p, hist_p - cur, has_well=has_well,
                                   max_aqueous=fluid.aqueous_volume_max_unsupported,
                                   water=fluid.water_component, co2=fluid.co2_component)
        block = slice(1 + 2 * (nc - 1), 1 + 3 * (nc - 1))
        cur[:, block] = hist_p[:, block] - limited[:, block]
    return w, hist_p, hist_ps, cur
