"""Comparison helpers for flash results (labels exact, values 1e-10 relative).

Values are compared where the reference defines them physically: phase
fractions and saturations of every slot, and compositions, Z, molar densities
of the phases present; K_go where oil and gas are present, K_wo where oil and
water are present.  Absent-slot placeholders (a composition copied from z,
Z = 1, K from a discarded iterate) depend on which recovery path produced the
result and are reported separately by `absent_slot_diff`.
"""
from __future__ import annotations

import numpy as np

REL_TOL = 1.0e-10   # north-star tolerance for FP64 values


def to_numpy(res) -> dict:
    if isinstance(res, dict):
        return {k: np.asarray(v) for k, v in res.items()}
    return {k: getattr(res, k).cpu().numpy() for k in (
        "presence", "converged", "status", "iterations", "beta", "saturation", "x", "Z",
        "rho_m", "k_go", "k_wo")}


def rel_diff(a, b, floor):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)
    d = np.abs(a - b) / scale
    d[(a == b)] = 0.0
    return d


def _present(presence):
    bits = np.asarray(presence).astype(np.int64)
    return np.stack([(bits >> p) & 1 for p in range(3)], axis=1).astype(bool)


def per_state_value_diff(mine: dict, ref: dict) -> np.ndarray:
    pres = _present(ref["presence"])
    n = len(pres)
    per = np.zeros(n)
    per = np.maximum(per, rel_diff(mine["beta"], ref["beta"], 1.0).max(axis=1))
    per = np.maximum(per, rel_diff(mine["saturation"], ref["saturation"], 1.0).max(axis=1))
    for key in ("Z", "rho_m"):
        d = np.where(pres, rel_diff(mine[key], ref[key], 1.0), 0.0)
        per = np.maximum(per, d.max(axis=1))
    dx = rel_diff(mine["x"], ref["x"], 1.0).max(axis=2)
    per = np.maximum(per, np.where(pres, dx, 0.0).max(axis=1))
    og = pres[:, 0] & pres[:, 1]
    ow = pres[:, 0] & pres[:, 2]
    per = np.maximum(per, np.where(og, rel_diff(mine["k_go"], ref["k_go"], 1e-300).max(axis=1), 0.0))
    per = np.maximum(per, np.where(ow, rel_diff(mine["k_wo"], ref["k_wo"], 1e-300).max(axis=1), 0.0))
    return per


def absent_slot_diff(mine: dict, ref: dict) -> float:
    pres = _present(ref["presence"])
    dx = rel_diff(mine["x"], ref["x"], 1.0).max(axis=2)
    dz = rel_diff(mine["Z"], ref["Z"], 1.0)
    return float(max(np.where(~pres, dx, 0.0).max(initial=0.0),
                     np.where(~pres, dz, 0.0).max(initial=0.0)))


def compare_flash(mine: dict, ref: dict):
    """Returns (n_status_mismatch, n_label_mismatch, max_rel, worst_index, compared_mask)."""
    status_ok = (mine["status"] == ref["status"]) & (mine["converged"] == ref["converged"])
    conv = status_ok & (ref["status"] == 0)
    label_ok = mine["presence"] == ref["presence"]
    n_status = int((~status_ok).sum())
    n_label = int((conv & ~label_ok).sum())
    sel = conv & label_ok
    worst = 0.0
    worst_i = -1
    if sel.any():
        per = per_state_value_diff(mine, ref)
        per[~sel] = 0.0
        worst_i = int(np.argmax(per))
        worst = float(per[worst_i])
    return n_status, n_label, worst, worst_i, sel
