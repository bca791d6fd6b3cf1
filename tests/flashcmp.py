"""Comparison helpers for flash results (labels exact, values 1e-10 relative)."""
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


def compare_flash(mine: dict, ref: dict):
    """Returns (n_status_mismatch, n_label_mismatch, max_rel, worst_index, ok_mask)."""
    status_ok = (mine["status"] == ref["status"]) & (mine["converged"] == ref["converged"])
    conv = status_ok & (ref["status"] == 0)
    label_ok = mine["presence"] == ref["presence"]
    n_status = int((~status_ok).sum())
    n_label = int((conv & ~label_ok).sum())
    sel = conv & label_ok
    worst = 0.0
    worst_i = -1
    if sel.any():
        per = np.zeros(len(sel))
        for key, floor in (("beta", 1.0), ("saturation", 1.0), ("x", 1.0), ("Z", 1.0),
                           ("rho_m", 1.0), ("k_go", 1e-300), ("k_wo", 1e-300)):
            d = rel_diff(mine[key], ref[key], floor).reshape(len(sel), -1).max(axis=1)
            per = np.maximum(per, d)
        per[~sel] = 0.0
        worst_i = int(np.argmax(per))
        worst = float(per[worst_i])
    return n_status, n_label, worst, worst_i, sel
