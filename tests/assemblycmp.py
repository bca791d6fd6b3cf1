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


def structural_zero_mismatch(mine, ref):
    return int(np.sum((ref == 0.0) != (mine == 0.0)))
