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


def gibbs_energy(ref, P, T, out) -> np.ndarray:
    """Dimensionless Gibbs energy sum_p beta_p sum_i x_pi (ln x_pi + ln phi_pi) of each
    flash result, over the slots with beta > 0, with the reference build's EOS
    (`ref`: refharness.Reference) evaluated for slot p in role p (the role of a
    canonical slot).  Two converged answers of one state with the same labels whose
    energies agree are equally valid minima of the same problem."""
    beta = np.asarray(out["beta"], np.float64)
    x = np.asarray(out["x"], np.float64)
    n, _, nc = x.shape
    idx, slot = np.nonzero(beta > 0.0)
    g = np.zeros(n)
    if len(idx) == 0:
        return g
    xs = np.maximum(x[idx, slot], 1e-300)
    _, lnphi, _, status = ref.eos_phase(np.asarray(P, np.float64)[idx], np.asarray(T, np.float64)[idx],
                                        xs, slot.astype(np.int32))
    term = beta[idx, slot] * np.sum(xs * (np.log(xs) + lnphi), axis=1)
    term[status != 0] = np.nan
    np.add.at(g, idx, term)
    return g


def gibbs_diff(ref, P, T, mine, theirs) -> np.ndarray:
    """|G(mine) - G(theirs)| / max(1, |G(theirs)|) per state (gibbs_energy)."""
    a = gibbs_energy(ref, P, T, mine)
    b = gibbs_energy(ref, P, T, theirs)
    return np.abs(a - b) / np.maximum(1.0, np.abs(b))


def update_as_flash(P, primary, phase_state, nc) -> dict:
    """The beta / x fields of updated cell states (primary: p, x_o[0..nc-2], x_g[..],
    x_w[..], ...; phase state: flag, K_go, K_wo, z, beta[3], ...), for gibbs_energy."""
    primary = np.asarray(primary, np.float64)
    n = len(primary)
    x = np.zeros((n, 3, nc))
    for p in range(3):
        xs = primary[:, 1 + p * (nc - 1):1 + (p + 1) * (nc - 1)]
        x[:, p, :nc - 1] = xs
        x[:, p, nc - 1] = 1.0 - xs.sum(axis=1)
    beta = np.asarray(phase_state, np.float64)[:, 1 + 3 * nc:4 + 3 * nc]
    return {"beta": beta, "x": x}


def update_gibbs_diff(ref, T, mine_prim, mine_ps, ref_prim, ref_ps, nc) -> np.ndarray:
    """gibbs_diff of two updated states of the same cells (pressure of `ref_prim`)."""
    P = np.asarray(ref_prim, np.float64)[:, 0]
    return gibbs_diff(ref, P, T, update_as_flash(P, mine_prim, mine_ps, nc),
                      update_as_flash(P, ref_prim, ref_ps, nc))
