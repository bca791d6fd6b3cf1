"""Per-state libm conditioning of the reference's outputs (test helper).

tests/golden/libm_sensitivity.json (written by tests/golden/make_sensitivity.py)
lists, per test set, the states whose reference result is conditioned by the
math library: results that move by more than round-off under 1-ulp changes of
exp/log/pow/cos/acos (their `band`: the largest relative movement over eight
seeded perturbation runs), states that reach the Soreide-Whitson exact-mass
allocation recovery, and states where a lowest-Gibbs selection compared two
candidates whose reduced Gibbs energies agree to 1e-12.  The GPU evaluates those
functions with the device math library, which is within 1 ulp of the host's but
not bit-identical, so every listed state is held to a bound of its own instead
of the common tolerance:

  * the reference's answer:    values within max(tol, BAND_FACTOR x band), where
    `band` is the spread of the perturbed runs that stay on that answer;
  * another recorded outcome:  a state whose perturbed runs reached more than one
    split (`outcomes`: distinct phase-fraction vectors, the reference's first)
    passes when the GPU's phase fractions match one of them within OUTCOME_TOL:
    the reference's own algorithm returns that answer under 1-ulp changes of
    its math library;
  * near-tie selection or allocation recovery, where the caller supplies the
    two answers' dimensionless Gibbs energies (flashcmp.gibbs_diff): an energy
    difference within GIBBS_TOL (a degenerate split, e.g. mass moved between
    two phases of one composition, moves the values at an unchanged energy).

Status and labels are compared exactly for every state by the callers.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

PATH = Path(__file__).resolve().parent / "golden" / "libm_sensitivity.json"
BAND_FACTOR = 4.0
GIBBS_TOL = 1.0e-9
OUTCOME_TOL = 1.0e-6


def entries(name: str) -> dict[int, dict]:
    data = json.loads(PATH.read_text())
    entry = data["sets"].get(name)
    return {} if entry is None else {int(k): v for k, v in entry["conditioned"].items()}


def conditioned(name: str) -> set[int]:
    return set(entries(name))


def bounds(name: str, n: int, tol: float) -> np.ndarray:
    """Per-state relative bound: tol, or the listed state's own bound."""
    lim = np.full(n, tol)
    for i, e in entries(name).items():
        if i >= n:
            continue
        lim[i] = max(tol, BAND_FACTOR * float(e["band"]))
    return lim


def matched_outcome(entry: dict, beta) -> int:
    """Index of the recorded outcome whose phase fractions match `beta` within
    OUTCOME_TOL, or -1."""
    if beta is None:
        return -1
    for k, b in enumerate(entry.get("outcomes", [])):
        if np.max(np.abs(np.asarray(b) - np.asarray(beta))) <= OUTCOME_TOL:
            return k
    return -1


def split_value_failures(per_state: np.ndarray, compared: np.ndarray, tol: float, name: str,
                         gibbs=None, beta=None):
    """(failures, conditioned) among compared states: failures exceed their own
    bound (tol for unlisted states) and match no other recorded outcome;
    conditioned states exceed tol but stay within what is recorded for them.
    `gibbs` (optional, per state): relative difference of the two answers'
    Gibbs energies (settles near-tie and allocation-recovery states); `beta`
    (optional, n x 3): the GPU's phase fractions, matched against the recorded
    outcomes.  Prints each conditioned state's value and how it was settled."""
    per_state = np.asarray(per_state)
    lim = bounds(name, len(per_state), tol)
    listed = entries(name)
    over_tol = compared & (per_state > tol)
    failures, within = [], []
    for i in np.nonzero(over_tol)[0]:
        i = int(i)
        e = listed.get(i)
        k = matched_outcome(e, None if beta is None else beta[i]) if e is not None else -1
        if per_state[i] <= lim[i]:
            within.append(i)
            print(f"  state {i}: rel diff {per_state[i]:.3e}, bound {lim[i]:.3e}")
        elif k > 0:
            within.append(i)
            print(f"  state {i}: rel diff {per_state[i]:.3e}; phase fractions match recorded "
                  f"outcome {k} of the reference's algorithm under 1-ulp libm changes")
        elif (e is not None and (e["gibbs_tie"] or e["allocation_recovery"]) and gibbs is not None
              and gibbs[i] <= GIBBS_TOL):
            within.append(i)
            print(f"  state {i}: rel diff {per_state[i]:.3e} over {lim[i]:.3e}; Gibbs energies "
                  f"agree to {gibbs[i]:.1e} (near-tie / allocation state)")
        else:
            failures.append(i)
            extra = "" if gibbs is None else f", Gibbs energy diff {gibbs[i]:.1e}"
            print(f"  state {i}: rel diff {per_state[i]:.3e}, bound {lim[i]:.3e}{extra} FAILS")
    return failures, within
