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

  * band only:                 max(tol, BAND_FACTOR x band);
  * near-tie selection or allocation recovery (equally converged answers, one
    kept by the last bits of the arithmetic, each stopped at the flash's 1e-9
    fugacity tolerance):       max(tol, BAND_FACTOR x band, TIE_BOUND).

Status and labels are compared exactly for every state by the callers.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

PATH = Path(__file__).resolve().parent / "golden" / "libm_sensitivity.json"
BAND_FACTOR = 4.0
TIE_BOUND = 1.0e-6


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
        b = max(tol, BAND_FACTOR * float(e["band"]))
        if e["gibbs_tie"] or e["allocation_recovery"]:
            b = max(b, TIE_BOUND)
        lim[i] = b
    return lim


def split_value_failures(per_state: np.ndarray, compared: np.ndarray, tol: float, name: str):
    """(failures, conditioned) among compared states: failures exceed their own
    bound (tol for unlisted states); conditioned states exceed tol but stay
    within their listed bound.  Prints each conditioned state's value and bound."""
    per_state = np.asarray(per_state)
    lim = bounds(name, len(per_state), tol)
    over_tol = compared & (per_state > tol)
    failures = [int(i) for i in np.nonzero(compared & (per_state > lim))[0]]
    within = [int(i) for i in np.nonzero(over_tol & (per_state <= lim))[0]]
    for i in within + failures:
        print(f"  state {i}: rel diff {per_state[i]:.3e}, bound {lim[i]:.3e}")
    return failures, within
