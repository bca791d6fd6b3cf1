"""Writes tests/golden/libm_sensitivity.json: which flash states of the SW and
CPA test sets, and which cells of the SW and closure phase-update grids, have
results conditioned by the math library.

For each state the host build of this library's flash (tests/golden/
sw_sensitivity.cpp; on the host it reproduces the reference build to ~1e-15)
is re-run under eight seeded 1-ulp perturbations of exp/log/pow/cos/acos
(64 for states that reach the allocation recovery or a near tie, whose
alternative answers can be rare; the distinct phase-fraction vectors reached
are recorded as the state's `outcomes`).
States whose values move by more than 1e-12 (relative, floor 1) — or whose
label moves — are recorded with that band, and so is every state whose flash
reaches the Soreide-Whitson exact-mass allocation recovery (a multi-start
Gibbs minimisation, 23 seeds, argmin with strict comparisons, each run stopped
at the 1e-9 fugacity tolerance): which seed wins, and so which of several
equally converged answers and K-value records is returned, follows 1-ulp
differences of the math library; and every state where any lowest-Gibbs
selection (stable reduced set, SW role permutations, allocation seeds) compared
two candidates whose reduced Gibbs energies agree to 1e-12 — equivalent
solutions, one reached by Newton to ~1e-13 and the other by SSI stopped at the
1e-9 tolerance, kept or replaced by the last bits of the arithmetic.  Those states' reference values are defined
only up to the host libm; the GPU tests compare their status and labels
exactly and report, without asserting, their value differences.

    python tests/golden/make_sensitivity.py [set names: recompute only those]
"""
from __future__ import annotations

import json
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from nonenone_amd import fluids  # noqa: E402
from nonenone_amd.workloads import batched_flash_states  # noqa: E402

THRESHOLD = 1.0e-12
SALTS = 8              # perturbed runs per state
SENSITIVE_SALTS = 64   # per allocation-recovery or near-tie state
SETS = {
    # name: (fluid factory, source) — golden fixture file or the live seed-7 set
    "flash_nc3_sw": (lambda: fluids.h2o_co2_nc10("SW"), "golden"),
    "flash_nc6_sw": (lambda: fluids.six_component("SW"), "golden"),
    "live_nc3_sw": (lambda: fluids.h2o_co2_nc10("SW"), "live"),
    "live_nc6_sw": (lambda: fluids.six_component("SW"), "live"),
    "flash_nc3_cpa": (fluids.cpa_h2o_co2_nc10, "golden"),
    "live_nc3_cpa": (fluids.cpa_h2o_co2_nc10, "live"),
}


def build(tmp: Path) -> Path:
    exe = tmp / "sw_sensitivity"
    cmd = ["hipcc", "-std=c++17", "-O2", "-ffp-contract=off", "-x", "hip", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           str(HERE / "sw_sensitivity.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def states(name, fluid, source):
    if source == "golden":
        g = np.load(HERE / f"{name}.npz")
        return g["P"], g["T"], g["z"]
    return batched_flash_states(256, fluid, seed=7, near_critical_fraction=0.1)


def parse(stdout: str) -> dict:
    """Listed states: label flips, band above THRESHOLD, allocation recovery, a
    near tie, or more than one outcome.  `outcomes`: the distinct phase-fraction
    vectors the perturbed runs that keep the label reached (the unperturbed one
    first); `band`: the spread of the runs that stay in the first outcome."""
    flagged = {}
    for line in stdout.split("\n"):
        if not line.strip():
            continue
        f = line.split()
        idx, _, flips, band, alloc, tie, nout = f[:7]
        vals = [float(v) for v in f[7:]]
        outcomes = [vals[3 * k:3 * k + 3] for k in range(int(nout))]
        if (int(flips) > 0 or float(band) > THRESHOLD or int(alloc) or int(tie) or
                len(outcomes) > 1):
            flagged[idx] = {"label_flips": int(flips), "band": float(band),
                            "allocation_recovery": int(alloc), "gibbs_tie": int(tie),
                            "outcomes": outcomes}
    return flagged


UPDATE_SETS = {
    # name: (fluid factory, well unknown) — the phase-update inputs of the grid
    # tests (60 x 20, seed 21)
    "update_sw_60x20_seed21": (lambda: fluids.h2o_co2_nc10("SW"), False),
    "update_sw6_60x20_seed21": (lambda: fluids.six_component("SW"), False),
    # PR with the Water-role closures (IAPWS-IF97 + Garcia volume; the viscosity
    # closure does not enter the phase update)
    "update_praq_60x20_seed21": (
        lambda: fluids.with_aqueous_closures(fluids.h2o_co2_nc10("PR")), False),
    "update_praq_60x20_seed21_well": (
        lambda: fluids.with_aqueous_closures(fluids.h2o_co2_nc10("PR")), True),
}


def update_sets(exe: Path, tmp: Path, only=None) -> dict:
    """The phase-update sets of UPDATE_SETS (all, or the names in `only`)."""
    import assemblycmp
    import refharness
    from nonenone_amd._abi import FlashOptions

    out = {}
    for name, (factory, has_well) in UPDATE_SETS.items():
        if only is not None and name not in only:
            continue
        fluid = factory()
        ref = refharness.Reference(fluid, FlashOptions.defaults())
        w, _, hist_ps, cur = assemblycmp.reference_grid_states(fluid, 60, 20, 21, ref,
                                                               has_well=has_well)
        (tmp / "desc.bin").write_bytes(bytes(fluid))
        with open(tmp / "cells.bin", "wb") as f:
            np.array([len(cur), int(has_well)], np.int64).tofile(f)
            np.ascontiguousarray(w["T"], np.float64).tofile(f)
            np.ascontiguousarray(cur, np.float64).tofile(f)
            np.ascontiguousarray(hist_ps, np.float64).tofile(f)
        res = subprocess.run([str(exe), "update", str(tmp / "desc.bin"), str(tmp / "cells.bin"),
                              str(SALTS), str(SENSITIVE_SALTS)], check=True, capture_output=True,
                             text=True)
        flagged = parse(res.stdout)
        print(f"{name}: {len(flagged)} of {len(cur)} cells libm-conditioned", flush=True)
        out[name] = {"n": int(len(cur)), "conditioned": flagged}
    return out


def flash_set(exe: Path, tmp: Path, name: str) -> tuple[str, dict]:
    factory, source = SETS[name]
    fluid = factory()
    P, T, z = states(name, fluid, source)
    d = tmp / name
    d.mkdir()
    (d / "desc.bin").write_bytes(bytes(fluid))
    with open(d / "states.bin", "wb") as f:
        np.array([len(P)], np.int64).tofile(f)
        np.ascontiguousarray(P, np.float64).tofile(f)
        np.ascontiguousarray(T, np.float64).tofile(f)
        np.ascontiguousarray(z, np.float64).tofile(f)
    res = subprocess.run([str(exe), "flash", str(d / "desc.bin"), str(d / "states.bin"),
                          str(SALTS), str(SENSITIVE_SALTS)], check=True, capture_output=True,
                         text=True)
    flagged = parse(res.stdout)
    print(f"{name}: {len(flagged)} of {len(P)} states libm-conditioned", flush=True)
    return name, {"n": int(len(P)), "conditioned": flagged}


def main() -> None:
    """All sets, or (arguments) only the named sets, merged into the existing file."""
    from concurrent.futures import ThreadPoolExecutor

    only = set(sys.argv[1:]) or None
    path = HERE / "libm_sensitivity.json"
    out = {"threshold": THRESHOLD, "salts": SALTS, "sensitive_salts": SENSITIVE_SALTS, "sets": {}}
    if only is not None:
        out["sets"] = json.loads(path.read_text())["sets"]
    names = [n for n in SETS if only is None or n in only]
    with tempfile.TemporaryDirectory() as td:
        tmp = Path(td)
        exe = build(tmp)
        with ThreadPoolExecutor(max_workers=max(1, len(names))) as pool:
            futures = [pool.submit(flash_set, exe, tmp, name) for name in names]
            out["sets"].update(update_sets(exe, tmp, only))
            for fut in futures:
                name, entry = fut.result()
                out["sets"][name] = entry
    path.write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main()
