"""The phase update's three-pass stability test (record -> lockstep trials -> replay)
gives bit-identical cells to the serial phase update.

Runs on the CPU: tests/lockstep_check.cpp compiles this library's phase update
for the host and drives the trial state machine one residual evaluation at a
time, as one lane of the GPU's lockstep pass does.  The GPU tests of the phase
update (test_assembly.py) check the three GPU passes against the reference.
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from nonenone_amd import fluids
from nonenone_amd._abi import FlashOptions

import refharness

HERE = Path(__file__).resolve().parent
GOLDEN = HERE / "golden"


@pytest.fixture(scope="module")
def check_exe(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path_factory.mktemp("lockstep") / "lockstep_check"
    subprocess.run([hipcc, "-std=c++17", "-O1", "-ffp-contract=off", "-x", "hip",
                    "--cuda-host-only", str(HERE / "lockstep_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    return exe


def _run(exe, fluid, T, primary, phase_state, has_well, tmp):
    fd = tmp / "fluid.bin"
    fd.write_bytes(bytes(fluid))
    cells = tmp / "cells.bin"
    n = len(T)
    with open(cells, "wb") as f:
        np.array([n, int(has_well)], np.int64).tofile(f)
        np.ascontiguousarray(T, np.float64).tofile(f)
        np.ascontiguousarray(primary, np.float64).tofile(f)
        np.ascontiguousarray(phase_state, np.float64).tofile(f)
    out = subprocess.run([str(exe), str(fd), str(cells)], capture_output=True, text=True)
    rows = np.array([[int(v) for v in line.split()] for line in out.stdout.splitlines()])
    assert rows.shape == (n, 3), out.stderr
    print(out.stderr.strip())
    return rows


def test_fixture_cells_lockstep_identical(check_exe, tmp_path):
    g = dict(np.load(GOLDEN / "assembly_nc3_pr_12x8.npz"))
    rows = _run(check_exe, fluids.h2o_co2_nc10("PR"), g["T"], g["cur_primary"],
                g["cur_phase_state"], False, tmp_path)
    assert rows[:, 2].all(), np.nonzero(rows[:, 2] == 0)[0]
    assert rows[:, 1].sum() > len(rows) // 2   # most cells take the recorded path


@pytest.mark.skipif(not refharness.available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("model", ["CPA", "SW", "PR6"])
def test_reference_history_cells_lockstep_identical(check_exe, tmp_path, model):
    from nonenone_amd import workloads
    from nonenone_amd.assembly import assign_flash_result, perturb_newton_state
    fluid = {"CPA": fluids.cpa_h2o_co2_nc10, "SW": lambda: fluids.h2o_co2_nc10("SW"),
             "PR6": lambda: fluids.six_component("PR")}[model]()
    w = workloads.grid_workload(12, 10, 1, 10.0, 10.0, 5.0, fluid, seed=31)
    ref = refharness.Reference(fluid, FlashOptions.defaults())
    flash = ref.flash(w["P"], w["T"], w["z"], threads=4)
    primary, ps = assign_flash_result(flash, w["P"], has_well=True)
    pert = perturb_newton_state(primary, ps, fluid.nc, seed=32, push_fraction=0.1)
    rows = _run(check_exe, fluid, w["T"], pert, ps, True, tmp_path)
    assert rows[:, 2].all(), np.nonzero(rows[:, 2] == 0)[0]
    if model != "SW":
        # (Soreide-Whitson cells of two or more phases certify their perturbed
        # compositions with a restricted flash first and finish in the first pass)
        assert rows[:, 1].sum() > len(rows) // 2
