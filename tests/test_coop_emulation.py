"""The wavefront-per-state (cooperative) flash gives the serial flash's results
bit for bit.

Runs on the CPU: tests/coop_emulation.cpp compiles this library's flash for the
host with NN_HOST_WAVE_EMULATION, where 64 host threads play the 64 lanes of a
wavefront (every ballot and shuffle is a rendezvous of all of them), and compares
lane 0's result (the one the cooperative pass stores) with the serial flash
(kModePlain, the reference's control flow) on golden-fixture states, including
states that reach the lane-parallel RR3 guesses and the Soreide-Whitson
allocation seeds.
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from nonenone_amd import fluids

HERE = Path(__file__).resolve().parent
GOLDEN = HERE / "golden"
# PR-CPA's cooperative density root (a shuffle per bisection-tree node) is too
# slow to emulate with host threads; the GPU test of the cooperative pass checks
# it against the fixtures.
CASES = {"flash_nc3_pr": lambda: fluids.h2o_co2_nc10("PR"),
         "flash_nc3_sw": lambda: fluids.h2o_co2_nc10("SW"),
         "flash_nc6_sw": lambda: fluids.six_component("SW")}


@pytest.fixture(scope="module")
def emulator(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path_factory.mktemp("coop") / "coop_emulation"
    subprocess.run([hipcc, "-std=c++17", "-O1", "-ffp-contract=off", "-x", "hip",
                    "--cuda-host-only", "-pthread", str(HERE / "coop_emulation.cpp"), "-o",
                    str(exe)], check=True, capture_output=True)
    return exe


@pytest.mark.parametrize("name", sorted(CASES))
def test_cooperative_flash_equals_serial_flash(emulator, tmp_path, name):
    g = dict(np.load(GOLDEN / f"{name}.npz"))
    # every state that the fixtures' SW/CPA conditioning survey marks as reaching
    # the allocation recovery or a near tie, plus the first states of the set
    import libm_conditioning
    listed = sorted(libm_conditioning.conditioned(name))
    idx = np.unique(np.r_[np.arange(16), listed[:16]]).astype(np.int64)
    fluid = CASES[name]()
    fd = tmp_path / "fluid.bin"
    fd.write_bytes(bytes(fluid))
    st = tmp_path / "states.bin"
    with open(st, "wb") as f:
        np.array([len(idx)], np.int64).tofile(f)
        for key in ("P", "T", "z"):
            np.ascontiguousarray(g[key][idx], np.float64).tofile(f)
    out = subprocess.run([str(emulator), str(fd), str(st)], capture_output=True, text=True,
                         check=True, timeout=600)
    rows = [line.split() for line in out.stdout.splitlines()]
    assert len(rows) == len(idx)
    differ = [(int(idx[int(r[0])]), r[2:]) for r in rows if r[1] != "1"]
    assert not differ, differ[:4]
