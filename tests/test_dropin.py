"""Drop-in check of the FlashBackend seam.

oracle/_ref/dropin_check (built by `make -C oracle dropin` from the reference's
own headers) runs the reference's FullyCompositionalThreePhaseEquilibrium::
updatePhaseState (state/three_phase_equilibrium.hpp:282-496) on every cell with
its default CubicThreePhaseFlash and again with GpuThreePhaseFlash
(include/nonenone_amd_flash_backend.hpp), which serves flash, flashRestricted
and stabilityTest from this library's GPU kernels through the C ABI.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from nonenone_amd import fluids
from nonenone_amd._abi import FlashOptions

import assemblycmp
import refharness

ROOT = Path(__file__).resolve().parents[1]
DROPIN = ROOT / "oracle" / "_ref" / "dropin_check"


def test_backend_header_compiles_without_reference_types():
    """include/nonenone_amd.hpp (the reference-free C++ wrapper) is self-contained."""
    src = '#include "nonenone_amd.hpp"\nint main() { return 0; }\n'
    proc = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__",
                           "-I", str(ROOT / "include"), "-I", "/opt/rocm/include", "-x", "c++", "-"],
                          input=src, capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr[-2000:]


@pytest.mark.gpu
@pytest.mark.skipif(not (refharness.available() and DROPIN.exists()),
                    reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("model", ["PR", "CPA"])
def test_gpu_flash_backend_dropin_matches_default_backend(model, tmp_path):
    fluid = fluids.cpa_h2o_co2_nc10() if model == "CPA" else fluids.h2o_co2_nc10(model)
    ref = refharness.Reference(fluid, FlashOptions.defaults())
    w, _, hist_ps, cur = assemblycmp.reference_grid_states(fluid, 20, 10, 41, ref)
    path = tmp_path / "cells.bin"
    with open(path, "wb") as f:
        f.write(bytes(fluid))
        np.array([len(w["T"])], np.int64).tofile(f)
        np.ascontiguousarray(w["T"], np.float64).tofile(f)
        np.ascontiguousarray(cur, np.float64).tofile(f)
        np.ascontiguousarray(hist_ps, np.float64).tofile(f)
    proc = subprocess.run([str(DROPIN), str(path)], capture_output=True, text=True, timeout=300)
    print(proc.stdout.strip(), proc.stderr[-500:])
    assert proc.returncode == 0, proc.stdout + proc.stderr[-2000:]
