"""The PR-CPA density root with certified skips (nonenone_amd/csrc/cpa.hpp,
cpaCertifiedSign) returns the plain bracket-and-bisection search's root bit for
bit (cubic_eos.hpp:2090-2178).

Runs on the CPU: the same source is compiled for the host twice, with
NN_CPA_CERTIFIED_SKIP=0 (every bracket point and midpoint evaluated with the
exact site solve, the reference's search) and =1 (the build the GPU runs), and
the two builds' outputs are compared byte for byte:
  * Z and ln(phi) of single-phase evaluations over a spread of states
    (tests/cpa_host_check.cpp);
  * the whole phase update of a PR-CPA grid with a well unknown
    (tests/phase_update_dump.cpp): statuses, flags, primaries and phase states.
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from nonenone_amd import fluids

import refharness

HERE = Path(__file__).resolve().parent


def _hipcc():
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    return hipcc


def _build(src, exe, skip):
    subprocess.run([_hipcc(), "-std=c++17", "-O1", "-ffp-contract=off",
                    f"-DNN_CPA_CERTIFIED_SKIP={skip}", "-x", "hip", "--cuda-host-only",
                    str(HERE / src), "-o", str(exe)], check=True, capture_output=True)


def _states(n, nc, seed):
    rng = np.random.default_rng(seed)
    P = np.exp(rng.uniform(np.log(1.0e5), np.log(8.0e7), n))
    T = rng.uniform(275.0, 650.0, n)
    alphas = [np.ones(nc), np.r_[8.0, np.ones(nc - 1)], np.r_[1.0, 8.0, np.ones(nc - 2)]]
    x = np.concatenate([rng.dirichlet(a, n // 3 + 1) for a in alphas])[:n]
    rng.shuffle(x)
    role = rng.integers(0, 3, n).astype(np.int32)
    return P, T, x, role


def test_certified_density_root_is_bit_identical_single_phase(tmp_path):
    fluid = fluids.cpa_h2o_co2_nc10()
    n = 30000
    P, T, x, role = _states(n, fluid.nc, seed=5)
    fd = tmp_path / "fluid.bin"
    fd.write_bytes(bytes(fluid))
    st = tmp_path / "states.bin"
    with open(st, "wb") as f:
        np.array([n], np.int64).tofile(f)
        for a in (P, T, x):
            np.ascontiguousarray(a, np.float64).tofile(f)
        role.tofile(f)
    outs = []
    for skip in (0, 1):
        exe = tmp_path / f"cpa_host_check_{skip}"
        _build("cpa_host_check.cpp", exe, skip)
        outs.append(subprocess.run([str(exe), str(fd), str(st)], capture_output=True, text=True,
                                   check=True).stdout.splitlines())
    assert len(outs[0]) == n and len(outs[1]) == n
    status = np.array([int(line.split()[0]) for line in outs[0]])
    # the sample covers both branches and evaluations that succeed
    assert (status == 0).sum() > 0.9 * n
    differ = [i for i in range(n) if outs[0][i] != outs[1][i]]
    assert not differ, [(i, outs[0][i], outs[1][i]) for i in differ[:5]]


def test_certified_density_root_is_bit_identical_phase_update(tmp_path):
    if not refharness.available():
        pytest.skip("reference build (oracle/_ref) not built")
    import assemblycmp
    from nonenone_amd._abi import FlashOptions

    fluid = fluids.cpa_h2o_co2_nc10()
    ref = refharness.Reference(fluid, FlashOptions.defaults())
    w, _, hist_ps, cur = assemblycmp.reference_grid_states(fluid, 24, 20, 41, ref, has_well=True,
                                                           nz=3, hydrostatic=True)
    n = w["ncell"]
    fd = tmp_path / "fluid.bin"
    fd.write_bytes(bytes(fluid))
    cells = tmp_path / "cells.bin"
    with open(cells, "wb") as f:
        np.array([n, 1], np.int64).tofile(f)
        for a in (w["T"], cur, hist_ps):
            np.ascontiguousarray(a, np.float64).tofile(f)
    blobs = []
    for skip in (0, 1):
        exe = tmp_path / f"phase_update_dump_{skip}"
        _build("phase_update_dump.cpp", exe, skip)
        out = tmp_path / f"out_{skip}.bin"
        subprocess.run([str(exe), str(fd), str(cells), str(out)], check=True, capture_output=True)
        blobs.append(out.read_bytes())
    np_, nps = 3 * fluid.nc + 2, 3 * fluid.nc + 8
    rec = np.dtype([("prim", "<f8", np_), ("ps", "<f8", nps), ("status", "<i4"), ("flags", "u1")])
    a, b = (np.frombuffer(x, rec) for x in blobs)
    assert len(a) == n
    # the grid exercises the stability test and the re-flash (flag 4: missing phase unstable)
    assert ((a["flags"] & 4) != 0).sum() > n // 10
    assert blobs[0] == blobs[1], np.nonzero(a.tobytes() != b.tobytes())
