"""Host build of the device kernels of one Newton state against the reference.

tests/kernel_emulation.cpp compiles the __global__ bodies of the phase update's
passes (lane-per-cell, lockstep trials, replay, re-flash, and the cooperative
pass on 64 emulated lanes), the accumulation and the residual + BSR assembly as
host code and runs them in the library's launch sequence, with the library's
buffer sizes.  With the host math library the kernels' arithmetic is the
reference's, so the results must equal the reference build bit for bit:
statuses, flags, primaries, phase states, accumulation, residual, BSR and upwind
decisions.  (Under -fsanitize=address,undefined the same program checks the
kernels' memory accesses; a build with those flags is not part of this test.)"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import assemblycmp
import refharness
from nonenone_amd import fluids
from nonenone_amd._abi import FlashOptions

HERE = Path(__file__).resolve().parent
DT = 86400.0


@pytest.fixture(scope="module")
def emu_exe(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path_factory.mktemp("kernel_emulation") / "kernel_emulation"
    subprocess.run([hipcc, "-std=c++17", "-O1", "-ffp-contract=off", "-x", "hip",
                    "--cuda-host-only", str(HERE / "kernel_emulation.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    return exe


def _run(exe, fluid, w, cur, hist_p, hist_ps, r_prim, r_ps, r_acc, has_well, tmp):
    t = w["topo"]
    n, nconn, nnzb = len(w["T"]), len(t["conn_neighbor"]), len(t["col_blocks"])
    (tmp / "fluid.bin").write_bytes(bytes(fluid))
    with open(tmp / "case.bin", "wb") as f:
        np.array([n, int(has_well), nconn, nnzb], np.int64).tofile(f)
        np.array([DT]).tofile(f)
        for a, dt in [(w["T"], np.float64), (cur, np.float64), (hist_ps, np.float64),
                      (hist_p, np.float64), (r_prim, np.float64), (r_ps, np.float64),
                      (r_acc, np.float64), (w["volume"], np.float64), (w["porosity"], np.float64),
                      (t["conn_offsets"], np.int64), (t["conn_neighbor"], np.int64),
                      (t["conn_trans"], np.float64), (t["conn_gdz"], np.float64),
                      (t["row_offsets"], np.int64), (t["col_blocks"], np.int64)]:
            np.ascontiguousarray(a, dt).tofile(f)
    proc = subprocess.run([str(exe), str(tmp / "fluid.bin"), str(tmp / "case.bin"),
                           str(tmp / "out.bin")], capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr
    print(proc.stderr.strip())
    nc = fluid.nc
    npv, nps = 3 * nc + 1 + (1 if has_well else 0), 3 * nc + 8
    b = (tmp / "out.bin").read_bytes()
    out, off = {}, 0
    for name, dt, shape in [("status", np.int32, (n,)), ("flags", np.uint8, (n,)),
                            ("primary", np.float64, (n, npv)), ("phase_state", np.float64, (n, nps)),
                            ("acc", np.float64, (n, nc)), ("residual", np.float64, (n, npv)),
                            ("bsr", np.float64, (nnzb, npv, npv)),
                            ("decisions", np.uint8, (nconn, 3)), ("asm_status", np.int32, (n,))]:
        cnt = int(np.prod(shape))
        out[name] = np.frombuffer(b, dt, cnt, off).reshape(shape)
        off += cnt * np.dtype(dt).itemsize
    assert off == len(b)
    return out


@pytest.mark.skipif(not refharness.available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("model,has_well", [("PR", True), ("CPA", False), ("PR6", False)])
def test_host_build_of_the_kernels_matches_reference_build(emu_exe, tmp_path, model, has_well):
    fluid = {"PR": lambda: fluids.h2o_co2_nc10("PR"), "CPA": fluids.cpa_h2o_co2_nc10,
             "PR6": lambda: fluids.six_component("PR")}[model]()
    ref = refharness.Reference(fluid, FlashOptions.defaults())
    w, hist_p, hist_ps, cur = assemblycmp.reference_grid_states(fluid, 12, 8, 21, ref,
                                                                has_well=has_well)
    r_prim, r_ps, r_status, r_flags = ref.phase_update(w["T"], cur, hist_ps, has_well, threads=8)
    args = (w["topo"], w["volume"], w["porosity"], w["T"])
    _, _, _, r_acc = ref.assembly(*args, hist_p, hist_ps, np.zeros((len(w["T"]), 3)), DT,
                                  has_well=has_well, want_residual=False, want_bsr=False,
                                  threads=8)
    r_res, r_bsr, r_dec, _ = ref.assembly(*args, r_prim, r_ps, r_acc, DT, has_well=has_well,
                                          threads=8)
    out = _run(emu_exe, fluid, w, cur, hist_p, hist_ps, r_prim, r_ps, r_acc, has_well, tmp_path)
    assert np.array_equal(out["status"], r_status)
    assert np.array_equal(out["flags"], r_flags)
    assert np.array_equal(out["primary"], r_prim)
    assert np.array_equal(out["phase_state"], r_ps)
    assert np.all(out["asm_status"] == 0)
    assert np.array_equal(out["decisions"], r_dec)
    acc_err = np.abs(out["acc"] - r_acc) / np.abs(r_acc)
    res_err = assemblycmp.residual_error(out["residual"], r_res)
    jac_err = assemblycmp.bsr_error(out["bsr"], r_bsr)
    print(f"{model}: accumulation {acc_err.max():.2e}, residual {res_err.max():.2e}, "
          f"Jacobian {jac_err.max():.2e}")
    if model == "CPA":
        # the PR-CPA properties take the density root's value through its two
        # implicit-function corrections (root + slope terms, cubic_eos.hpp:2050-2370),
        # evaluated on derivative numbers: last-bit differences from the reference's
        # scalar evaluation
        assert acc_err.max() <= 1e-14 and res_err.max() <= 1e-13 and jac_err.max() <= 1e-13
        assert assemblycmp.structural_zero_mismatch(out["bsr"], r_bsr) == 0
    else:
        assert np.array_equal(out["acc"], r_acc)
        assert np.array_equal(out["residual"], r_res)
        assert np.array_equal(out["bsr"], r_bsr)
