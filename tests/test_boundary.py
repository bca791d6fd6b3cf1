"""The boundary's host-side duties of the reference runtime: the all-or-none phase
update (natural_petsc_runtime_state.inc:166-334) and the per-cell column insertion
of the Jacobian (natural_petsc_runtime_jacobian.inc:279-299)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from nonenone_amd import _abi, fluids, workloads
from nonenone_amd._abi import FlashOptions

import assemblycmp
import refharness

DT = 86400.0


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


@pytest.mark.skipif(not _abi.library_path().exists(), reason="library not built")
@pytest.mark.parametrize("has_well", [False, True])
def test_cell_columns_reproduce_the_matrix(has_well):
    """nn_bsr_cell_columns regroups the BSR values by column in the reference's
    insertion order; adding every column's blocks into a dense matrix, as
    MatSetValuesBlocked(..., ADD_VALUES) does, gives back the matrix the BSR
    values describe, and each block lands exactly once."""
    nx, ny, nz = 5, 4, 3
    t = workloads.structured_topology(nx, ny, nz, 10.0, 10.0, 5.0, np.full(nx * ny * nz, 1e-13))
    n = nx * ny * nz
    npv = 11 if has_well else 10
    nnzb = int(t["row_offsets"][-1])
    rng = np.random.Generator(np.random.PCG64(3))
    bsr = rng.standard_normal((nnzb, npv, npv))
    values = np.zeros_like(bsr)
    rows = np.zeros(nnzb, np.int64)
    offs = np.zeros(n + 1, np.int64)
    rc = _abi.lib().nn_bsr_cell_columns(n, npv, _p(t["conn_offsets"]), _p(t["conn_neighbor"]),
                                        _p(t["row_offsets"]), _p(t["col_blocks"]), _p(bsr),
                                        _p(values), _p(rows), _p(offs))
    assert rc == 0
    assert offs[-1] == nnzb
    dense_ref = np.zeros((n * npv, n * npv))
    for r in range(n):
        for k in range(t["row_offsets"][r], t["row_offsets"][r + 1]):
            c = t["col_blocks"][k]
            dense_ref[r * npv:(r + 1) * npv, c * npv:(c + 1) * npv] = bsr[k]
    dense = np.zeros_like(dense_ref)
    hits = np.zeros((n, n), np.int64)
    for c in range(n):
        seg = range(offs[c], offs[c + 1])
        assert rows[offs[c]] == c                      # self first
        assert list(rows[offs[c] + 1:offs[c + 1]]) == list(
            t["conn_neighbor"][t["conn_offsets"][c]:t["conn_offsets"][c + 1]])
        for q in seg:
            r = rows[q]
            dense[r * npv:(r + 1) * npv, c * npv:(c + 1) * npv] += values[q]
            hits[r, c] += 1
    assert np.array_equal(dense, dense_ref)
    assert hits.max() == 1 and hits.sum() == nnzb


@pytest.mark.skipif(not _abi.library_path().exists(), reason="library not built")
def test_cell_columns_refuse_a_pattern_without_the_block():
    nx, ny, nz = 3, 3, 1
    t = workloads.structured_topology(nx, ny, nz, 10.0, 10.0, 5.0, np.full(9, 1e-13))
    cols = t["col_blocks"].copy()
    cols[t["row_offsets"][4]] = 8            # first block of row 4 no longer (4, 3)
    nnzb = int(t["row_offsets"][-1])
    bsr = np.zeros((nnzb, 10, 10))
    out = np.zeros_like(bsr)
    rows = np.zeros(nnzb, np.int64)
    offs = np.zeros(10, np.int64)
    rc = _abi.lib().nn_bsr_cell_columns(9, 10, _p(t["conn_offsets"]), _p(t["conn_neighbor"]),
                                        _p(t["row_offsets"]), _p(cols), _p(bsr), _p(out),
                                        _p(rows), _p(offs))
    assert rc == _abi.NN_E_INVALID_ARGUMENT


@pytest.mark.gpu
@pytest.mark.skipif(not refharness.available(), reason="reference build (oracle/_ref) not present")
def test_gpu_transactional_update_commits_all_or_none():
    """All cells fine: the committed vectors equal the in-place update bit for bit and
    the report counts every cell and the StableReducedSet updates.  One cell whose
    phase-state flag is invalid (the update's own decoding rejects it, status
    THERMO_ERROR, no device fault): nothing is committed, primary and phase state
    stay what they were, and the report names the lowest failing cell."""
    import torch
    from nonenone_amd.flash import ThreePhaseFlash

    fluid = fluids.h2o_co2_nc10("PR")
    ref = refharness.Reference(fluid, FlashOptions.defaults())
    w, hist_p, hist_ps, cur = assemblycmp.reference_grid_states(fluid, 30, 20, 21, ref,
                                                                has_well=True)
    ctx = ThreePhaseFlash(fluid)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a)).cuda()  # noqa: E731
    p_in, s_in = dev(cur), dev(hist_ps)
    status_a, flags_a = ctx.update_phase_state(w["T"], p_in, s_in, True)   # in place
    p_t, s_t = dev(cur), dev(hist_ps)
    status_b, flags_b, rep = ctx.update_state_with_report(w["T"], p_t, s_t, True)
    print("report", rep)
    assert torch.equal(status_a, status_b) and torch.equal(flags_a, flags_b)
    assert int((status_a != 0).sum()) == 0
    assert rep["succeeded"] == 1 and rep["first_failure_cell"] == -1
    assert rep["updated_cells"] == len(cur)
    assert rep["stable_reduced_cells"] == int(((flags_a.cpu().numpy() >> 5) & 1).sum())
    assert torch.equal(p_t, p_in) and torch.equal(s_t, s_in)
    # two invalid cells: the lower index is the first failure, nothing committed
    bad = hist_ps.copy()
    bad[[417, 233], 0] = 0.0
    p_f, s_f = dev(cur), dev(bad)
    status_c, flags_c, rep = ctx.update_state_with_report(w["T"], p_f, s_f, True)
    print("report", rep)
    assert rep["succeeded"] == 0 and rep["first_failure_cell"] == 233
    assert rep["first_failure_status"] == _abi.NN_STATUS_THERMO_ERROR
    assert rep["updated_cells"] == 233
    good = flags_c.cpu().numpy()[:233]
    assert rep["stable_reduced_cells"] == int(((good >> 5) & 1).sum())
    assert torch.equal(p_f, dev(cur)) and torch.equal(s_f, dev(bad))
    sc = status_c.cpu().numpy()
    assert sc[233] == _abi.NN_STATUS_THERMO_ERROR and sc[417] == _abi.NN_STATUS_THERMO_ERROR
    assert np.count_nonzero(sc) == 2
