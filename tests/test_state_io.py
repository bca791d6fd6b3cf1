"""Cell-state samples and the reference's sample CSV text
(output/include/output/core/format.hpp:19-42, writer.hpp:58-93,
output/include/output/petsc/cpgrid_output_access.hpp:42-140).

Pinned by the reference's own output test (test/src/unit/output_test.cpp:
88-143: the exact text of solution_samples.csv, phase_state_samples.csv and a
mass sample file) and, where the reference build exists, by its
writeIndexedCsvHeader / writeCsvValues on random and special values; the
device gather is checked against numpy indexing on the GPU."""
from __future__ import annotations

import numpy as np
import pytest

import refharness
from nonenone_amd import _abi, state_io

pytestmark = pytest.mark.skipif(not _abi.library_path().exists(), reason="library not built")


def test_sample_files_match_the_reference_output_test(tmp_path):
    # output_test.cpp:88-92 writes these rows; :130-143 and :165-170 expect this text
    w = state_io.SampleWriter(tmp_path)
    w.write_solution_samples(np.array([1.25, -2.5]))
    w.write_phase_state_samples(np.array([0.125]))
    assert (tmp_path / "solution_samples.csv").read_text() == "value_0,value_1\n1.25,-2.5\n"
    assert (tmp_path / "phase_state_samples.csv").read_text() == "value_0\n0.125\n"
    assert state_io.row_text([1.0e-6, 2.0e-6]) == "1e-06,2e-06\n"
    assert state_io.row_text([-2.0, -200.0, -6.0]) == "-2,-200,-6\n"


def test_rows_read_back_and_malformed_rows_are_refused(tmp_path):
    rng = np.random.default_rng(4)
    rows = rng.standard_normal((5, 17)) * 10.0 ** rng.integers(-12, 12, (5, 17))
    w = state_io.SampleWriter(tmp_path)
    for r in rows:
        w.write_solution_samples(r)
    back = state_io.read_samples(tmp_path / "solution_samples.csv")
    assert back.shape == rows.shape
    # 16 significant digits: each value within half a unit in the 16th digit
    assert np.all(np.abs(back - rows) <= 5e-16 * np.abs(rows) * 10)
    # what was read prints to the same text again
    text = (tmp_path / "solution_samples.csv").read_text().splitlines(keepends=True)[1:]
    assert [state_io.row_text(b) for b in back] == text
    with pytest.raises(RuntimeError):
        state_io.parse_row("1.0,abc\n")
    with pytest.raises(RuntimeError):
        state_io.parse_row("1.0;2.0\n")


@pytest.mark.skipif(not refharness.available(), reason="reference build (oracle/_ref) not present")
def test_csv_text_matches_the_reference_writer():
    rng = np.random.default_rng(9)
    cases = [rng.standard_normal(11) * 10.0 ** rng.integers(-300, 300, 11),
             np.array([0.0, -0.0, 1.0, -1.0, 0.1, 1.0 / 3.0, 2.0 ** -1074, 1.7976931348623157e308,
                       np.inf, -np.inf, 123456789012345678.0, 1e-5, 1e16, 1e15]),
             rng.uniform(0.0, 1.0, 40), np.array([7.0])]
    for values in cases:
        mine = state_io.header_text(len(values)) + state_io.row_text(values)
        assert mine == refharness.csv_samples(values)


@pytest.mark.gpu
def test_gpu_state_sample_gathers_cell_major():
    import torch

    rng = np.random.default_rng(2)
    ncell, dof = 5000, 17
    host = rng.standard_normal((ncell, dof))
    vec = torch.as_tensor(host).cuda()
    cells = rng.integers(0, ncell, 333)
    variables = np.array([0, 16, 3, 3, 9], np.int32)
    got = state_io.sample(vec, cells, variables).cpu().numpy()
    assert np.array_equal(got, host[cells][:, variables].reshape(-1))
    with pytest.raises(IndexError):
        state_io.sample(vec, [0, ncell], [1])
    with pytest.raises(IndexError):
        state_io.sample(vec, [0], [dof])
    assert state_io.sample(vec, [], [1]).numel() == 0
