"""Cell-state samples and their CSV files, as the reference's output module writes
them (output/include/output/core/writer.hpp:58-93, format.hpp:19-42,
output/include/output/petsc/cpgrid_output_access.hpp:42-140).

A sample is the values of chosen variables of chosen cells of a primary or
phase-state array, cell-major: value (i, k) is variable vars[k] of cell cells[i].
`sample` gathers it on the device (nn_state_sample); `SampleWriter` appends it to
solution_samples.csv / phase_state_samples.csv with the reference's header and
number format (nn_csv_sample_header / nn_csv_sample_row); `read_samples` reads
such a file back (nn_csv_parse_row).
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from . import _abi


def sample(vec, cells, variables):
    """Device gather of a sample from an (ncell, dof) torch array; raises on an
    index outside the array (the reference throws std::out_of_range)."""
    import torch

    from .flash import _ptr, _stream

    ncell, dof = vec.shape
    dev = vec.device
    c = torch.as_tensor(np.asarray(cells, np.int64), device=dev)
    v = torch.as_tensor(np.asarray(variables, np.int32), device=dev)
    out = torch.empty(len(c) * len(v), dtype=torch.float64, device=dev)
    bad = torch.empty(1, dtype=torch.int32, device=dev)
    L = _abi.lib()
    _abi.check(L.nn_state_sample(_ptr(vec), ncell, dof, _ptr(c), len(c), _ptr(v), len(v),
                                 _ptr(out), _ptr(bad), _stream()), "nn_state_sample")
    if int(bad.item()) != 0:
        raise IndexError("requested output cell or variable outside the state array")
    return out


def header_text(count: int) -> str:
    L = _abi.lib()
    n = ctypes.c_int64()
    _abi.check(L.nn_csv_sample_header(count, None, 0, ctypes.byref(n)), "nn_csv_sample_header")
    buf = ctypes.create_string_buffer(n.value + 1)
    _abi.check(L.nn_csv_sample_header(count, buf, n.value + 1, ctypes.byref(n)),
               "nn_csv_sample_header")
    return buf.value.decode()


def row_text(values) -> str:
    values = np.ascontiguousarray(values, np.float64)
    L = _abi.lib()
    n = ctypes.c_int64()
    p = ctypes.c_void_p(values.ctypes.data)
    _abi.check(L.nn_csv_sample_row(p, len(values), None, 0, ctypes.byref(n)), "nn_csv_sample_row")
    buf = ctypes.create_string_buffer(n.value + 1)
    _abi.check(L.nn_csv_sample_row(p, len(values), buf, n.value + 1, ctypes.byref(n)),
               "nn_csv_sample_row")
    return buf.value.decode()


def parse_row(line: str) -> np.ndarray:
    L = _abi.lib()
    raw = line.encode()
    cap = raw.count(b",") + 1
    out = np.zeros(cap)
    n = ctypes.c_int64()
    _abi.check(L.nn_csv_parse_row(raw, ctypes.c_void_p(out.ctypes.data), cap, ctypes.byref(n)),
               "nn_csv_parse_row")
    return out[:n.value]


class SampleWriter:
    """OutputWriter's solution / phase-state sample files in `directory`: the
    header is written with the first row, one row per sampling time."""

    FILES = {"solution": "solution_samples.csv", "phase_state": "phase_state_samples.csv"}

    def __init__(self, directory):
        self.directory = Path(directory)
        self.directory.mkdir(parents=True, exist_ok=True)
        self._header = {}
        for kind, name in self.FILES.items():
            (self.directory / name).write_text("")
            self._header[kind] = False

    def _write(self, kind, values):
        if hasattr(values, "detach"):
            values = values.detach().cpu().numpy()
        with open(self.directory / self.FILES[kind], "a") as f:
            if not self._header[kind]:
                f.write(header_text(len(values)))
                self._header[kind] = True
            f.write(row_text(values))

    def write_solution_samples(self, values):
        self._write("solution", values)

    def write_phase_state_samples(self, values):
        self._write("phase_state", values)


def read_samples(path) -> np.ndarray:
    """Rows of a sample file (header checked) as a (rows, count) array."""
    lines = Path(path).read_text().splitlines(keepends=True)
    if not lines:
        return np.zeros((0, 0))
    count = lines[0].count(",") + 1
    if lines[0] != header_text(count):
        raise ValueError(f"{path}: not a sample file header")
    rows = [parse_row(line) for line in lines[1:]]
    if any(len(r) != count for r in rows):
        raise ValueError(f"{path}: row length differs from the header")
    return np.array(rows).reshape(len(rows), count)
