"""Serving batch merge plan (host code of nonenone_amd/csrc/serve_plan.hpp): steps
of one fluid merge only when their primary / phase-state layouts and well flags
agree, so a merged phase update never copies one member's layout into another's
buffers.  Runs on the CPU."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


def test_merge_plan_separates_layouts(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "serve_plan_check"
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", str(HERE / "serve_plan_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"
