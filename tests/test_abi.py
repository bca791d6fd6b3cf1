"""The C-ABI library loads and exports every symbol include/nonenone_amd.h declares."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from nonenone_amd import _abi

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "nonenone_amd.h").read_text()
    return sorted(set(re.findall(r"\b(nn_[a-z_]+)\s*\(", text)))


def test_header_declarations_match_python_mirror():
    assert declared_symbols() == sorted(_abi.EXPORTED_SYMBOLS)


@pytest.mark.skipif(not _abi.library_path().exists(), reason="library not built")
def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(str(_abi.library_path()))
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


@pytest.mark.skipif(not _abi.library_path().exists(), reason="library not built")
def test_default_options_match_reference_defaults():
    L = _abi.lib()
    o = _abi.FlashOptions()
    L.nn_flash_options_default(ctypes.byref(o))
    d = _abi.FlashOptions.defaults()
    for name, _ in _abi.FlashOptions._fields_:
        assert getattr(o, name) == getattr(d, name), name


def test_struct_sizes_are_stable():
    # nn_fluid_desc / nn_flash_options layouts are part of the ABI.
    assert ctypes.sizeof(_abi.FlashOptions) == 9 * 4 + 4 + 8 * 8
    assert ctypes.sizeof(_abi.FlashOut) == 11 * 8
