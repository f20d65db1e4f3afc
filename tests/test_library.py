"""The C-ABI library loads and exports every entry point the header declares."""

import re
from pathlib import Path

from torchls_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "torchls_hip.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|void)\s+(tlh_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "tlh_normal_matvec" in names and "tlh_pack_blocks" in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
