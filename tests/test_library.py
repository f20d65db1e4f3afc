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


def test_python_constants_match_the_header():
    """The host-side plan mirrors the kernels' tile geometry; a mismatch silently
    mis-assigns primary runs, so every #define the Python side restates must agree."""
    from torchls_amd import plan

    defines = {k: int(v) for k, v in re.findall(r"^#define\s+(TLH_\w+)\s+(-?\d+)\b", HEADER.read_text(), re.M)}
    restated = [name for name in dir(_lib) if name.startswith("TLH_") and name in defines]
    assert {"TLH_WAVE_THREADS", "TLH_WAVE_TILE", "TLH_TILE_THREADS", "TLH_F32", "TLH_F64"} <= set(restated)
    for name in restated:
        assert getattr(_lib, name) == defines[name], name
    assert plan.WAVE_THREADS == defines["TLH_WAVE_THREADS"]
    assert plan.WAVE_TILE == defines["TLH_WAVE_TILE"]
    assert plan.THREADS == defines["TLH_TILE_THREADS"]
