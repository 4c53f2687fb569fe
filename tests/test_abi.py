"""The C ABI library loads and exports every entry point include/nts.h
declares (no compute: runs without a GPU)."""
import ctypes as C
import re
from pathlib import Path

from nanotracer_amd import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "nts.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(nts_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "nts_project" in syms and "nts_bin" in syms and "nts_composite_backward" in syms
    assert len(syms) >= 40


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_table_matches_header():
    assert sorted(_lib.EXPORTED) == declared_symbols()


def test_library_is_gfx950():
    # the embedded offload bundle names its target
    assert b"gfx950" in _lib.LIB_PATH.read_bytes()


def test_camera_struct_is_96_bytes():
    assert C.sizeof(_lib.Camera) == 96
