"""The C-ABI library loads and exports every entry point include/tensordem_amd.h declares."""
import ctypes
import re

import pytest

from conftest import ROOT
from tensordem_amd import _lib


def declared_functions():
    text = (ROOT / "include" / "tensordem_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tdm_\w+)\s*\(", text)))


def test_header_declarations_are_bound():
    assert declared_functions() == sorted(_lib.exported_symbols())


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_struct_mirrors_match_library():
    lib = _lib.load()
    mb, ab = ctypes.c_size_t(), ctypes.c_size_t()
    assert lib.tdm_abi_sizes(ctypes.byref(mb), ctypes.byref(ab)) == 0
    assert mb.value == ctypes.sizeof(_lib.Member)
    assert ab.value == ctypes.sizeof(_lib.Arrays)


def test_host_only_entry_points():
    lib = _lib.load()
    assert lib.tdm_version() >= 100
    assert lib.tdm_plate_tiles(1024, 128, 32) == 32 * 16
    assert lib.tdm_plate_tiles(7, 3, 16) == 1
    assert lib.tdm_plate_tiles(7, 3, 24) == _lib.TDM_EINVAL
    buf = ctypes.create_string_buffer(128)
    assert lib.tdm_last_error(buf, 128) == 0 and b"tile_x" in buf.value
    assert lib.tdm_plate_workspace_bytes(1000, 4, 1, 100, 8, 8) > 0
    with pytest.raises(RuntimeError):
        _lib.check(lib.tdm_plate_step(None, 1, 0, 0, None))


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "absent.so")
    with pytest.raises(ImportError):
        _lib.load()


def test_bed_host_only_entry_points():
    lib = _lib.load()
    assert lib.tdm_bed_blocks(257) == 2
    assert lib.tdm_bed_blocks(0) == _lib.TDM_EINVAL
    args = (1000, 4, 1, 300, 16, 8)
    total = lib.tdm_bed_workspace_bytes(*args)
    offs = [ctypes.c_size_t() for _ in range(5)]
    assert lib.tdm_bed_workspace_layout(*args, *[ctypes.byref(o) for o in offs]) == 0
    cand, xi, mask, count, xb = (o.value for o in offs)
    # cand [2][16][1000] int32, then xi [2][16][1000] double2, mask [2][1000], ncand [2][1000]
    assert 0 < cand and cand % 256 == 0
    assert xi - cand >= 2 * 16 * 1000 * 4
    assert mask - xi >= 2 * 16 * 1000 * 16
    assert count - mask >= 2 * 1000 * 4
    assert xb - count >= 2 * 1000 and xb % 256 == 0
    assert total >= xb + 16 * 1000
