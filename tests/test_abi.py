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


def _bed_arrays(**kw):
    a = _lib.BedArrays()
    a.n, a.nmembers, a.nblocks, a.ncells, a.cand_slots, a.cell_cap = 1000, 1, 4, 100, 16, 8
    for f in ("x", "v", "x_alt", "v_alt", "radius", "xi_wall", "accounting", "members",
              "workspace"):
        setattr(a, f, 256)          # never dereferenced: validation fails first
    a.workspace_bytes = 1
    for k, v in kw.items():
        setattr(a, k, v)
    return a


@pytest.mark.parametrize("field,value,msg", [
    ("x", 0, b"null array"),
    ("n", 0, b"empty bed batch"),
    ("cand_slots", 17, b"cand_slots"),
    ("cand_slots", 0, b"cand_slots"),
    ("cell_cap", 0, b"cell_cap"),
])
def test_bed_entry_points_reject_bad_descriptors(field, value, msg):
    lib = _lib.load()
    a = _bed_arrays(**{field: value})
    for call in (lambda: lib.tdm_bed_step(ctypes.byref(a), 1, 0, None),
                 lambda: lib.tdm_bed_setup(ctypes.byref(a), None),
                 lambda: lib.tdm_bed_energy(ctypes.byref(a), None)):
        assert call() == _lib.TDM_EINVAL
        buf = ctypes.create_string_buffer(256)
        lib.tdm_last_error(buf, 256)
        assert msg in buf.value


def test_bed_workspace_too_small_is_reported():
    lib = _lib.load()
    a = _bed_arrays()          # 1-byte workspace for 1000 disks
    assert lib.tdm_bed_step(ctypes.byref(a), 1, 0, None) == _lib.TDM_ESPACE
    assert lib.tdm_bed_step(None, 1, 0, None) == _lib.TDM_EINVAL


def test_disk_host_only_entry_points():
    from tensordem_amd.granular2d import DiskArrays, DiskParams
    lib = _lib.load()
    pb, ab = ctypes.c_size_t(), ctypes.c_size_t()
    assert lib.tdm_disk_abi_sizes(ctypes.byref(pb), ctypes.byref(ab)) == 0
    assert pb.value == ctypes.sizeof(DiskParams)
    assert ab.value == ctypes.sizeof(DiskArrays)
    assert lib.tdm_disk_blocks(257) == 2
    small = lib.tdm_disk_workspace_bytes(1000, 400, 16, 24, 4, 0)
    assert small > 1000 * 24 * 2 * (4 + 16 + 8)
    assert lib.tdm_disk_workspace_bytes(1000, 400, 16, 24, 4, 20) > small
    with pytest.raises(RuntimeError):
        _lib.check(lib.tdm_disk_step(None, None, 1, 0, None))


def test_fragment_entry_point_checks_arguments_on_the_host():
    lib = _lib.load()
    n = 1024 * 128
    need = lib.tdm_fragments_workspace_bytes(n)
    assert need >= 7 * 4 * n + 4 * 256 * (n // 256)
    assert lib.tdm_fragments_workspace_bytes(0) == 0
    null = [None] * 3
    assert lib.tdm_plate_fragments(*null, 4, 4, *([None] * 5), None, 0, None) == _lib.TDM_EINVAL
    assert lib.tdm_plate_fragments(*null, 1 << 13, 1 << 12, *([None] * 5), None, 0,
                                   None) == _lib.TDM_EINVAL
