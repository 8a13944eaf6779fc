"""ctypes binding of the C ABI in include/tensordem_amd.h.

The shared library `libtensordem_amd.so` is built in-tree by
`tensordem_amd.build.build()` (hipcc, gfx950).  Loading fails with an
ImportError when it is missing: there is no CPU or PyTorch fallback.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libtensordem_amd.so"

TDM_OK, TDM_EINVAL, TDM_EHIP, TDM_ESPACE = 0, -1, -2, -3
ST_NONFINITE, ST_STALE_LIST, ST_CAND_OVERFLOW, ST_CELL_OVERFLOW = 1, 2, 4, 8

NACC, NLAST = 13, 10
ACC_EXT_WORK, ACC_RELEASE, ACC_ACQUIRED, ACC_TOOL_WORK, ACC_DAMPING = range(5)
ACC_EXT_IMP, ACC_SUP_IMP, ACC_DRAG_IMP, ACC_TOOL_IMP = 5, 7, 9, 11

c_double_p = ctypes.c_void_p


class Member(ctypes.Structure):
    """Mirror of `tdm_plate_member`."""
    _fields_ = [
        ("radius", ctypes.c_double), ("bond_stiffness", ctypes.c_double),
        ("contact_stiffness", ctypes.c_double), ("breaking_strain", ctypes.c_double),
        ("shear_breaking_strain", ctypes.c_double), ("contact_damping", ctypes.c_double),
        ("drag", ctypes.c_double), ("tool_radius", ctypes.c_double), ("dt", ctypes.c_double),
        ("mass", ctypes.c_double), ("tool_start", ctypes.c_double * 2),
        ("tool_velocity", ctypes.c_double * 2), ("time", ctypes.c_double),
        ("skin", ctypes.c_double), ("cell_size", ctypes.c_double),
        ("grid_origin", ctypes.c_double * 2), ("fp32_margin", ctypes.c_double),
        ("step_count", ctypes.c_int64),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("first", ctypes.c_int32),
        ("bond_first", ctypes.c_int32), ("nbonds", ctypes.c_int32),
        ("hull_first", ctypes.c_int32), ("hull_m", ctypes.c_int32), ("ncx", ctypes.c_int32),
        ("ncy", ctypes.c_int32), ("cell_first", ctypes.c_int32), ("tile_x", ctypes.c_int32),
        ("block_first", ctypes.c_int32), ("nblocks", ctypes.c_int32), ("mode", ctypes.c_int32),
        ("tab", ctypes.c_int32), ("status", ctypes.c_int32), ("has_loads", ctypes.c_int32),
        ("per_disk_strength", ctypes.c_int32), ("pad_", ctypes.c_int32),
    ]


class Arrays(ctypes.Structure):
    """Mirror of `tdm_plate_arrays`."""
    _fields_ = [
        ("x", ctypes.c_void_p), ("v", ctypes.c_void_p), ("x_alt", ctypes.c_void_p),
        ("v_alt", ctypes.c_void_p), ("fixed", ctypes.c_void_p), ("links", ctypes.c_void_p),
        ("fwd_bond", ctypes.c_void_p), ("alive", ctypes.c_void_p),
        ("failure_mode", ctypes.c_void_p), ("failure_time_s", ctypes.c_void_p),
        ("failure_extension_m", ctypes.c_void_p), ("failure_energy_J", ctypes.c_void_p),
        ("loads", ctypes.c_void_p), ("hull", ctypes.c_void_p), ("eps_t_disk", ctypes.c_void_p),
        ("eps_s_disk", ctypes.c_void_p), ("accounting", ctypes.c_void_p),
        ("last", ctypes.c_void_p), ("force", ctypes.c_void_p), ("reaction", ctypes.c_void_p),
        ("energy", ctypes.c_void_p), ("members", ctypes.c_void_p), ("workspace", ctypes.c_void_p),
        ("workspace_bytes", ctypes.c_size_t),
        ("n", ctypes.c_int32), ("nbonds", ctypes.c_int32), ("nmembers", ctypes.c_int32),
        ("nblocks", ctypes.c_int32), ("ncells", ctypes.c_int32), ("cand_slots", ctypes.c_int32),
        ("cell_cap", ctypes.c_int32), ("max_tile_x", ctypes.c_int32),
    ]


class BedMember(ctypes.Structure):
    """Mirror of `tdm_bed_member`."""
    _fields_ = [
        ("e_star", ctypes.c_double), ("g_star", ctypes.c_double), ("damping", ctypes.c_double),
        ("friction", ctypes.c_double), ("wall_friction", ctypes.c_double),
        ("rho_pi_t", ctypes.c_double), ("gravity_y", ctypes.c_double), ("dt", ctypes.c_double),
        ("floor_y", ctypes.c_double), ("back_wall_x", ctypes.c_double),
        ("right_wall_x", ctypes.c_double), ("right_wall_v", ctypes.c_double),
        ("time", ctypes.c_double),
        ("skin", ctypes.c_double), ("cell_size", ctypes.c_double),
        ("grid_origin", ctypes.c_double * 2), ("step_count", ctypes.c_int64),
        ("first", ctypes.c_int32), ("n", ctypes.c_int32), ("use_floor", ctypes.c_int32),
        ("use_back_wall", ctypes.c_int32), ("ncx", ctypes.c_int32), ("ncy", ctypes.c_int32),
        ("cell_first", ctypes.c_int32), ("block_first", ctypes.c_int32),
        ("nblocks", ctypes.c_int32), ("mode", ctypes.c_int32), ("tab", ctypes.c_int32),
        ("list_parity", ctypes.c_int32), ("status", ctypes.c_int32), ("rebuilds", ctypes.c_int32),
        ("use_right_wall", ctypes.c_int32), ("integrator", ctypes.c_int32),
        ("verlet_start", ctypes.c_int32), ("pad_", ctypes.c_int32),
    ]


class BedArrays(ctypes.Structure):
    """Mirror of `tdm_bed_arrays`."""
    _fields_ = [
        ("x", ctypes.c_void_p), ("v", ctypes.c_void_p), ("x_alt", ctypes.c_void_p),
        ("v_alt", ctypes.c_void_p), ("radius", ctypes.c_void_p), ("xi_wall", ctypes.c_void_p),
        ("accounting", ctypes.c_void_p), ("energy", ctypes.c_void_p), ("members", ctypes.c_void_p),
        ("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_size_t),
        ("n", ctypes.c_int32), ("nmembers", ctypes.c_int32), ("nblocks", ctypes.c_int32),
        ("ncells", ctypes.c_int32), ("cand_slots", ctypes.c_int32), ("cell_cap", ctypes.c_int32),
        ("fp32_histories", ctypes.c_int32), ("verlet", ctypes.c_int32),
    ]


class D3Params(ctypes.Structure):
    """Mirror of `tdm_d3_params`."""
    _fields_ = [(name, ctypes.c_double) for name in (
        "radius", "mass", "dt", "drag", "kb", "kc", "kt", "mu", "eps_t", "eps_s", "hz_pair",
        "hz_pair_u", "hz_rigid", "hz_kt", "zeta2", "gamma_mobile", "gamma_pinned", "cdamp",
        "tool_r", "platen_k", "platen_c", "top_z0", "bot_z0", "top_v", "bot_v", "lat_k", "lat_c",
        "mg", "cell_w", "skin")] + [(name, ctypes.c_int32) for name in (
        "hertz", "damp", "friction", "mindlin_kt", "top", "bot", "lateral", "gravity", "n",
        "nbonds", "ncells", "cell_cap", "kh", "nblocks")]


class D3Arrays(ctypes.Structure):
    """Mirror of `tdm_d3_arrays`."""
    _fields_ = [(name, ctypes.c_void_p) for name in (
        "x", "v", "x_alt", "v_alt", "x0", "fixed", "nbr", "bond_of_slot", "pairs", "alive", "thr",
        "failure_mode", "failure_time_s", "failure_extension_m", "failure_energy_J", "loads",
        "force", "pair_energy", "tool", "clock", "out", "workspace")] + [
        ("workspace_bytes", ctypes.c_size_t), ("evals", ctypes.c_int64),
        ("hist_flips", ctypes.c_int64), ("steps", ctypes.c_int64)]


D3_REACTION, D3_TOP_REACTION, D3_BOTTOM_REACTION, D3_LATERAL = 0, 3, 4, 5
D3_FRICTION_POWER, D3_CONTACTS, D3_BREAKS, D3_STATUS, D3_REBUILDS, D3_NOUT = 9, 10, 11, 12, 13, 14
D3_UPDATE_FRACTURE, D3_UPDATE_HISTORY = 1, 2
ST_HIST_OVERFLOW = 16
ST_INTEGRATOR = 32


BED_DASHPOT, BED_FRICTION, BED_WALL_DASHPOT, BED_WALL_FRICTION, BED_WALL_IMPULSE = 0, 1, 2, 3, 4
BED_CONTACTS, BED_WALL_CONTACTS, BED_WALL_WORK, BED_NACC = 6, 7, 8, 9

# every entry point declared in include/tensordem_amd.h, with its ctypes signature
_SIGNATURES = {
    "tdm_version": ([], ctypes.c_int),
    "tdm_last_error": ([ctypes.c_char_p, ctypes.c_size_t], ctypes.c_int),
    "tdm_abi_sizes": ([ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)],
                      ctypes.c_int),
    "tdm_plate_tiles": ([ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "tdm_plate_workspace_bytes": ([ctypes.c_int] * 6, ctypes.c_size_t),
    "tdm_plate_setup": ([ctypes.POINTER(Arrays), ctypes.c_void_p], ctypes.c_int),
    "tdm_plate_sync_links": ([ctypes.POINTER(Arrays), ctypes.c_void_p], ctypes.c_int),
    "tdm_plate_rebin": ([ctypes.POINTER(Arrays), ctypes.c_void_p], ctypes.c_int),
    "tdm_plate_step": ([ctypes.POINTER(Arrays), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                        ctypes.c_void_p], ctypes.c_int),
    "tdm_plate_step_timed": ([ctypes.POINTER(Arrays), ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                              ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "tdm_plate_forces": ([ctypes.POINTER(Arrays), ctypes.c_int, ctypes.c_int, ctypes.c_void_p],
                         ctypes.c_int),
    "tdm_plate_energy": ([ctypes.POINTER(Arrays), ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "tdm_plate_graph_create": ([ctypes.POINTER(Arrays), ctypes.c_int, ctypes.c_int,
                                ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "tdm_graph_launch": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "tdm_fragments_workspace_bytes": ([ctypes.c_int], ctypes.c_size_t),
    "tdm_plate_fragments": ([ctypes.c_void_p] * 3 + [ctypes.c_int] * 2 + [ctypes.c_void_p] * 6
                            + [ctypes.c_size_t, ctypes.c_void_p], ctypes.c_int),
    "tdm_plate_fragments_noalloc": ([ctypes.c_void_p] * 3 + [ctypes.c_int] * 2
                                    + [ctypes.c_void_p] * 6, ctypes.c_int),
    "tdm_bed_abi_sizes": ([ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)],
                          ctypes.c_int),
    "tdm_bed_blocks": ([ctypes.c_int], ctypes.c_int),
    "tdm_bed_workspace_bytes": ([ctypes.c_int] * 6, ctypes.c_size_t),
    "tdm_bed_workspace_layout": ([ctypes.c_int] * 6 + [ctypes.POINTER(ctypes.c_size_t)] * 5,
                                 ctypes.c_int),
    "tdm_bed_setup": ([ctypes.POINTER(BedArrays), ctypes.c_void_p], ctypes.c_int),
    "tdm_bed_rebin": ([ctypes.POINTER(BedArrays), ctypes.c_void_p], ctypes.c_int),
    "tdm_bed_step": ([ctypes.POINTER(BedArrays), ctypes.c_int, ctypes.c_int, ctypes.c_void_p],
                     ctypes.c_int),
    "tdm_bed_step_timed": ([ctypes.POINTER(BedArrays), ctypes.c_int, ctypes.c_int,
                            ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "tdm_bed_energy": ([ctypes.POINTER(BedArrays), ctypes.c_void_p], ctypes.c_int),
    "tdm_bed_velocity": ([ctypes.POINTER(BedArrays), ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "tdm_bed_graph_create": ([ctypes.POINTER(BedArrays), ctypes.c_int, ctypes.c_int,
                              ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "tdm_graph_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "tdm_disk_abi_sizes": ([ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)],
                           ctypes.c_int),
    "tdm_disk_blocks": ([ctypes.c_int], ctypes.c_int),
    "tdm_disk_workspace_bytes": ([ctypes.c_int] * 6, ctypes.c_size_t),
    "tdm_disk_setup": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "tdm_disk_rebin": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "tdm_disk_segments": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "tdm_disk_step": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                       ctypes.c_void_p], ctypes.c_int),
    "tdm_disk_inspect": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "tdm_disk_step_timed": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
                            ctypes.c_int),
    "tdm_disk_graph_create": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                               ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "tdm_d3_abi_sizes": ([ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)],
                         ctypes.c_int),
    "tdm_d3_blocks": ([ctypes.c_int], ctypes.c_int),
    "tdm_d3_workspace_bytes": ([ctypes.c_int] * 5, ctypes.c_size_t),
    "tdm_d3_setup": ([ctypes.c_void_p] * 3, ctypes.c_int),
    "tdm_d3_clear_status": ([ctypes.c_void_p] * 3, ctypes.c_int),
    "tdm_d3_clear_history": ([ctypes.c_void_p] * 3, ctypes.c_int),
    "tdm_d3_rebin": ([ctypes.c_void_p] * 3, ctypes.c_int),
    "tdm_d3_step": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
                    ctypes.c_int),
    "tdm_d3_forces": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
                      ctypes.c_int),
    "tdm_d3_energy": ([ctypes.c_void_p] * 3, ctypes.c_int),
    "tdm_d3_shear": ([ctypes.c_void_p] * 4, ctypes.c_int),
    "tdm_d3_graph_create": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                             ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "tdm_d3_step_timed": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
                          + [ctypes.POINTER(ctypes.c_double)] * 3, ctypes.c_int),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load the HIP library (once).  Raises ImportError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m tensordem_amd.build` "
            "(hipcc, gfx950). tensordem_amd has no CPU fallback.")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    mb, ab = ctypes.c_size_t(), ctypes.c_size_t()
    lib.tdm_abi_sizes(ctypes.byref(mb), ctypes.byref(ab))
    if mb.value != ctypes.sizeof(Member) or ab.value != ctypes.sizeof(Arrays):
        raise ImportError(f"ABI mismatch: library structs {mb.value}/{ab.value} B, binding "
                          f"{ctypes.sizeof(Member)}/{ctypes.sizeof(Arrays)} B")
    lib.tdm_bed_abi_sizes(ctypes.byref(mb), ctypes.byref(ab))
    if mb.value != ctypes.sizeof(BedMember) or ab.value != ctypes.sizeof(BedArrays):
        raise ImportError(f"ABI mismatch: library bed structs {mb.value}/{ab.value} B, binding "
                          f"{ctypes.sizeof(BedMember)}/{ctypes.sizeof(BedArrays)} B")
    lib.tdm_d3_abi_sizes(ctypes.byref(mb), ctypes.byref(ab))
    if mb.value != ctypes.sizeof(D3Params) or ab.value != ctypes.sizeof(D3Arrays):
        raise ImportError(f"ABI mismatch: library 3-D structs {mb.value}/{ab.value} B, binding "
                          f"{ctypes.sizeof(D3Params)}/{ctypes.sizeof(D3Arrays)} B")
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def check(rc: int) -> None:
    """Raise RuntimeError with the library's message for a nonzero status."""
    if rc == TDM_OK:
        return
    buf = ctypes.create_string_buffer(512)
    load().tdm_last_error(buf, len(buf))
    raise RuntimeError(f"tensordem_amd error {rc}: {buf.value.decode(errors='replace')}")
