"""ctypes binding of libprt.so (the C-ABI declared in include/prt.h).

The library is built in-tree by ``parlllelrtree_amd.build``; loading fails
loudly when it is missing -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_size_t
from ctypes import c_uint32, c_uint64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libprt.so"

PRT_MAX_LEVELS = 40
PRT_MAX_FANOUT = 64
PRT_MAX_LAYERS = 4
PRT_TUNE_BUCKETS = 6

ERRORS = {0: "OK", -1: "PRT_E_INVAL", -2: "PRT_E_NOMEM", -3: "PRT_E_HIP",
          -4: "PRT_E_OVERFLOW", -5: "PRT_E_UNSUPPORTED", -6: "PRT_E_DEVICE"}
PRT_E_DEVICE = -6
PRT_STATUS_LOOKBACK = 1
PRT_STATUS_INEXACT = 2


class PrtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class BUTreeStruct(ctypes.Structure):
    _fields_ = [
        ("fanout", c_int32), ("num_levels", c_int32), ("num_rects", c_int64),
        ("num_nodes", c_int64), ("level_start", c_int64 * PRT_MAX_LEVELS),
        ("level_count", c_int64 * PRT_MAX_LEVELS), ("keys", c_void_p), ("perm", c_void_p),
        ("node_count", c_void_p), ("node_start", c_void_p), ("node_end", c_void_p),
        ("entry_mbr", c_void_p), ("entry_ref", c_void_p),
    ]


class TDTreeStruct(ctypes.Structure):
    _fields_ = [
        ("fanout", c_int32), ("num_levels", c_int32), ("num_rects", c_int64),
        ("num_slots", c_int64), ("level_offset", c_int64 * (PRT_MAX_LEVELS + 1)),
        ("slot_child", c_void_p), ("slot_mbr", c_void_p), ("perm", c_void_p),
    ]


class TuneStruct(ctypes.Structure):
    _fields_ = [
        ("bucket_max_verts", c_int32 * PRT_TUNE_BUCKETS),
        ("lanes_per_pair", c_int32 * PRT_TUNE_BUCKETS),
        ("waves_per_block", c_int32 * PRT_TUNE_BUCKETS),
        ("list_factor", c_int32 * PRT_TUNE_BUCKETS),
    ]


class PipelineIO(ctypes.Structure):
    _fields_ = [
        ("xyA", c_void_p), ("offA", c_void_p), ("nA", c_int64),
        ("xyB", c_void_p), ("offB", c_void_p), ("nB", c_int64),
        ("max_verts", c_int32), ("fanout", c_int32), ("tree", c_int32), ("reserved", c_int32),
        ("cap", c_int64), ("pairs", c_void_p), ("area", c_void_p), ("count", c_void_p),
        ("total", c_void_p), ("nonempty", c_void_p),
    ]


# name -> (restype, argtypes); every symbol of include/prt.h plus the host
# generators of csrc/gen.cpp
SIGNATURES = {
    "prt_version": (c_char_p, []),
    "prt_last_error": (c_char_p, []),
    "prt_device_status": (c_int, [c_void_p, c_int, c_void_p]),
    "prt_debug_lookback": (c_int, [c_int32, c_int64]),
    "prt_debug_contours": (c_int, [c_int32, POINTER(TuneStruct)]),
    "prt_profile_enable": (None, [c_int]),
    "prt_profile_read": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int]),
    "prt_morton_center30": (c_int, [c_void_p, c_int64, c_void_p, c_void_p]),
    "prt_morton_ref4d": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p]),
    "prt_sort_workspace_size": (c_size_t, [c_int64]),
    "prt_sort_pairs_u32": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_void_p, c_void_p,
                                   c_size_t, c_void_p]),
    "prt_bu_plan": (c_int, [c_int64, c_int, POINTER(BUTreeStruct)]),
    "prt_bu_workspace_size": (c_size_t, [c_int64, c_int]),
    "prt_bu_build": (c_int, [c_void_p, c_int64, c_int, POINTER(BUTreeStruct), c_void_p,
                             c_size_t, c_void_p]),
    "prt_bu_ref4d_workspace_size": (c_size_t, [c_int64, c_int]),
    "prt_bu_build_ref4d": (c_int, [c_void_p, c_int64, c_int, c_int64, POINTER(BUTreeStruct),
                                   c_void_p, c_void_p, c_size_t, c_void_p]),
    "prt_td_plan": (c_int, [c_int64, c_int, POINTER(TDTreeStruct)]),
    "prt_td_workspace_size": (c_size_t, [c_int64, c_int]),
    "prt_td_build": (c_int, [c_void_p, c_int64, c_int, POINTER(TDTreeStruct), c_void_p,
                             c_size_t, c_void_p]),
    "prt_td_layers_workspace_size": (c_size_t, [c_int, c_int64, c_int]),
    "prt_td_build_layers": (c_int, [c_int, POINTER(c_void_p), c_int64, c_int,
                                    POINTER(TDTreeStruct), c_void_p, c_size_t, c_void_p]),
    "prt_join_workspace_size": (c_size_t, [c_int64]),
    "prt_join_bu": (c_int, [POINTER(BUTreeStruct), c_void_p, c_int64, c_void_p, c_int64,
                            c_void_p, c_void_p, c_size_t, c_void_p]),
    "prt_join_td": (c_int, [POINTER(TDTreeStruct), c_void_p, c_int64, c_void_p, c_int64,
                            c_void_p, c_void_p, c_size_t, c_void_p]),
    "prt_join_bu_noalloc": (c_int, [POINTER(BUTreeStruct), c_void_p, c_int64, c_void_p, c_int64,
                                    c_void_p, c_void_p]),
    "prt_search_stats_bu": (c_int, [POINTER(BUTreeStruct), c_void_p, c_int64, c_void_p,
                                    c_void_p]),
    "prt_search_stats_td": (c_int, [POINTER(TDTreeStruct), c_void_p, c_int64, c_void_p,
                                    c_void_p]),
    "prt_sibling_overlap_bu": (c_int, [POINTER(BUTreeStruct), c_void_p, c_void_p]),
    "prt_sibling_overlap_td": (c_int, [POINTER(TDTreeStruct), c_void_p, c_void_p]),
    "prt_overlay_plain_workspace_size": (c_size_t, [c_int64, c_int32]),
    "prt_overlay_plain": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_int32, c_void_p, c_void_p, c_size_t, c_void_p]),
    "prt_overlay_plain_op": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_int32, c_int32, c_void_p, c_void_p, c_size_t, c_void_p]),
    "prt_overlay_batched_op": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_int32, c_int32, POINTER(TuneStruct), c_void_p, c_void_p,
                                       c_size_t, c_void_p]),
    "prt_tune_default": (c_int, [POINTER(TuneStruct)]),
    "prt_overlay_batched_workspace_size": (c_size_t, [c_int64, c_int32]),
    "prt_overlay_batched": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_int32, POINTER(TuneStruct), c_void_p, c_void_p, c_size_t,
                                    c_void_p]),
    "prt_overlay_batched_overflow": (c_int, [c_void_p, c_size_t, c_int64, c_int32,
                                             POINTER(c_int64), c_void_p]),
    "prt_device_lds_limit": (c_int, []),
    "prt_overlay_area_grad": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "prt_overlay_area_grad_f32": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                          c_void_p, c_int32, c_void_p, c_void_p, c_void_p,
                                          c_void_p, c_void_p]),
    "prt_overlay_team_lds": (c_int64, [c_int32, c_int32]),
    "prt_overlay_batched_dev": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                        c_void_p, c_void_p, c_int32, POINTER(TuneStruct),
                                        c_void_p, c_void_p, c_size_t, c_void_p]),
    "prt_overlay_total_workspace_size": (c_size_t, []),
    "prt_overlay_total": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                  c_size_t, c_void_p]),
    "prt_overlay_contours_workspace_size": (c_size_t, [c_int64, c_int32, c_int64]),
    "prt_overlay_contours_stage_size": (c_size_t, [c_int64, c_int64]),
    "prt_overlay_contours_count": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                           c_void_p, c_int32, c_int32, c_void_p, c_void_p,
                                           c_void_p, c_size_t, c_void_p, c_size_t, c_void_p]),
    "prt_overlay_contours_write": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                           c_void_p, c_int32, c_int32, c_void_p, c_void_p,
                                           c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p, c_size_t, c_void_p, c_size_t, c_void_p]),
    "prt_poly_mbr_dev": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    "prt_pipeline_workspace_size": (c_size_t, [POINTER(PipelineIO)]),
    "prt_pipeline_run": (c_int, [POINTER(PipelineIO), POINTER(TuneStruct), c_void_p, c_size_t,
                                 c_void_p]),
    "prt_pipeline_capture": (c_int, [POINTER(PipelineIO), POINTER(TuneStruct), c_void_p,
                                     c_size_t, c_void_p, POINTER(c_void_p)]),
    "prt_pipeline_launch": (c_int, [c_void_p, c_void_p]),
    "prt_pipeline_destroy": (c_int, [c_void_p]),
    # 2 file names, 20 output pointers (rect1..4, num_contours, rect*_query,
    # num_querys, num/prefix/vertex of both layers, min_x, min_y, holes)
    "prt_read_data": (c_int, [c_char_p, c_char_p] + [c_void_p] * 20),
    "prt_run": (c_int, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                        c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                        POINTER(c_int64)]),
    # host-only reference text formats (csrc/io.cpp)
    "prt_read_rects_text": (c_int, [c_char_p, POINTER(c_int64), c_void_p, c_void_p]),
    "prt_write_rects_text": (c_int, [c_char_p, c_int64, c_void_p, c_void_p]),
    "prt_rects_normalize": (c_int, [c_void_p, c_int64, POINTER(c_int64), POINTER(c_int64),
                                    POINTER(c_int64)]),
    "prt_read_polys_text": (c_int, [c_char_p, POINTER(c_int64), POINTER(c_int64), c_void_p,
                                    c_void_p, c_void_p, c_void_p, c_void_p]),
    "prt_write_polys_text": (c_int, [c_char_p, c_int64, c_void_p, c_void_p, c_void_p]),
    # host-only generators (csrc/gen.cpp)
    "prt_gen_u01": (c_double, [c_uint64, c_uint64]),
    "prt_gen_rects_uniform": (None, [c_uint64, c_int64, c_double, c_double, c_void_p]),
    "prt_gen_windows": (None, [c_uint64, c_int64, c_double, c_void_p]),
    "prt_gen_rects_gmm": (None, [c_uint64, c_int64, c_int, c_double, c_double, c_double,
                                 c_void_p]),
    "prt_gen_stars_in_rects": (None, [c_uint64, c_void_p, c_void_p, c_int64, c_int, c_int,
                                      c_void_p, c_void_p]),
    "prt_gen_stars_grid": (None, [c_uint64, c_int64, c_int, c_double, c_double, c_double,
                                  c_double, c_int, c_int, c_void_p, c_void_p]),
    "prt_gen_stars_gmm": (None, [c_uint64, c_int64, c_int, c_double, c_double, c_double,
                                 c_double, c_double, c_int, c_int, c_void_p, c_void_p]),
    "prt_gen_star_pairs": (None, [c_uint64, c_int64, c_int, c_int, c_double, c_void_p,
                                  c_void_p]),
    "prt_poly_mbr": (None, [c_void_p, c_void_p, c_int64, c_void_p]),
}

_LIB: ctypes.CDLL | None = None


def lib() -> ctypes.CDLL:
    """Load libprt.so; raises when it has not been built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m parlllelrtree_amd.build`")
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def check(rc: int) -> None:
    if rc != 0:
        raise PrtError(rc, lib().prt_last_error().decode())
