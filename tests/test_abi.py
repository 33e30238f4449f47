"""C-ABI boundary on the CPU: libprt.so loads, exports every symbol that
include/prt.h declares, and its host-only planning functions agree with the
oracle.  No kernel is launched here."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from parlllelrtree_amd import _lib, gen

HEADER = Path(__file__).resolve().parents[1] / "include" / "prt.h"


def declared_symbols():
    text = HEADER.read_text()
    decl = r"^(?:const\s+)?[a-z_0-9]+\s*\*?\s*(prt_[a-z0-9_]+)\s*\("
    return sorted(set(re.findall(decl, text, flags=re.M)))


def test_header_symbols_exported():
    lib = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), f"{name} missing from libprt.so"
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_version_and_error():
    lib = _lib.lib()
    assert lib.prt_version().startswith(b"prt ")
    st = _lib.BUTreeStruct()
    assert lib.prt_bu_plan(0, 16, ctypes.byref(st)) == -1
    assert b"out of range" in lib.prt_last_error()
    assert lib.prt_bu_plan(10, 1, ctypes.byref(st)) == -1
    assert lib.prt_bu_plan(10, 65, ctypes.byref(st)) == -1


@pytest.mark.parametrize("n,f", [(1, 16), (2, 2), (17, 16), (1 << 20, 16), (1 << 24, 16),
                                 (123457, 10), (999, 64)])
def test_plans_match_oracle(n, f):
    lib = _lib.lib()
    st = _lib.BUTreeStruct()
    assert lib.prt_bu_plan(n, f, ctypes.byref(st)) == 0
    lv = O.bu_levels(n, f)
    assert st.num_levels == len(lv)
    assert [(st.level_start[i], st.level_count[i]) for i in range(st.num_levels)] == lv
    td = _lib.TDTreeStruct()
    assert lib.prt_td_plan(n, f, ctypes.byref(td)) == 0
    off = O.td_levels(n, f)
    assert [td.level_offset[i] for i in range(td.num_levels + 1)] == off
    assert lib.prt_bu_workspace_size(n, f) > 0
    assert lib.prt_td_workspace_size(n, f) > 0


def test_td_layers_workspace_and_argument_checks():
    lib = _lib.lib()
    for n, f in ((1, 16), (70001, 16), (1 << 20, 16), (1000, 10), (300000, 2)):
        one = lib.prt_td_workspace_size(n, f)
        assert lib.prt_td_layers_workspace_size(1, n, f) == one
        two = lib.prt_td_layers_workspace_size(2, n, f)
        # power-of-two fanouts sort the layers together; others reuse one
        # layer's workspace layer after layer
        assert two >= 2 * one - 4096 if f & (f - 1) == 0 else two == one
    assert lib.prt_td_layers_workspace_size(0, 100, 16) == 0
    assert lib.prt_td_layers_workspace_size(_lib.PRT_MAX_LAYERS + 1, 100, 16) == 0
    trees = (_lib.TDTreeStruct * 2)()
    ptrs = (ctypes.c_void_p * 2)()
    assert lib.prt_td_build_layers(0, ptrs, 100, 16, trees, None, 0, None) == -1
    assert b"layers not in" in lib.prt_last_error()
    assert lib.prt_td_build_layers(2, ptrs, 1 << 30, 16, trees, None, 0, None) == -1
    assert lib.prt_td_build_layers(2, ptrs, 100, 16, trees, None, 0, None) == -1
    assert b"null device pointer" in lib.prt_last_error()


def test_contour_stage_sizes():
    lib = _lib.lib()
    for npairs in (1, 1000, 100_000):
        sizes = [lib.prt_overlay_contours_stage_size(npairs, k)
                 for k in (8, 9, 1000, 1 << 20, (1 << 31) - 1)]
        assert all(b > a for a, b in zip(sizes, sizes[1:]))
        # 20 bytes per vertex slot, 16 per contour record (one per 3 slots)
        # and 16 per pair at least
        assert sizes[3] >= (1 << 20) * 20 + (1 << 20) // 3 * 16 + npairs * 16
    assert lib.prt_overlay_contours_stage_size(10, 7) == 0
    assert lib.prt_overlay_contours_stage_size(10, 1 << 31) == 0
    assert lib.prt_overlay_contours_stage_size(-1, 8) == 0
    # the stage is optional and checked on the host before any launch
    assert lib.prt_overlay_contours_count(None, 1, None, None, None, None, 8, 1, None, None,
                                          None, 0, None, 0, None) == -1


def test_known_layout_sizes():
    lib = _lib.lib()
    st = _lib.BUTreeStruct()
    lib.prt_bu_plan(1 << 20, 16, ctypes.byref(st))
    assert st.num_nodes == 69905
    td = _lib.TDTreeStruct()
    lib.prt_td_plan(1 << 20, 16, ctypes.byref(td))
    assert td.num_slots == 1118480 and td.num_levels == 5


def test_generators_deterministic_and_in_range():
    a = gen.rects_uniform(gen.seed(0), 1000)
    b = gen.rects_uniform(gen.seed(0), 1000)
    np.testing.assert_array_equal(a, b)
    w = a[:, 2] - a[:, 0]
    assert (w > 0.99e-4).all() and (w < 1.01e-3).all()
    g = gen.stars_grid(gen.seed(2), n=2000)
    k = np.diff(g.off)
    assert k.min() >= 4 and k.max() <= 512
    assert 15 < np.mean(k) < 45
    m = gen.poly_mbr(g)
    for i in range(0, 2000, 97):
        p = g.poly(i)
        assert m[i, 0] <= p[:, 0].min() and m[i, 2] >= p[:, 0].max()
        assert m[i, 1] <= p[:, 1].min() and m[i, 3] >= p[:, 1].max()


def test_generator_thread_independence():
    # 70000 > the single-thread threshold: the split must not change values
    big = gen.rects_uniform(gen.seed(4), 70000)
    small = gen.rects_uniform(gen.seed(4), 1000)
    np.testing.assert_array_equal(big[:1000], small)


def test_tune_shapes_checked_before_any_launch():
    # the default shapes are valid (teams of whole waves, 128 or 256 lanes,
    # are one per block), and a call with an invalid shape is refused before
    # anything is launched (the device pointers here are never touched)
    lib = _lib.lib()
    t = _lib.TuneStruct()
    assert lib.prt_tune_default(ctypes.byref(t)) == 0
    for b in range(_lib.PRT_TUNE_BUCKETS):
        assert t.lanes_per_pair[b] in (8, 16, 32, 64, 128, 256)
        assert t.lanes_per_pair[b] <= 64 or t.waves_per_block[b] == 1
    fake = ctypes.c_void_p(256)

    def call(tune):
        return lib.prt_overlay_batched(fake, 1, fake, fake, fake, fake, 8, ctypes.byref(tune),
                                       fake, None, 0, None)

    for lanes, waves in [(128, 2), (256, 4), (48, 1), (512, 1)]:
        bad = _lib.TuneStruct()
        assert lib.prt_tune_default(ctypes.byref(bad)) == 0
        bad.lanes_per_pair[3], bad.waves_per_block[3] = lanes, waves
        assert call(bad) == -1, (lanes, waves)
        assert "tune" in lib.prt_last_error().decode()


def test_tune_json_round_trip_and_lds_fit():
    """The autotuner's cache format round-trips a tune, and its LDS fit rule
    follows the library's per-team LDS size: teams of whole waves (more than
    64 lanes) are one per block; smaller teams share a block's LDS."""
    from parlllelrtree_amd import autotune as at
    from parlllelrtree_amd import rtree
    from parlllelrtree_amd._lib import PRT_TUNE_BUCKETS
    t = rtree.default_tune()
    t.lanes_per_pair[1], t.waves_per_block[1], t.list_factor[1] = 32, 2, 4
    d = at.tune_to_json(t, {1: {"pairs": 7, "ms": 0.5}})
    assert d["buckets"][1] == {"max_verts": 64, "lanes": 32, "waves": 2, "list_factor": 4,
                               "pairs": 7, "ms": 0.5}
    u = at.tune_from_json(d)
    for b in range(PRT_TUNE_BUCKETS):
        assert (u.bucket_max_verts[b], u.lanes_per_pair[b], u.waves_per_block[b],
                u.list_factor[b]) == (t.bucket_max_verts[b], t.lanes_per_pair[b],
                                      t.waves_per_block[b], t.list_factor[b])
    tb = at._team_bytes(128, 1)
    assert tb > 0 and at._team_bytes(128, 2) > tb and at._team_bytes(256, 1) > tb
    assert at._fits(128, 128, 1, 1) and not at._fits(128, 128, 2, 1)
    per_block = 160 * 1024 // tb       # teams of the 128-vertex bucket a block's LDS holds
    assert at._fits(128, 64, per_block, 1) and not at._fits(128, 64, per_block + 1, 1)
    assert at._fits(128, 8, per_block // 8, 1) and not at._fits(128, 8, per_block // 8 + 1, 1)
