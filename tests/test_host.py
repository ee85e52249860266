"""Host-side checks that need no GPU: the C ABI library loads and exports
every symbol include/tnl.h declares, argument validation in the library
maps to the reference's exception types, and the Python front end raises
the reference's errors before any device work."""

import re
from pathlib import Path

import pytest
import torch

ROOT = Path(__file__).resolve().parent.parent


def header_symbols():
    text = (ROOT / "include" / "tnl.h").read_text()
    return sorted(set(re.findall(r"\b(tnl_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    from torchlinops_amd import _lib

    h = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 8
    for s in syms:
        assert hasattr(h, s), s
    assert set(syms) == set(_lib.SIGNATURES)
    assert h.tnl_abi_version() == 1


def test_plan_layout_is_host_only_and_consistent():
    from torchlinops_amd.functional._plan import plan_layout

    L = plan_layout(10_240_000, 3, (256, 256, 256), (4.0, 4.0, 4.0), 0)
    assert L.u_nbins == 32 ** 3
    assert tuple(L.ntaps) == (5, 5, 5)
    assert tuple(L.g_tile) == (8, 8, 8)
    assert L.touch_capacity == 8 * 10_240_000
    offs = [L.off_counts, L.off_starts, L.off_cursor, L.off_perm, L.off_slocs,
            L.off_suboff, L.off_sub, L.off_gcounts, L.off_gstarts, L.off_gcursor,
            L.off_tlist, L.off_itemoff, L.off_items, L.off_items_tmp, L.off_tpm]
    assert tuple(L.grid_size) == (256, 256, 256) and L.lds_budget == 53248
    assert all(o % 256 == 0 for o in offs)
    assert offs == sorted(offs) and L.bytes > offs[-1]
    # small grid: a single tile along each axis, one touch per point
    L2 = plan_layout(100, 2, (8, 8), (4.0, 4.0), 0)
    assert tuple(L2.g_ntiles)[1:] == (1, 1) and L2.touch_capacity == 100
    # tuning knobs reach the layout
    L3 = plan_layout(1000, 3, (64, 64, 64), (4.0,) * 3, 0,
                     tuning={"u_tile": (4, 8, 16), "g_cap": 512, "lds_budget": 131072})
    assert tuple(L3.u_tile) == (4, 8, 16) and L3.g_cap == 512 and L3.lds_budget == 131072
    with pytest.raises(ValueError, match="160 KiB"):
        plan_layout(10, 2, (8, 8), (4.0, 4.0), 0, tuning={"lds_budget": 1 << 20})
    # gridding accumulator: matrix cores for 3-D 8x8x8 tiles, LDS otherwise
    assert L.grid_accum == 2 and L2.grid_accum == 1
    # touch pairs of the one-wave gridding kernel: 3-D matrix-core plans only
    assert L.off_tpm > 0 and L2.off_tpm == 0
    assert L.bytes - L.off_tpm >= 8 * (L.touch_capacity + 4)
    L4 = plan_layout(1000, 3, (64, 64, 64), (4.0,) * 3, 0, tuning={"grid_accum": "lds"})
    assert L4.grid_accum == 1
    L5 = plan_layout(1000, 3, (64, 64, 64), (4.0,) * 3, 0, tuning={"g_tile": (8, 8, 16)})
    assert L5.grid_accum == 1
    with pytest.raises(ValueError, match="8x8x8"):
        plan_layout(1000, 3, (64, 64, 64), (4.0,) * 3, 0,
                    tuning={"g_tile": (8, 8, 16), "grid_accum": "mfma"})
    with pytest.raises(ValueError, match="grid_accum"):
        plan_layout(1000, 3, (64, 64, 64), (4.0,) * 3, 0, tuning={"grid_accum": "fast"})


def test_library_validation_maps_to_exceptions():
    from torchlinops_amd import _lib

    h = _lib.lib()
    g = _lib.i64_array([8, 8])
    w = _lib.f32_array([4.0, 4.0])
    with pytest.raises(ValueError, match="kernel"):
        _lib.check(h.tnl_ungrid(None, None, None, 1, 1, 2, g, w, 7, 1, 0, 0, 1.0, None, None, None))
    with pytest.raises(ValueError, match="norm"):
        _lib.check(h.tnl_grid(None, None, None, 1, 1, 2, g, w, 0, 3, 0, 0, 1.0, None, None, None))
    with pytest.raises(ValueError, match="ndim"):
        _lib.check(h.tnl_ungrid(None, None, None, 1, 1, 4, g, w, 0, 1, 0, 0, 1.0, None, None, None))
    big = _lib.f32_array([4.0, 40.0])
    with pytest.raises(ValueError, match="block width"):
        _lib.check(h.tnl_ungrid(None, None, None, 1, 1, 2, g, big, 0, 1, 0, 1, 1.0, None, None, None))
    nb = _lib.i64_array([0])
    with pytest.raises(ValueError, match="nblocks"):
        _lib.check(h.tnl_unfold(None, None, 4, 1, 1, nb, _lib.i64_array([3]),
                                _lib.i64_array([2]), _lib.i64_array([1]), None))
    with pytest.raises(RuntimeError, match="inconsistent"):
        _lib.check(h.tnl_fold(None, None, 0, 1, 1, _lib.i64_array([2]), _lib.i64_array([3]),
                              _lib.i64_array([9]), _lib.i64_array([1]), None))


def test_index_and_wrap_entry_points_validate_before_launch():
    """The sampling and wrapped pad/crop entry points reject bad arguments
    on the host (no pointer is touched, no kernel is launched)."""
    from torchlinops_amd import _lib

    h = _lib.lib()
    with pytest.raises(ValueError, match="elem_bytes"):
        _lib.check(h.tnl_index(1, 1, 1, 3, 1, 8, 4, None))
    with pytest.raises(ValueError, match=">= 0"):
        _lib.check(h.tnl_index(None, None, None, 8, -1, 8, 4, None))
    with pytest.raises(ValueError, match="empty grid"):
        _lib.check(h.tnl_index(1, 1, 1, 8, 1, 0, 4, None))
    assert h.tnl_index(None, None, None, 8, 1, 8, 0, None) == 0  # no points: no work
    with pytest.raises(ValueError, match="heavy"):
        _lib.check(h.tnl_index_adjoint(1, 1, 1, None, 2, 64, 1, 0, 1, 1, 8, 4, None))
    with pytest.raises(ValueError, match="heavy"):
        _lib.check(h.tnl_index_adjoint(1, 1, 1, 1, 2, 0, 1, 0, 1, 1, 8, 4, None))
    with pytest.raises(ValueError, match="dtype"):
        _lib.check(h.tnl_index_adjoint(1, 1, 1, None, 0, 0, 1, 9, 1, 1, 8, 4, None))
    with pytest.raises(ValueError, match="no complex form"):
        _lib.check(h.tnl_index_adjoint(1, 1, 1, None, 0, 0, 1, _lib.TNL_I32, 1, 1, 8, 4, None))
    for code in (_lib.TNL_I8, _lib.TNL_U8, _lib.TNL_I16, _lib.TNL_I32, _lib.TNL_I64):
        # integer element types are accepted (no batch: no work)
        assert h.tnl_index_adjoint(None, None, None, None, 0, 0, None, code, 0, 0, 8, 4,
                                   None) == 0
    im, pad = _lib.i64_array([8, 8]), _lib.i64_array([16, 6])
    with pytest.raises(ValueError, match="im_size <= pad_size"):
        _lib.check(h.tnl_pad_wrapped(1, 1, 8, 1, 2, im, pad, None, None))
    ok = _lib.i64_array([16, 16])
    with pytest.raises(ValueError, match="ndim"):
        _lib.check(h.tnl_crop_wrapped(1, 1, 8, 1, 4, im, ok, None, None))
    with pytest.raises(ValueError, match="weight"):
        _lib.check(h.tnl_pad_wrapped(1, 1, 16, 1, 2, im, ok, 1, None))
    with pytest.raises(ValueError, match="elem_bytes"):
        _lib.check(h.tnl_crop_wrapped(1, 1, 2, 1, 2, im, ok, None, None))


def test_functional_errors_before_device_work():
    from torchlinops_amd import functional as F

    vals = torch.zeros(8, 8, dtype=torch.complex64)
    locs = torch.zeros(5, 2)
    with pytest.raises(ValueError, match="Unrecognized kernel"):
        F.interpolate(vals, locs, 4.0, kernel="gauss")
    with pytest.raises(ValueError, match="same length"):
        F.interpolate(vals, locs, (4.0, 4.0, 4.0))
    with pytest.raises(ValueError, match="last dim of locs"):
        F.interpolate_adjoint(torch.zeros(5), locs, (8, 8, 8), 4.0)
    with pytest.raises(ValueError, match="block width"):
        F.interpolate(vals, locs, (4.0, 40.0))
    with pytest.raises(ValueError, match="Found 0 in nblocks"):
        F.array_to_blocks(torch.zeros(4, 4), (5, 5), (1, 1))
    with pytest.raises(ValueError, match="dtypes must match"):
        F.array_to_blocks(torch.zeros(9, 9), (3, 3), (1, 1), out=torch.zeros(1, dtype=torch.float64))
    with pytest.raises(RuntimeError, match="Fold expected input"):
        F.blocks_to_array(torch.zeros(2, 2, 3, 3), (9, 9), (3, 3), (2, 2))
    # host tensors take the host path; ops with no host path still refuse them
    assert F.interpolate(vals, locs, 4.0).shape == (5,)
    assert F.array_to_blocks(torch.zeros(9, 9), (3, 3), (1, 1)).shape == (7, 7, 3, 3)
    from torchlinops_amd.linops.toeplitz import pad_wrapped

    with pytest.raises(RuntimeError, match="HIP devices only"):
        pad_wrapped(vals, (8, 8), (16, 16))
    with pytest.raises(RuntimeError, match="one device"):
        F.interpolate(vals, locs.to("meta"), 4.0)


def test_get_nblocks_doc_examples():
    from torchlinops_amd.functional import get_nblocks

    assert get_nblocks((5, 5), (3, 3), (1, 1)) == (3, 3)
    assert get_nblocks((5, 5), (3, 3), (2, 2)) == (2, 2)
    assert get_nblocks((6, 6), (3, 3), (2, 2)) == (2, 2)
    assert get_nblocks((7, 7), (3, 3), (2, 2)) == (3, 3)
    assert get_nblocks((10, 10), (8, 8), (4, 4)) == (1, 1)


def test_plan_cache_never_builds_under_graph_capture(monkeypatch):
    """While a stream is being captured a plan build would only be recorded:
    the cache must not build or store one then (a miss takes the plan-free
    kernels), must hand out and pin a plan it already holds, and must report
    every plan used to an active recorder."""
    from torchlinops_amd import config
    from torchlinops_amd.functional import _plan

    _plan.clear_plan_cache()
    _plan.release_plan_pins()
    built = []

    def fake_build(locs2d, ndim, grid_size, width, pad_code, tuning=None):
        built.append(locs2d)
        return _plan.Plan(layout=None, workspace=torch.empty(1, dtype=torch.uint8))

    monkeypatch.setattr(_plan, "build_plan", fake_build)
    locs = torch.rand(100, 3) * 16
    args = (locs, locs, 3, (16, 16, 16), (4.0,) * 3, 1, 0)
    capturing = {"on": True}
    monkeypatch.setattr(_plan, "_capturing", lambda device: capturing["on"])
    with config.using(interp_plan="always"):
        # capture, cache miss: no build, no store, plan-free kernels
        assert _plan.get_plan(*args) is None
        assert built == []
        # eager: builds and caches
        capturing["on"] = False
        plan = _plan.get_plan(*args)
        assert plan is not None and len(built) == 1
        # capture, cache hit: same plan, pinned, recorded, no build
        capturing["on"] = True
        sink = []
        with _plan.record_plans(sink):
            assert _plan.get_plan(*args) is plan
        assert sink == [plan] and len(built) == 1
        assert any(p is plan for p in _plan._cache._pinned.values())
        # clearing the cache keeps the pin (a live graph may point at it)
        _plan.clear_plan_cache()
        assert any(p is plan for p in _plan._cache._pinned.values())
        assert _plan.get_plan(*args) is None and len(built) == 1
    _plan.release_plan_pins()
    assert not _plan._cache._pinned


def test_grid_static_rejects_plans_beyond_its_scratch():
    """tnl_grid_static refuses, before any device work, a plan whose point
    records exceed TNL_GRID_STATIC_SCRATCH_BYTES (64 MiB = 1 048 576
    records of 64 bytes); within the limit it needs the records a plan's
    tnl_grid_scratch_bytes reports: one channel record per point and chunk
    of 8 complex channels, plus one window record per point for plans with
    touch pairs (3-D, 8x8x8 tiles)."""
    import ctypes

    from torchlinops_amd import _lib
    from torchlinops_amd.functional._plan import plan_layout

    h = _lib.lib()
    text = (ROOT / "include" / "tnl.h").read_text()
    assert "#define TNL_GRID_STATIC_SCRATCH_BYTES (64ll << 20)" in text
    L = plan_layout(1_100_000, 3, (64, 64, 64), (4.0,) * 3, 0)  # 8x8x8 tiles: matrix cores
    need = ctypes.c_int64(0)
    _lib.check(h.tnl_grid_scratch_bytes(L, 8, 1, ctypes.byref(need)))
    assert L.off_tpm > 0
    assert need.value == 1_100_000 * 64 * 2 > (64 << 20)
    L2 = plan_layout(1_100_000, 2, (256, 256), (4.0,) * 2, 0)  # no touch pairs in 2-D
    _lib.check(h.tnl_grid_scratch_bytes(L2, 8, 1, ctypes.byref(need)))
    assert L2.off_tpm == 0 and need.value == 1_100_000 * 64
    args = [None, None, None, 8, 1_100_000, 3, _lib.i64_array((64, 64, 64)),
            _lib.f32_array((4.0,) * 3), 0, 1, 0, 1, 8.99, L, None, None]
    with pytest.raises(ValueError, match="more than the static 67108864 bytes"):
        _lib.check(h.tnl_grid_static(*args))
