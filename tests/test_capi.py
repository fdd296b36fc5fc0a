"""The C ABI library without a GPU: it loads, exports every function the
public headers declare, and rejects invalid layouts with the reference's
error conditions before touching the device."""
import ctypes
import os
import re

import numpy as np
import pytest

import grid2grid_amd as g2g
from grid2grid_amd import _lib
from grid2grid_amd.layout import Layout, scalapack_layout

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(g2g_\w+)\s*\(", text, re.M))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared_functions()
    assert len(names) >= 15
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(_lib.PROTOTYPES) >= names


def test_version():
    """The library reports the ABI version its header declares."""
    text = open(os.path.join(os.path.dirname(__file__), "..", "include", "grid2grid_amd.h")).read()
    want = int(re.search(r"#define G2G_ABI_VERSION (\d+)", text).group(1))
    assert _lib.load().g2g_version() == want == 4


def _plan_err(init, fin, dtype=g2g.FLOAT64, trans="N"):
    with pytest.raises(g2g.G2GError) as e:
        g2g.Plan(init, 0x1000, fin, 0x100000, dtype, trans)
    return e.value


def test_rejects_bad_trans_and_dtype():
    a = scalapack_layout(8, 8, 2, 2, 2, 2)
    assert _plan_err(a, a, trans="X").code == 1
    assert _plan_err(a, a, dtype=7).code == 1


def test_rejects_mismatched_extent():
    """The reference asserts the grids span the same matrix (grid_cover.cpp:70)."""
    a = scalapack_layout(8, 8, 2, 2, 2, 2)
    b = scalapack_layout(8, 9, 2, 2, 2, 2)
    e = _plan_err(a, b)
    assert e.code == 1 and "differ" in str(e)


def test_rejects_transposed_shape_mismatch():
    a = scalapack_layout(8, 6, 2, 2, 2, 2)
    assert _plan_err(a, a, trans="T").code == 1


def test_rejects_owner_out_of_range():
    """rank outside the rank grid (reference scalapack_layout.cpp:12-15, 43-48)."""
    a = scalapack_layout(8, 8, 2, 2, 2, 2)
    bad = Layout(a.rows_split, a.cols_split, np.where(a.owners == 3, 9, a.owners), 4, a.blk_row, a.blk_col,
                 a.blk_stride, a.blk_offset, a.arena_elems)
    assert "owner" in str(_plan_err(bad, a))


def test_rejects_block_outside_grid():
    """block coordinates outside the cover (reference grid_cover.hpp:57-62)."""
    a = scalapack_layout(8, 8, 2, 2, 2, 2)
    row = a.blk_row.copy()
    row[0] = 99
    bad = Layout(a.rows_split, a.cols_split, a.owners, 4, row, a.blk_col, a.blk_stride, a.blk_offset,
                 a.arena_elems)
    assert "grid" in str(_plan_err(bad, a))


def test_rejects_missing_and_duplicate_blocks():
    a = scalapack_layout(8, 8, 2, 2, 2, 2)
    miss = Layout(a.rows_split, a.cols_split, a.owners, 4, a.blk_row[1:], a.blk_col[1:], a.blk_stride[1:],
                  a.blk_offset[1:], a.arena_elems)
    assert "no block" in str(_plan_err(miss, a))
    dup = Layout(a.rows_split, a.cols_split, a.owners, 4, np.r_[a.blk_row, a.blk_row[:1]],
                 np.r_[a.blk_col, a.blk_col[:1]], np.r_[a.blk_stride, a.blk_stride[:1]],
                 np.r_[a.blk_offset, a.blk_offset[:1]], a.arena_elems)
    assert "two blocks" in str(_plan_err(dup, a))


def test_rejects_bad_splits_and_stride():
    """negative or inverted intervals (reference interval.cpp:8-16)."""
    a = scalapack_layout(8, 8, 2, 2, 2, 2)
    rs = a.rows_split.copy()
    rs[2] = rs[1]
    bad = Layout(rs, a.cols_split, a.owners, 4, a.blk_row, a.blk_col, a.blk_stride, a.blk_offset, a.arena_elems)
    assert "increasing" in str(_plan_err(bad, a))
    st = a.blk_stride.copy()
    st[3] = 1
    bad2 = Layout(a.rows_split, a.cols_split, a.owners, 4, a.blk_row, a.blk_col, st, a.blk_offset, a.arena_elems)
    assert "stride" in str(_plan_err(bad2, a))


def test_tuning_bounds():
    with pytest.raises(g2g.G2GError):
        g2g.set_tuning(tile_bytes=100)
    g2g.set_tuning(0, -1)


def test_last_error_is_thread_local_message():
    a = scalapack_layout(8, 8, 2, 2, 2, 2)
    _plan_err(a, a, trans="Q")
    assert b"trans" in _lib.load().g2g_last_error()


def test_layout_descriptor_is_kept_per_address():
    """A layout keeps the descriptors it built (the drop-in call in a loop
    rebuilds nothing); another arena address gets its own, whose block
    pointers move with the base."""
    from grid2grid_amd.layout import scalapack_layout
    lay = scalapack_layout(100, 80, 16, 16, 2, 2, "C", lld="rows")
    a = lay.descriptor(1 << 20, 8)
    assert lay.descriptor(1 << 20, 8) is a
    b = lay.descriptor(1 << 21, 8)
    assert b is not a
    np.testing.assert_array_equal(b.blocks["data"] - a.blocks["data"], np.uint64(1 << 20))
    assert lay.descriptor(1 << 20, 4) is not a  # the element size is part of the address arithmetic



def test_missing_library_fails_loudly(monkeypatch):
    """Without the built HIP library nothing falls back to a CPU path: loading
    it, and with it creating a plan, raises."""
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", os.path.join(os.path.dirname(_lib.LIB_PATH), "absent.so"))
    with pytest.raises(ImportError):
        _lib.load()
    lay = scalapack_layout(8, 8, 4, 4, 1, 1, "C")
    with pytest.raises(ImportError):
        g2g.Plan(lay, 0, lay, 0, g2g.FLOAT64)
