"""The fused gather's index arithmetic replayed on the host (no GPU): for the
layouts the GPU tests use, every work item's classification, queue walk,
copy slot, element-queue slot, transpose tile (both phases, LDS included)
and scaled element is checked against the transform's definition -- every
final element receives exactly its source element, nothing outside the
final blocks is written, LDS reads hit cells written by the same tile, queue
walks stay within the queue. The replay runs the kernels' own __host__
__device__ index functions (grid2grid_amd/csrc/g2g_index.hpp), so a kernel
variant whose indexing is wrong fails here before it reaches a GPU."""
import ctypes
import os

import pytest

import shapes
import grid2grid_amd as g2g
from grid2grid_amd import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "cpp", "bin", "libgather_replay.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="tests/cpp not built (grid2grid_amd.build)")

FEAT_COPY, FEAT_TRANS, FEAT_ALL = 0, 1, 2
DTYPES = [g2g.FLOAT32, g2g.FLOAT64, g2g.COMPLEX64, g2g.COMPLEX128]
_lib = None


def _run(init, fin, dtype, trans, scaled, tile_bytes, threads, feat, chunk_w, plant, order=0, chunk_h=0, split=0):
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(LIB)
        _lib.replay_gather.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_char, ctypes.c_int,
                                       ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                       ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int64)]
    if feat is None:
        feat = FEAT_ALL if scaled else (FEAT_TRANS if trans != "N" else FEAT_COPY)
    es = g2g.DTYPE_SIZE[dtype]
    # addresses as on a device: the two arenas far apart, 256-byte aligned
    di = init.descriptor(1 << 40, es)
    df = fin.descriptor(1 << 44, es)
    msg = ctypes.create_string_buffer(4096)
    stats = (ctypes.c_int64 * 6)()
    errors = _lib.replay_gather(ctypes.addressof(di.struct), ctypes.addressof(df.struct), dtype, trans.encode(),
                                int(scaled), tile_bytes, threads, feat, chunk_w, order, chunk_h, split, plant, msg,
                                len(msg),
                                stats)
    return errors, msg.value.decode(), list(stats)


def _replay(init, fin, dtype, trans="N", scaled=False, tile_bytes=65536, threads=512, feat=None, chunk_w=0,
            order=0, chunk_h=0, split=0):
    errors, msg, stats = _run(init, fin, dtype, trans, scaled, tile_bytes, threads, feat, chunk_w, 0, order, chunk_h,
                              split)
    assert errors == 0, (f"{errors} errors ({init.m}x{init.n} dtype {dtype} trans {trans} scaled {scaled} "
                         f"tile {tile_bytes} threads {threads} feat {feat} chunks {chunk_w} order {order} "
                         f"rows {chunk_h} split {split}):\n" + msg)
    return stats


CONFIGS = [(4096, 256), (65536, 512), (1 << 20, 512), (12288, 256)]


@pytest.mark.parametrize("name", list(shapes.layout_pairs()))
def test_replay_untransposed(name):
    init, fin = shapes.layout_pairs()[name]
    for dtype in DTYPES:
        for tb, th in CONFIGS:
            _replay(init, fin, dtype, tile_bytes=tb, threads=th)
            _replay(init, fin, dtype, tile_bytes=tb, threads=th, feat=FEAT_ALL)
        _replay(init, fin, dtype, scaled=True, tile_bytes=4096, threads=256)
        for w in (1, 3, 16):
            _replay(init, fin, dtype, tile_bytes=8192, threads=512, chunk_w=w)
        for w, tb in ((0, 40960), (2, 12288), (16, 49152)):
            _replay(init, fin, dtype, tile_bytes=tb, threads=512, chunk_w=w, order=1)
        for w, h, o in ((0, 3, 0), (8, 16, 1), (3, 7, 1)):
            _replay(init, fin, dtype, tile_bytes=12288, threads=512, chunk_w=w, order=o, chunk_h=h)
        for sp, w, o in ((2, 0, 0), (4, 8, 1), (8, 0, 1)):
            _replay(init, fin, dtype, tile_bytes=12288, threads=256, chunk_w=w, order=o, split=sp)


@pytest.mark.parametrize("name", list(shapes.transposed_pairs()))
def test_replay_transposed(name):
    init, fin = shapes.transposed_pairs()[name]
    for dtype in DTYPES:
        for trans in ("T", "C") if dtype in (g2g.COMPLEX64, g2g.COMPLEX128) else ("T",):
            for tb, th in CONFIGS:
                st = _replay(init, fin, dtype, trans, tile_bytes=tb, threads=th)
                assert st[1] == 0  # no copy-queue slots in a transposed plan
            _replay(init, fin, dtype, trans, scaled=True, tile_bytes=65536, threads=512)
            q = 256 // g2g.DTYPE_SIZE[dtype]
            for w in (3, q, 2 * q):
                _replay(init, fin, dtype, trans, tile_bytes=65536, threads=512, chunk_w=w)
            for w, tb in ((0, 49152), (q, 40960)):
                _replay(init, fin, dtype, trans, tile_bytes=tb, threads=512, chunk_w=w, order=1)
            for sp in (2, 8):
                _replay(init, fin, dtype, trans, tile_bytes=65536, threads=512, split=sp)


def test_replay_cut_edges_take_both_paths():
    """The element-queue / LDS-tile cut: f32 pieces of 127 final rows go to
    the element queue and of 128 rows to LDS tiles, 8-byte pieces of 63 and
    64 rows likewise; both paths are exercised and exact."""
    init, fin = shapes.transposed_pairs()["bcT_cut_edges"]
    for dtype in DTYPES:
        st = _replay(init, fin, dtype, "T", tile_bytes=1 << 20, threads=512)
        assert st[2] > 0 and st[3] > 0, st  # element slots and transpose tile vectors


@pytest.mark.parametrize("dtype", [g2g.FLOAT32, g2g.FLOAT64, g2g.COMPLEX64])
def test_replay_ragged_columns(dtype):
    """Columns of 4-88 bytes: the dword path and the overlapping last slot,
    whose repeated writes carry the same source element."""
    init, fin, _ = shapes.ragged_pair(dtype)
    for tb, th in CONFIGS:
        st = _replay(init, fin, dtype, tile_bytes=tb, threads=th)
        assert st[1] > 0
    _replay(init, fin, dtype, tile_bytes=4096, threads=512, chunk_w=8)


def test_replay_benchmark_block_cyclic_4096():
    """The 4096^2 f64 block-cyclic benchmark pair at full size (16 384
    pieces), on the pieces and on chunks of 16 columns, and with the last
    quarter of its work items (256 CUs x 4 items of 32 KiB) split."""
    init, fin, dtype, trans = workloads.bc4096()
    st = _replay(init, fin, dtype, tile_bytes=65536, threads=512)
    assert st[0] == 2048
    _replay(init, fin, dtype, tile_bytes=65536, threads=512, chunk_w=16)
    _replay(init, fin, dtype, tile_bytes=49152, threads=512, chunk_w=2, order=1)
    for sp in (2, 4):
        st = _replay(init, fin, dtype, tile_bytes=32768, threads=256, chunk_w=64, split=sp)
        assert st[0] == 3072 + 1024 * sp


def test_replay_mixed_requests():
    """Transforms of the mixed-request distribution (sizes up to 600), as the
    batch tiles them (64 KiB work items, 512 threads)."""
    import numpy as np
    rng = np.random.default_rng(91)
    for k in range(24):
        init, fin, dtype, trans = workloads.random_request(rng, dtypes=DTYPES, lo=16, hi=600, p_trans=0.5)
        _replay(init, fin, dtype, trans, tile_bytes=65536, threads=512)


@pytest.mark.parametrize("dtype", [g2g.FLOAT32, g2g.FLOAT64, g2g.COMPLEX128])
def test_replay_irregular_small(dtype):
    """The irregular benchmark construction at 6000 x 400 (ragged splits,
    padded leading dimensions, shuffled blocks: columns start at every dword
    of a 128-byte line), on the pieces and on column chunks, in union and in
    destination order."""
    init, fin = shapes.irregular_small()
    for tb, th in CONFIGS:
        st = _replay(init, fin, dtype, tile_bytes=tb, threads=th)
        assert st[1] > 0, st
    _replay(init, fin, dtype, tile_bytes=65536, threads=512, chunk_w=16)
    for w, tb in ((0, 40960), (2, 49152)):
        _replay(init, fin, dtype, tile_bytes=tb, threads=512, chunk_w=w, order=1)
    for w, h in ((8, 256), (32, 128)):
        _replay(init, fin, dtype, tile_bytes=65536, threads=512, chunk_w=w, order=1, chunk_h=h)
    _replay(init, fin, dtype, tile_bytes=40960, threads=512, chunk_w=2, order=1, split=2)


def test_replay_random_streams():
    """The random requests and tunings of the GPU test
    test_random_requests_random_streams, replayed: every work item, queue
    slot and transpose tile of each stream (column and row chunks, either
    order, any work-item size, split tails) checked against the transform's
    definition."""
    feats = {"N": FEAT_COPY, "T": FEAT_TRANS, "C": FEAT_TRANS}
    for k, (init, fin, dtype, trans, tu) in enumerate(shapes.random_stream_cases()):
        tile, th, ub, nt, mode, w, order, h, sp = tu
        _replay(init, fin, dtype, trans, tile_bytes=tile, threads=th, feat=feats[trans], chunk_w=w, order=order,
                chunk_h=h, split=sp)


@pytest.mark.parametrize("plant,case", [(1, "ragged"), (2, "ragged"), (3, "cut"), (4, "cut"), (5, "cut"), (6, "arb"),
                                        (1, "irr"), (2, "irr"), (6, "irr")])
def test_replay_catches_planted_mistakes(plant, case):
    """A typical indexing mistake planted into the replayed arithmetic (the
    copy queue's last slot not pulled back, leading dimensions swapped, an
    element tile read along the wrong axis, a transpose tile one column off,
    the LDS row padding dropped, work items overlapping by a column) is
    reported."""
    if case == "ragged":
        init, fin, _ = shapes.ragged_pair(g2g.FLOAT32)
        args = (init, fin, g2g.FLOAT32, "N", False, 4096, 256, FEAT_COPY, 0)
    elif case == "cut":
        init, fin = shapes.transposed_pairs()["bcT_cut_edges"]
        args = (init, fin, g2g.FLOAT64, "T", False, 1 << 20, 512, FEAT_TRANS, 0)
    elif case == "irr":
        init, fin = shapes.irregular_small()
        args = (init, fin, g2g.FLOAT32, "N", False, 65536, 512, FEAT_COPY, 0)
    else:
        init, fin = shapes.layout_pairs()["arbitrary"]
        args = (init, fin, g2g.FLOAT64, "N", False, 4096, 512, FEAT_COPY, 0)
    assert _run(*args, 0)[0] == 0
    errors, msg, _ = _run(*args, plant)
    assert errors > 0, f"planted mistake {plant} not caught"
