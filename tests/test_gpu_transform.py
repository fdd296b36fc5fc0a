"""GPU parity: the HIP transform (fused gather and plain staged form),
called through the C ABI, against the CPU oracle on the same bytes.

Arenas are filled with random bytes (NaN payloads, -0.0, denormals
included) so bit-exactness is checked on arbitrary patterns; every byte of
the destination arena is compared, so ld padding and gaps between blocks
must stay untouched."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
import shapes
import grid2grid_amd as g2g
from grid2grid_amd.layout import scalapack_layout, arbitrary_layout

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DTYPES = [g2g.FLOAT32, g2g.FLOAT64, g2g.COMPLEX64, g2g.COMPLEX128]


def _dev(host):
    return torch.from_numpy(host).to("cuda")


def _rand_arena(layout, dtype, rng):
    return rng.integers(0, 256, size=layout.arena_elems * g2g.DTYPE_SIZE[dtype], dtype=np.uint8)


def _layout_pairs():
    return shapes.layout_pairs()


def _transposed_pairs():
    return shapes.transposed_pairs()


def _expected(init, ia, fin, fa, dtype, trans="N", alpha=None, beta=None):
    out = fa.copy()
    O.transform(init, ia.copy(), fin, out, dtype, trans, alpha, beta)
    return out


def _run(init, ia, fin, fa, dtype, trans="N", alpha=None, beta=None, staged=False):
    ti, tf = _dev(ia), _dev(fa)
    plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans)
    if staged:
        plan.transform_staged(alpha, beta)
    else:
        plan.transform(alpha, beta)
    torch.cuda.synchronize()
    got = tf.cpu().numpy()
    # the source is never modified
    np.testing.assert_array_equal(ti.cpu().numpy(), ia)
    plan.close()
    return got


@pytest.mark.parametrize("staged", [False, True], ids=["fused", "staged"])
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("name", list(_layout_pairs()))
def test_transform_bit_exact(name, dtype, staged):
    init, fin = _layout_pairs()[name]
    rng = np.random.default_rng(hash((name, dtype)) % 2**32)
    ia, fa = _rand_arena(init, dtype, rng), _rand_arena(fin, dtype, rng)
    print(f"case {name} dtype {dtype} staged {staged}", flush=True)
    want = _expected(init, ia, fin, fa, dtype)
    got = _run(init, ia, fin, fa, dtype, staged=staged)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("staged", [False, True], ids=["fused", "staged"])
@pytest.mark.parametrize("trans", ["T", "C"])
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("name", list(_transposed_pairs()))
def test_transform_transposed_bit_exact(name, dtype, trans, staged):
    init, fin = _transposed_pairs()[name]
    rng = np.random.default_rng(hash((name, dtype, trans)) % 2**32)
    ia, fa = _rand_arena(init, dtype, rng), _rand_arena(fin, dtype, rng)
    print(f"case {name} dtype {dtype} trans {trans} staged {staged}", flush=True)
    want = _expected(init, ia, fin, fa, dtype, trans)
    got = _run(init, ia, fin, fa, dtype, trans, staged=staged)
    np.testing.assert_array_equal(got, want)


SCALARS = {
    g2g.FLOAT32: (np.float32(0.75), np.float32(-1.25)),
    g2g.FLOAT64: (0.3, 1.7),
    g2g.COMPLEX64: (np.complex64(0.5 - 1.5j), np.complex64(-0.25 + 2.0j)),
    g2g.COMPLEX128: (0.7 + 0.1j, -1.1 + 0.9j),
}


@pytest.mark.parametrize("staged", [False, True], ids=["fused", "staged"])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("dtype", DTYPES)
def test_scaled_bit_exact(dtype, trans, staged):
    """beta*dst + alpha*op(src) on finite seeded values: identical bits to the
    oracle (no FMA contraction on the GPU)."""
    pairs = _layout_pairs() if trans == "N" else _transposed_pairs()
    alpha, beta = SCALARS[dtype]
    es = g2g.DTYPE_SIZE[dtype]
    for name in list(pairs)[:3]:
        init, fin = pairs[name]
        ia = np.zeros(init.arena_elems * es, dtype=np.uint8)
        fa = np.zeros(fin.arena_elems * es, dtype=np.uint8)
        O.fill(init, ia, dtype, 7)
        O.fill(fin, fa, dtype, 8)
        want = _expected(init, ia, fin, fa, dtype, trans, alpha, beta)
        got = _run(init, ia, fin, fa, dtype, trans, alpha, beta, staged=staged)
        np.testing.assert_array_equal(got, want, err_msg=name)


@pytest.mark.parametrize("name", ["bc_small", "bc_grid_change", "arbitrary", "many_ranks"])
def test_wire_order_matches_oracle(name):
    """Counts per (source, destination) and every piece's offset in the send
    and receive buffers equal the reference's communication_data (via the
    oracle, itself pinned to the reference)."""
    init, fin = _layout_pairs()[name]
    dtype = g2g.FLOAT64
    rng = np.random.default_rng(5)
    ia, fa = _rand_arena(init, dtype, rng), _rand_arena(fin, dtype, rng)
    ti, tf = _dev(ia), _dev(fa)
    plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype)
    counts = plan.counts()
    w = plan.wire()
    P = max(init.n_ranks, fin.n_ranks)
    for r in range(P):
        msgs, c_send, _ = O.plan_rank(init, ia, fin, fa, dtype, "N", r, 0)
        np.testing.assert_array_equal(counts[r], c_send)
        want = {(m.r0, m.c0): m.offset for m in msgs if m.offset >= 0}
        sel = (w["src"] == r) & (w["dst"] != r)
        got = {(int(a), int(b)): int(o) for a, b, o in zip(w["row0"][sel], w["col0"][sel], w["send_off"][sel])}
        assert got == want
        msgs, c_recv, _ = O.plan_rank(init, ia, fin, fa, dtype, "N", r, 1)
        np.testing.assert_array_equal(counts[:, r], c_recv)
        want = {(m.r0, m.c0): m.offset for m in msgs if m.offset >= 0}
        sel = (w["dst"] == r) & (w["src"] != r)
        got = {(int(a), int(b)): int(o) for a, b, o in zip(w["row0"][sel], w["col0"][sel], w["recv_off"][sel])}
        assert got == want
    plan.close()


_hip = None


def _hip_memcpy_d2h(ptr, nbytes):
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    out = np.zeros(nbytes, dtype=np.uint8)
    if nbytes:
        assert _hip.hipMemcpy(out.ctypes.data, ptr, nbytes, 2) == 0
    return out


@pytest.mark.parametrize("trans", ["N", "T"])
def test_staged_send_buffers_match_reference_packing(trans):
    """After the plain form ran, each virtual rank's packed send buffer holds
    the same bytes as the reference's copy_to_buffer (via the oracle)."""
    name = "bc_grid_change" if trans == "N" else "bcT"
    init, fin = (_layout_pairs() if trans == "N" else _transposed_pairs())[name]
    dtype = g2g.COMPLEX64
    rng = np.random.default_rng(9)
    ia, fa = _rand_arena(init, dtype, rng), _rand_arena(fin, dtype, rng)
    ti, tf = _dev(ia), _dev(fa)
    plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans)
    plan.transform_staged()
    torch.cuda.synchronize()
    for r in range(max(init.n_ranks, fin.n_ranks)):
        ptr, n = plan.send_buffer(r)
        got = _hip_memcpy_d2h(ptr, n * g2g.DTYPE_SIZE[dtype])
        want = O.pack_rank(init, ia, fin, fa, dtype, trans, r)
        np.testing.assert_array_equal(got, want)
    plan.close()


def test_repeat_and_reuse_plan():
    """A plan is reusable; running it twice (and staged after fused) gives
    the same bytes."""
    init, fin = _layout_pairs()["arbitrary"]
    dtype = g2g.FLOAT32
    rng = np.random.default_rng(1)
    ia, fa = _rand_arena(init, dtype, rng), _rand_arena(fin, dtype, rng)
    want = _expected(init, ia, fin, fa, dtype)
    ti, tf = _dev(ia), _dev(fa)
    plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype)
    for k in range(3):
        tf.copy_(_dev(fa))
        (plan.transform_staged if k == 1 else plan.transform)()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(tf.cpu().numpy(), want)
    plan.close()


@pytest.mark.parametrize("tile_bytes,launch_mode", [(4096, 0), (4096, 1), (65536, 1), (1 << 20, 0), (4096, 2),
                                                     (65536, 2), (4096, 3), (1 << 20, 3)])
def test_work_item_sizes(tile_bytes, launch_mode):
    """Results do not depend on the work-item size or launch mode."""
    init, fin = _layout_pairs()["arbitrary"]
    dtype = g2g.COMPLEX128
    rng = np.random.default_rng(77)
    ia, fa = _rand_arena(init, dtype, rng), _rand_arena(fin, dtype, rng)
    want = _expected(init, ia, fin, fa, dtype)
    g2g.set_tuning(tile_bytes, launch_mode)
    try:
        got = _run(init, ia, fin, fa, dtype)
        got_t = None
        initT, finT = _transposed_pairs()["arbT"]
        iaT, faT = _rand_arena(initT, dtype, rng), _rand_arena(finT, dtype, rng)
        wantT = _expected(initT, iaT, finT, faT, dtype, "C")
        got_t = _run(initT, iaT, finT, faT, dtype, "C")
    finally:
        g2g.set_tuning(0, -1)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(got_t, wantT)


@pytest.mark.parametrize("dtype", [g2g.FLOAT32, g2g.FLOAT64, g2g.COMPLEX64])
def test_ragged_column_tails(dtype):
    """Pieces of 1-11 rows (4 to 88 bytes a column), blocks with 0-3 elements
    of ld padding and gaps, both arenas filled with random bytes: the
    copy queue moves a column whose byte length is not a multiple of 16 with
    a last 16-byte slot that overlaps the one before it, and a column shorter
    than 16 bytes dword by dword. Every byte of the final arena (padding and
    gaps included) matches the oracle, for every work-item size, launch
    mode and the column-chunked stream."""
    init, fin, rng = shapes.ragged_pair(dtype)
    ia, fa = _rand_arena(init, dtype, rng), _rand_arena(fin, dtype, rng)
    want = _expected(init, ia, fin, fa, dtype)
    ti = _dev(ia)
    plan = None
    try:
        for tb, th, ub, nt, mode, chunks in ((4096, 256, 64, 0, 0, 0), (4096, 512, 128, 1, 1, 0),
                                             (65536, 512, 64, 1, 1, 0), (65536, 256, 128, 0, 2, 0),
                                             (4096, 512, 64, 0, 3, 0), (4096, 512, 64, 1, 1, 8)):
            print(f"dtype {dtype} tuning {tb}:{th}:{ub}:{nt}:{mode}:{chunks}", flush=True)
            tf = _dev(fa)
            plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, "N")
            plan.set_tuning(tb, th, ub, nt, mode, column_chunks=chunks)
            plan.transform()
            torch.cuda.synchronize()
            np.testing.assert_array_equal(tf.cpu().numpy(), want)
            plan.close()
            plan = None
    finally:
        if plan is not None:
            plan.close()
    np.testing.assert_array_equal(ti.cpu().numpy(), ia)


@pytest.mark.parametrize("dtype", DTYPES)
def test_transposed_whole_tiles(dtype):
    """Transposed pieces of 260 (and 264) final rows by 150 (and 146) final
    columns: whole transpose tiles moved several per pass (the last pass
    partly filled), then a bottom strip of 4 (8) final rows and a right strip
    of the remaining final columns; both workgroup sizes (2 and 4 tiles per
    pass), every launch mode, at work-item sizes that give each workgroup one
    column group of a piece or whole pieces."""
    init = scalapack_layout(300, 524, 150, 260, 2, 2, "C")
    fin = scalapack_layout(524, 300, 524, 300, 1, 1, "C")
    rng = np.random.default_rng(31)
    ia, fa = _rand_arena(init, dtype, rng), _rand_arena(fin, dtype, rng)
    for trans in (["T", "C"] if dtype in (g2g.COMPLEX64, g2g.COMPLEX128) else ["T"]):
        want = _expected(init, ia, fin, fa, dtype, trans)
        ti = _dev(ia)
        for threads, tile_bytes, mode in ((256, 65536, 0), (512, 65536, 1), (256, 1 << 20, 2), (512, 1 << 22, 0),
                                          (512, 1 << 20, 3), (256, 1 << 22, 1)):
            print(f"whole tiles dtype {dtype} trans {trans} threads {threads} tile_bytes {tile_bytes} mode {mode}",
                  flush=True)
            tf = _dev(fa)
            plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans)
            plan.set_tuning(tile_bytes, threads, 128, 1, mode)
            plan.transform()
            torch.cuda.synchronize()
            np.testing.assert_array_equal(tf.cpu().numpy(), want)
            plan.close()


@pytest.mark.parametrize("name", ["bc_small", "bc_grid_change", "arbitrary", "tall_thin", "many_ranks"])
def test_column_chunked_stream(name):
    """The gather over column chunks (pieces cut into chunks of w columns,
    ordered down each column interval) and over the pieces or chunks in
    destination order writes exactly what the pieces in their own order
    write, for odd widths, work-item sizes that are not powers of two and
    every launch mode, also cut into row chunks and with the last round of
    work items split; the width, order, row chunk height and tail split are
    reported back; negative widths or heights, an unknown order and a tail
    split other than 0, 2, 4 or 8 are refused."""
    init, fin = _layout_pairs()[name]
    dtype = g2g.FLOAT32
    rng = np.random.default_rng(17)
    ia, fa = _rand_arena(init, dtype, rng), _rand_arena(fin, dtype, rng)
    want = _expected(init, ia, fin, fa, dtype)
    ti = _dev(ia)
    for w, mode, tile, order, h, sp in ((1, 0, 4096, 0, 0, 0), (3, 1, 65536, 1, 0, 2), (16, 2, 4096, 0, 0, 4),
                                        (64, 3, 1 << 20, 1, 0, 0), (3, 0, 1 << 20, 0, 0, 8), (0, 1, 40960, 1, 0, 0),
                                        (0, 0, 4096, 1, 0, 2), (0, 3, 49152, 1, 0, 4), (16, 1, 12288, 1, 0, 0),
                                        (0, 1, 4096, 0, 5, 8), (8, 0, 12288, 1, 16, 0), (3, 3, 65536, 1, 3, 4),
                                        (32, 1, 8192, 0, 64, 2)):
        print(f"chunks {name} w {w} mode {mode} tile {tile} order {order} rows {h} split {sp}", flush=True)
        tf = _dev(fa)
        plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype)
        plan.set_tuning(tile, 512, 128, 0, mode, column_chunks=w, stream_order=order, row_chunks=h, tail_split=sp)
        t = plan.tuning
        assert t["column_chunks"] == w and t["stream_order"] == order and t["row_chunks"] == h
        assert t["tail_split"] == sp
        plan.transform()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(tf.cpu().numpy(), want)
        # back to the pieces in their own order, whole work items
        tf.copy_(_dev(fa))
        plan.set_tuning(tile, 256, 64, 1, mode, column_chunks=0)
        t = plan.tuning
        assert t["column_chunks"] == 0 and t["stream_order"] == 0 and t["row_chunks"] == 0 and t["tail_split"] == 0
        plan.transform()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(tf.cpu().numpy(), want)
        plan.close()
    for bad in ({"column_chunks": -1}, {"stream_order": 2}, {"stream_order": -1}, {"row_chunks": -1},
                {"tail_split": 3}, {"tail_split": -2}, {"tail_split": 16}):
        with pytest.raises(g2g.G2GError):
            plan = g2g.Plan(init, ti.data_ptr(), fin, ti.data_ptr(), dtype)
            try:
                plan.set_tuning(65536, 256, 64, 0, 0, **bad)
            finally:
                plan.close()


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("name", ["bcT", "arbT", "bcT_mixed_paths", "bcT_short_tiles", "bcT_mid_tiles", "bcT_cut_edges"])
def test_column_chunked_transposed(name, dtype):
    """Transposed pieces cut into chunks of final columns (source offsets
    along the storage rows), in union and in destination order: whole LDS
    tiles, strips and the element queue give the oracle's bytes for chunk
    widths that split tiles and groups."""
    init, fin = _transposed_pairs()[name]
    rng = np.random.default_rng(41)
    ia, fa = _rand_arena(init, dtype, rng), _rand_arena(fin, dtype, rng)
    trans = "C" if dtype in (g2g.COMPLEX64, g2g.COMPLEX128) else "T"
    want = _expected(init, ia, fin, fa, dtype, trans)
    ti = _dev(ia)
    q = 256 // g2g.DTYPE_SIZE[dtype]
    for w, mode, tile, order in ((q, 0, 65536, 0), (3, 1, 4096, 1), (2 * q, 2, 1 << 20, 0), (40, 3, 65536, 0),
                                 (0, 1, 49152, 1), (q, 3, 40960, 1)):
        print(f"transposed chunks {name} dtype {dtype} w {w} mode {mode} order {order}", flush=True)
        tf = _dev(fa)
        plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans)
        plan.set_tuning(tile, 512, 128, 1, mode, column_chunks=w, stream_order=order)
        assert plan.tuning["column_chunks"] == w and plan.tuning["stream_order"] == order
        with pytest.raises(g2g.G2GError):  # row chunks: untransposed plans only
            plan.set_tuning(tile, 512, 128, 1, mode, column_chunks=w, stream_order=order, row_chunks=16)
        assert plan.tuning["column_chunks"] == w and plan.tuning["stream_order"] == order
        plan.transform()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(tf.cpu().numpy(), want)
        plan.close()


def test_tall_pieces_autotune_row_chunks_and_shape_class():
    """Tall pieces (4 KiB columns: 512-row f64 blocks, and one block per side):
    the autotuner's pick, which may cut them into row bands, gives the oracle's
    bytes; a plan of the same shape created afterwards starts from the same
    stream (column chunks, row chunks, order) and work items (size, tail
    split) and gives them too."""
    dtype = g2g.FLOAT64
    rng = np.random.default_rng(61)
    for init, fin in ((scalapack_layout(2048, 96, 512, 32, 2, 1, "C"), scalapack_layout(2048, 96, 1024, 48, 2, 2, "C")),
                      (scalapack_layout(1536, 40, 1536, 40, 1, 1, "C"), scalapack_layout(1536, 40, 1536, 40, 1, 1, "C"))):
        ia, fa = _rand_arena(init, dtype, rng), _rand_arena(fin, dtype, rng)
        want = _expected(init, ia, fin, fa, dtype)
        ti, tf = _dev(ia), _dev(fa)
        plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype)
        best = plan.autotune(reps=1)
        print(f"tall pieces {init.m}x{init.n}: {best}", flush=True)
        tf.copy_(_dev(fa))
        plan.transform()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(tf.cpu().numpy(), want)
        # a new plan of the shape class starts from the recorded stream
        ti2, tf2 = _dev(ia), _dev(fa)
        plan2 = g2g.Plan(init, ti2.data_ptr(), fin, tf2.data_ptr(), dtype)
        t2 = plan2.tuning
        for k in ("column_chunks", "row_chunks", "stream_order", "tile_bytes", "tail_split"):
            assert t2[k] == best[k], (k, t2, best)
        plan2.transform()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(tf2.cpu().numpy(), want)
        # and every row-chunk height the tuner does not try, on the same plan
        for h in (1, 7, 300):
            tf2.copy_(_dev(fa))
            plan2.set_tuning(8192, 256, 128, 1, 1, column_chunks=5, stream_order=h % 2, row_chunks=h)
            plan2.transform()
            torch.cuda.synchronize()
            np.testing.assert_array_equal(tf2.cpu().numpy(), want, err_msg=f"row chunks {h}")
        plan2.close()
        plan.close()


def test_column_chunked_batch_and_autotune():
    """A batch of chunked and unchunked plans, in union and destination
    order, gives every request's own result; the autotuner's choice (chunked
    or not, either order) too."""
    dtype = g2g.FLOAT64
    rng = np.random.default_rng(23)
    cases = []
    for k, name in enumerate(["bc_small", "bc_grid_change", "arbitrary", "tall_thin"]):
        init, fin = _layout_pairs()[name]
        ia, fa = _rand_arena(init, dtype, rng), _rand_arena(fin, dtype, rng)
        ti, tf = _dev(ia), _dev(fa)
        plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype)
        if k % 2 == 0:
            plan.set_tuning(65536, 256, 64, 0, 1, column_chunks=2 + k, stream_order=k // 2)
        elif k == 3:
            plan.set_tuning(40960, 512, 128, 0, 1, stream_order=1)
        elif k == 1:
            plan.set_tuning(16384, 512, 64, 0, 1, column_chunks=4, row_chunks=6)
        cases.append((plan, ti, tf, fa, _expected(init, ia, fin, fa, dtype)))
    batch = g2g.Batch([c[0] for c in cases])
    batch.execute()
    torch.cuda.synchronize()
    for k, (plan, ti, tf, fa, want) in enumerate(cases):
        np.testing.assert_array_equal(tf.cpu().numpy(), want, err_msg=f"batched request {k}")
    batch.close()
    for k, (plan, ti, tf, fa, want) in enumerate(cases):
        tf.copy_(_dev(fa))
        best = plan.autotune(reps=1)
        print(f"autotuned request {k}: {best}", flush=True)
        tf.copy_(_dev(fa))
        plan.transform()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(tf.cpu().numpy(), want, err_msg=f"autotuned request {k}")
        plan.close()


@pytest.mark.parametrize("shape", [(0, 40), (30, 0), (0, 0)])
def test_empty_matrix_is_a_no_op(shape):
    """An empty dimension plans to no pieces; transforms (fused, staged,
    transposed), volume and counts succeed and write nothing."""
    m, n = shape
    init = scalapack_layout(m, n, 4, 3, 2, 2, "C")
    fin = scalapack_layout(m, n, 5, 7, 2, 2, "C")
    dtype = g2g.FLOAT64
    rng = np.random.default_rng(5)
    ia = _rand_arena(init, dtype, rng) if init.arena_elems else np.zeros(8, dtype=np.uint8)
    fa = rng.integers(0, 256, size=64, dtype=np.uint8)
    ti, tf = _dev(ia), _dev(fa)
    plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype)
    assert plan.info.n_pieces == 0 and plan.info.bytes == 0
    plan.transform()
    plan.transform_staged()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(tf.cpu().numpy(), fa)
    assert not np.any(plan.counts())
    plan.close()
    initT = scalapack_layout(n, m, 4, 3, 2, 2, "C")
    planT = g2g.Plan(initT, ti.data_ptr(), fin, tf.data_ptr(), dtype, "T")
    planT.transform()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(tf.cpu().numpy(), fa)
    planT.close()
    assert g2g.comm_volume(init, fin) == {}


def test_random_requests_both_forms():
    """A sweep of random block-cyclic requests (sizes 8..300, blocks 16..256,
    grids up to 4x4, every type, 'N'/'T'/'C', some scaled), fused and staged,
    each against the oracle on random bytes."""
    from grid2grid_amd import workloads
    rng = np.random.default_rng(2024)
    for k in range(40):
        init, fin, dtype, trans = workloads.random_request(rng, dtypes=DTYPES, lo=8, hi=300, p_trans=0.5)
        if trans == "T" and dtype in (g2g.COMPLEX64, g2g.COMPLEX128) and rng.random() < 0.5:
            trans = "C"
        staged = bool(k % 2)
        alpha = beta = None
        if k % 5 == 0:
            alpha, beta = SCALARS[dtype]
            es = g2g.DTYPE_SIZE[dtype]
            ia = np.zeros(init.arena_elems * es, dtype=np.uint8)
            fa = np.zeros(fin.arena_elems * es, dtype=np.uint8)
            O.fill(init, ia, dtype, 100 + k)
            O.fill(fin, fa, dtype, 200 + k)
        else:
            ia, fa = _rand_arena(init, dtype, rng), _rand_arena(fin, dtype, rng)
        print(f"case {k}: {init.m}x{init.n} dtype {dtype} trans {trans} staged {staged} "
              f"scaled {alpha is not None}", flush=True)
        want = _expected(init, ia, fin, fa, dtype, trans, alpha, beta)
        got = _run(init, ia, fin, fa, dtype, trans, alpha, beta, staged=staged)
        np.testing.assert_array_equal(got, want, err_msg=f"case {k}")


def test_random_requests_random_streams():
    """Random requests (block-cyclic and irregular layouts, every type,
    'N'/'T'/'C') under random tunings: work-item sizes that are not powers of
    two, both workgroup sizes and unrolls, every launch mode, column chunks,
    row chunks (untransposed), union or destination order, each transform
    against the oracle on random bytes; then the same plans in one batch.
    The host replay runs the same cases (tests/test_gather_replay.py)."""
    rng = np.random.default_rng(4343)
    cases = []
    for k, (init, fin, dtype, trans, tu) in enumerate(shapes.random_stream_cases()):
        tile, th, ub, nt, mode, w, order, h, sp = tu
        ia, fa = _rand_arena(init, dtype, rng), _rand_arena(fin, dtype, rng)
        print(f"case {k}: {init.m}x{init.n} dtype {dtype} trans {trans} tuning {tu}", flush=True)
        want = _expected(init, ia, fin, fa, dtype, trans)
        ti, tf = _dev(ia), _dev(fa)
        plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans)
        plan.set_tuning(tile, th, ub, nt, mode, column_chunks=w, stream_order=order, row_chunks=h, tail_split=sp)
        plan.transform()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(tf.cpu().numpy(), want, err_msg=f"case {k}")
        cases.append((plan, ti, tf, fa, want))
    for plan, ti, tf, fa, want in cases:
        tf.copy_(_dev(fa))
    batch = g2g.Batch([c[0] for c in cases])
    batch.execute()
    torch.cuda.synchronize()
    for k, (plan, ti, tf, fa, want) in enumerate(cases):
        np.testing.assert_array_equal(tf.cpu().numpy(), want, err_msg=f"batched case {k}")
    batch.close()
    for c in cases:
        c[0].close()


def test_beyond_32_bit_element_indices():
    """65536 x 40000 f32 (2.6e9 elements, 10.5 GB per side): element offsets,
    piece addresses and the seeded values' indices all pass 2^31 and 2^32.
    Checked on the device element by element against v(seed, i, j)."""
    m, n = 65536, 40000
    init = scalapack_layout(m, n, 4096, 2500, 4, 2, "C", lld="rows")
    fin = scalapack_layout(m, n, 8192, 5000, 2, 4, "R", lld="rows")
    dtype = g2g.FLOAT32
    es = g2g.DTYPE_SIZE[dtype]
    assert m * n > 2**31 and init.arena_elems * es > 2**33
    ti = torch.empty(init.arena_elems * es, dtype=torch.uint8, device="cuda")
    tf = torch.zeros(fin.arena_elems * es, dtype=torch.uint8, device="cuda")
    g2g.fill_layout(init, ti.data_ptr(), dtype, 31)
    plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype)
    plan.transform()
    torch.cuda.synchronize()
    assert g2g.verify_layout(fin, tf.data_ptr(), dtype, 31, "N", n) == 0
    # the check sees a single flipped element (row 5 of the first block)
    tf.view(torch.int32)[int(fin.blk_offset[0]) + 5] ^= 1
    assert g2g.verify_layout(fin, tf.data_ptr(), dtype, 31, "N", n) == 1
    plan.close()
    del ti, tf
    torch.cuda.empty_cache()


def test_stream_probe():
    """The streaming reference passes: the copy copies, the write writes its
    pattern, the read leaves the source alone; bad arguments are refused."""
    n = 64 << 20
    src = torch.randint(0, 256, (n,), dtype=torch.uint8, device="cuda")
    keep = src.clone()
    dst = torch.zeros(n, dtype=torch.uint8, device="cuda")
    assert g2g.stream_probe(0, src.data_ptr(), dst.data_ptr(), n, reps=2) > 0
    torch.cuda.synchronize()
    assert torch.equal(dst, src)
    assert g2g.stream_probe(1, src.data_ptr(), dst.data_ptr(), n, reps=2) > 0
    assert g2g.stream_probe(2, src.data_ptr(), dst.data_ptr(), n, reps=2) > 0
    torch.cuda.synchronize()
    assert torch.equal(src, keep)
    words = dst.view(torch.int32).view(-1, 4)[:1024].cpu().numpy()
    assert (words[:, 0] == 0x3ff00000).all() and (words[:, 1] == 0).all()
    for bad in ((3, n), (0, n + 4), (0, 0)):
        with pytest.raises(g2g.G2GError):
            g2g.stream_probe(bad[0], src.data_ptr(), dst.data_ptr(), bad[1])


# (tile_bytes or None for the plan's own, threads, unroll_bytes, nontemporal,
# launch_mode, column_chunks, stream_order, row_chunks, tail_split): the
# pieces, plain column chunks, and the configurations the autotuner picks for
# these workloads (destination order, tail split)
FULL_SIZE_CASES = [("bc4096", (None, 512, 64, 1, 1, 0, 0, 0, 0)), ("bc4096", (32768, 256, 128, 1, 0, 64, 0, 0, 2)),
                   ("bc32768", (None, 512, 64, 1, 1, 0, 0, 0, 0)), ("bc32768", (None, 512, 64, 1, 1, 16, 0, 0, 0)),
                   ("bc32768", (65536, 512, 64, 1, 1, 32, 1, 0, 0)),
                   ("irregular", (None, 512, 64, 1, 1, 0, 0, 0, 0)), ("irregular", (None, 512, 64, 1, 1, 8, 0, 0, 0)),
                   ("irregular", (40960, 512, 128, 1, 1, 2, 1, 0, 0))]


@pytest.mark.parametrize("name,tu", FULL_SIZE_CASES)
def test_benchmark_workloads_full_size(name, tu):
    """The benchmark's own layout pairs at full size: 4096^2 f64, 32768^2 f64
    (8 GiB a side) and the irregular 40000x25000 f32 with padded, shuffled
    blocks, on the pieces, on column-chunked streams and in the configurations
    the autotuner picks for them. Every element of the final layout equals the
    seeded value of its source element (checked on the device), and no byte
    outside the final blocks (ld padding, gaps) is touched."""
    from grid2grid_amd import workloads
    init, fin, dtype, trans = getattr(workloads, name)()
    es = g2g.DTYPE_SIZE[dtype]
    ti = torch.empty(init.arena_elems * es, dtype=torch.uint8, device="cuda")
    tf = torch.full((fin.arena_elems * es,), 0xA5, dtype=torch.uint8, device="cuda")
    g2g.fill_layout(init, ti.data_ptr(), dtype, 11)
    plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans)
    tile, th, ub, nt, mode, w, order, h, sp = tu
    plan.set_tuning(tile or plan.tuning["tile_bytes"], th, ub, nt, mode, column_chunks=w, stream_order=order,
                    row_chunks=h, tail_split=sp)
    print(f"{name} tuning {tu}: {plan.info.n_pieces} pieces, {plan.info.n_tiles} work items", flush=True)
    plan.transform()
    torch.cuda.synchronize()
    assert g2g.verify_layout(fin, tf.data_ptr(), dtype, 11, trans, init.n) == 0
    mask = fin.element_mask()
    if not mask.all():
        inside = torch.from_numpy(mask).to("cuda")
        changed = (tf.view(-1, es) != 0xA5).any(dim=1)
        assert not bool((changed & ~inside).any()), "a byte outside the final blocks was written"
    plan.close()
    del ti, tf
    torch.cuda.empty_cache()
