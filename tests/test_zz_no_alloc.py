"""GPU: g2g_transform_plans_noalloc -- several plans in one launch with the
launch's descriptors, work records and counters in the library's static
device array -- gives the bytes of the existing entry point it varies
(g2g_batch_create + g2g_batch_execute; for plans on tuned streams, each
plan's g2g_transform) and of the oracle, on two calls in a row, the second
with fewer and smaller transforms; requests beyond the array's capacity are
refused before anything runs."""
import numpy as np
import pytest

import oracle_lib as O
import shapes
import grid2grid_amd as g2g
from grid2grid_amd import workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _cases(n, seed, hi):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        init, fin, dtype, trans = workloads.random_request(
            rng, dtypes=(g2g.FLOAT32, g2g.FLOAT64, g2g.COMPLEX64, g2g.COMPLEX128), lo=8, hi=hi, p_trans=0.4)
        if trans == "T" and dtype in (g2g.COMPLEX64, g2g.COMPLEX128) and k % 2:
            trans = "C"
        es = g2g.DTYPE_SIZE[dtype]
        scaled = k % 5 == 0
        if scaled:
            ia = O.special_arena(dtype, init.arena_elems, rng)
            fa = O.special_arena(dtype, fin.arena_elems, rng)
            alpha, beta = O.special_scalar(dtype, rng), O.special_scalar(dtype, rng)
        else:
            ia = rng.integers(0, 256, size=init.arena_elems * es, dtype=np.uint8)
            fa = rng.integers(0, 256, size=fin.arena_elems * es, dtype=np.uint8)
            alpha = beta = None
        want = fa.copy()
        O.transform(init, ia.copy(), fin, want, dtype, trans, alpha, beta)
        ti, tf = torch.from_numpy(ia).cuda(), torch.from_numpy(fa).cuda()
        plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans)
        out.append(dict(plan=plan, ti=ti, tf=tf, fa=fa, want=want, alpha=alpha, beta=beta))
    return out


def _reset(cases):
    for c in cases:
        c["tf"].copy_(torch.from_numpy(c["fa"]).cuda())
    torch.cuda.synchronize()


def _call(cases, stream):
    print(f"noalloc call: {len(cases)} plans", flush=True)
    g2g.transform_plans_noalloc([c["plan"] for c in cases], [c["alpha"] for c in cases],
                                [c["beta"] for c in cases], stream=stream)


def test_noalloc_matches_batch_two_calls():
    s = torch.cuda.Stream()
    first = _cases(40, 11, 400)
    second = _cases(12, 12, 60)  # fewer, smaller transforms: the array keeps the first call's contents
    for cases in (first, second):
        _reset(cases)
        batch = g2g.Batch([c["plan"] for c in cases], [c["alpha"] for c in cases], [c["beta"] for c in cases])
        batch.execute()
        torch.cuda.synchronize()
        via_batch = [c["tf"].cpu().numpy() for c in cases]
        batch.close()
        _reset(cases)
        with torch.cuda.stream(s):
            _call(cases, s)
        s.synchronize()
        for k, c in enumerate(cases):
            got = c["tf"].cpu().numpy()
            np.testing.assert_array_equal(got, via_batch[k], err_msg=f"request {k}: noalloc vs batch")
            np.testing.assert_array_equal(got, c["want"], err_msg=f"request {k}: noalloc vs oracle")
    for c in first + second:
        c["plan"].close()


def test_noalloc_refuses_beyond_capacity():
    """4097 plans, or more work items than the array holds, fail cleanly
    with nothing written."""
    cases = _cases(1, 13, 64)
    with pytest.raises(g2g.G2GError) as e:
        g2g.transform_plans_noalloc([cases[0]["plan"]] * 4097)
    assert e.value.code == 1 and "4096" in str(e.value)
    # 9000^2 f32 in 4096-byte work items: 79 102 work items
    from grid2grid_amd.layout import scalapack_layout
    init = scalapack_layout(9000, 9000, 1000, 1000, 2, 2, "C", lld="rows")
    fin = scalapack_layout(9000, 9000, 3000, 3000, 1, 3, "C", lld="rows")
    ti = torch.zeros(init.arena_elems * 4, dtype=torch.uint8, device="cuda")
    tf = torch.full((fin.arena_elems * 4,), 7, dtype=torch.uint8, device="cuda")
    g2g.set_tuning(4096, -1)
    try:
        big = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), g2g.FLOAT32)
    finally:
        g2g.set_tuning(0, -1)
    assert big.info.n_tiles > 73722
    with pytest.raises(g2g.G2GError) as e:
        g2g.transform_plans_noalloc([big])
    assert e.value.code == 1 and "work items" in str(e.value)
    torch.cuda.synchronize()
    assert bool((tf == 7).all())
    big.close()
    cases[0]["plan"].close()


def test_noalloc_tuned_streams_two_calls():
    """Plans on tuned streams (column chunks, row chunks, destination order,
    work-item sizes that are not powers of two) run through the allocation-free
    launch give the bytes of their own g2g_transform and of the oracle; the
    second call takes fewer plans with smaller inputs."""
    rng = np.random.default_rng(77)
    all_cases = []
    for k, (init, fin, dtype, trans, tu) in enumerate(shapes.random_stream_cases(seed=5151, count=18)):
        tile, th, ub, nt, mode, w, order, h, sp = tu
        es = g2g.DTYPE_SIZE[dtype]
        ia = rng.integers(0, 256, size=init.arena_elems * es, dtype=np.uint8)
        fa = rng.integers(0, 256, size=fin.arena_elems * es, dtype=np.uint8)
        want = fa.copy()
        O.transform(init, ia.copy(), fin, want, dtype, trans, None, None)
        ti, tf = torch.from_numpy(ia).cuda(), torch.from_numpy(fa).cuda()
        plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans)
        plan.set_tuning(tile, th, ub, nt, mode, column_chunks=w, stream_order=order, row_chunks=h, tail_split=sp)
        all_cases.append(dict(plan=plan, ti=ti, tf=tf, fa=fa, want=want, alpha=None, beta=None,
                              size=init.m * init.n * es, tuning=tu))
    by_size = sorted(all_cases, key=lambda c: c["size"])
    first = all_cases
    second = by_size[:6]  # fewer and the smallest: the array keeps the first call's contents
    s = torch.cuda.Stream()
    for cases in (first, second):
        _reset(cases)
        for c in cases:
            c["plan"].transform()
        torch.cuda.synchronize()
        own = [c["tf"].cpu().numpy() for c in cases]
        _reset(cases)
        with torch.cuda.stream(s):
            _call(cases, s)
        s.synchronize()
        for k, c in enumerate(cases):
            got = c["tf"].cpu().numpy()
            np.testing.assert_array_equal(got, own[k], err_msg=f"request {k} {c['tuning']}: noalloc vs transform")
            np.testing.assert_array_equal(got, c["want"], err_msg=f"request {k} {c['tuning']}: noalloc vs oracle")
    for c in all_cases:
        c["plan"].close()
