"""GPU: the drop-in call planned once per layout pair (the plan cache behind
g2g_transform_layouts), a plan's launch replayed from its hipGraph, and one
plan (and one batch) launched on several streams and from several host
threads at once: every result equals the oracle's."""
import threading

import numpy as np
import pytest

import oracle_lib as O
import grid2grid_amd as g2g
from grid2grid_amd.layout import scalapack_layout, arbitrary_layout

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _case(seed, dtype=g2g.FLOAT64, trans="N"):
    rng = np.random.default_rng(seed)
    init = arbitrary_layout(rng, 300, 260, 7, 6, 5, lo=10, hi=80, pad=(0, 9), gap=(0, 33))
    if trans == "N":
        fin = arbitrary_layout(rng, 300, 260, 5, 8, 6, lo=10, hi=80, pad=(0, 9), gap=(0, 33))
    else:
        fin = arbitrary_layout(rng, 260, 300, 5, 8, 6, lo=10, hi=80, pad=(0, 9), gap=(0, 33))
    es = g2g.DTYPE_SIZE[dtype]
    ia = rng.integers(0, 256, size=init.arena_elems * es, dtype=np.uint8)
    fa = rng.integers(0, 256, size=fin.arena_elems * es, dtype=np.uint8)
    want = fa.copy()
    O.transform(init, ia.copy(), fin, want, dtype, trans)
    return init, fin, ia, fa, want


def test_transform_layouts_plans_once():
    """The first call of a layout pair plans it, later calls with the same
    descriptors are cache hits; another pair, another dtype or another
    operation is a new plan; every result equals the oracle's."""
    g2g.plan_cache_clear()
    init, fin, ia, fa, want = _case(1)
    ti, tf = torch.from_numpy(ia).cuda(), torch.from_numpy(fa).cuda()
    for k in range(3):
        tf.copy_(torch.from_numpy(fa).cuda())
        g2g.transform_layouts(init, ti.data_ptr(), fin, tf.data_ptr(), g2g.FLOAT64)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(tf.cpu().numpy(), want)
        assert g2g.plan_cache_stats() == (k, 1, 1)
    # the same layouts at other addresses are another pair
    ti2, tf2 = torch.from_numpy(ia).cuda(), torch.from_numpy(fa).cuda()
    g2g.transform_layouts(init, ti2.data_ptr(), fin, tf2.data_ptr(), g2g.FLOAT64)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(tf2.cpu().numpy(), want)
    assert g2g.plan_cache_stats() == (2, 2, 2)
    # capacity 1: the least recently used plan leaves, results stay exact
    g2g.plan_cache_set_capacity(1)
    assert g2g.plan_cache_stats()[2] == 1
    tf.copy_(torch.from_numpy(fa).cuda())
    g2g.transform_layouts(init, ti.data_ptr(), fin, tf.data_ptr(), g2g.FLOAT64)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(tf.cpu().numpy(), want)
    g2g.plan_cache_set_capacity(64)
    g2g.plan_cache_clear()
    assert g2g.plan_cache_stats() == (0, 0, 0)


def test_transform_graph_replays():
    """A plan's graph gives the eager result on every replay, with eager work
    in between; changing alpha/beta captures again."""
    init, fin, ia, fa, want = _case(2, g2g.COMPLEX64, "C")
    ti, tf = torch.from_numpy(ia).cuda(), torch.from_numpy(fa).cuda()
    plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), g2g.COMPLEX64, "C")
    s = torch.cuda.Stream()
    for rep in range(3):
        tf.copy_(torch.from_numpy(fa).cuda())
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            plan.transform_graph(s)
        s.synchronize()
        np.testing.assert_array_equal(tf.cpu().numpy(), want, err_msg=f"replay {rep}")
        x = torch.randn(1 << 18, device="cuda")
        assert torch.isfinite(x.sum())
    alpha, beta = np.complex64(0.5 - 2j), np.complex64(1.5 + 0.25j)
    want2 = fa.copy()
    O.transform(init, ia.copy(), fin, want2, g2g.COMPLEX64, "C", alpha, beta)
    for rep in range(2):
        tf.copy_(torch.from_numpy(fa).cuda())
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            plan.transform_graph(s, alpha, beta)
        s.synchronize()
        np.testing.assert_array_equal(tf.cpu().numpy(), want2, err_msg=f"scaled replay {rep}")
    plan.close()


@pytest.mark.parametrize("launch_mode", [1, 3], ids=["persistent", "persistent_xcd"])
def test_one_plan_on_several_streams(launch_mode):
    """Persistent-mode launches on 4 streams at once: 4 plans of the same
    layout pair (one per stream, each with its own destination) and one more
    plan launched on all 4 streams (each launch of it takes the work counters
    of its stream); every destination equals the oracle's."""
    init, fin, ia, fa, want = _case(3)
    ti = torch.from_numpy(ia).cuda()
    g2g.set_tuning(4096, launch_mode)
    try:
        shared_tf = torch.from_numpy(fa).cuda()
        shared = g2g.Plan(init, ti.data_ptr(), fin, shared_tf.data_ptr(), g2g.FLOAT64)
        own = []
        for k in range(4):
            tf = torch.from_numpy(fa).cuda()
            own.append((g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), g2g.FLOAT64), tf))
    finally:
        g2g.set_tuning(0, -1)
    assert shared.tuning["launch_mode"] == launch_mode and shared.info.n_tiles > 64
    streams = [torch.cuda.Stream() for _ in range(4)]
    torch.cuda.synchronize()
    for rep in range(5):
        for k, s in enumerate(streams):
            with torch.cuda.stream(s):
                own[k][0].transform(stream=s)
                shared.transform(stream=s)  # the same plan on 4 streams at once
    torch.cuda.synchronize()
    np.testing.assert_array_equal(shared_tf.cpu().numpy(), want)
    for k, (p, tf) in enumerate(own):
        np.testing.assert_array_equal(tf.cpu().numpy(), want, err_msg=f"stream {k}")
        p.close()
    shared.close()


def test_threads_share_cached_plans():
    """Several host threads call the drop-in entry point for the same layout
    pair on their own streams (one plan from the cache, launched on every
    stream at once); each thread's destination equals the oracle's."""
    g2g.plan_cache_clear()
    init, fin, ia, fa, want = _case(4)
    ti = torch.from_numpy(ia).cuda()
    dests = [torch.from_numpy(fa).cuda() for _ in range(4)]
    errors = []
    torch.cuda.synchronize()
    g2g.set_tuning(4096, 1)

    def work(k):
        try:
            s = torch.cuda.Stream()
            for _ in range(4):
                g2g.transform_layouts(init, ti.data_ptr(), fin, dests[k].data_ptr(), g2g.FLOAT64, stream=s)
            s.synchronize()
        except Exception as e:  # reported below
            errors.append(e)

    try:
        th = [threading.Thread(target=work, args=(k,)) for k in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    finally:
        g2g.set_tuning(0, -1)
    assert not errors, errors
    torch.cuda.synchronize()
    for k, tf in enumerate(dests):
        np.testing.assert_array_equal(tf.cpu().numpy(), want, err_msg=f"thread {k}")
    hits, misses, entries = g2g.plan_cache_stats()
    assert hits + misses == 16 and entries == 4
    g2g.plan_cache_clear()
