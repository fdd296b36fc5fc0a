"""GPU: many transforms of mixed types and operations in one persistent
launch (batches, hipGraph replay) and through the serving loop; every result
must equal the oracle's result for that transform alone."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
import shapes
import grid2grid_amd as g2g
from grid2grid_amd import _lib, workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _requests(n, seed, dtypes=(g2g.FLOAT32, g2g.FLOAT64, g2g.COMPLEX64, g2g.COMPLEX128), hi=200):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        init, fin, dtype, trans = workloads.random_request(rng, dtypes=dtypes, lo=8, hi=hi, p_trans=0.4)
        if trans == "T" and dtype in (g2g.COMPLEX64, g2g.COMPLEX128) and rng.random() < 0.5:
            trans = "C"
        out.append((init, fin, dtype, trans))
    return out


class Case:
    """Host arenas (random bytes), device copies, expected result and plan."""

    def __init__(self, init, fin, dtype, trans, rng, scaled=False):
        es = g2g.DTYPE_SIZE[dtype]
        self.dtype, self.trans = dtype, trans
        if scaled:
            self.ia = np.zeros(init.arena_elems * es, dtype=np.uint8)
            self.fa = np.zeros(fin.arena_elems * es, dtype=np.uint8)
            O.fill(init, self.ia, dtype, int(rng.integers(1, 1 << 30)))
            O.fill(fin, self.fa, dtype, int(rng.integers(1, 1 << 30)))
            self.alpha, self.beta = {g2g.FLOAT32: (np.float32(1.5), np.float32(-0.5)),
                                     g2g.FLOAT64: (0.25, 3.0),
                                     g2g.COMPLEX64: (np.complex64(1 - 1j), np.complex64(0.5j)),
                                     g2g.COMPLEX128: (2 + 1j, -1 + 0.5j)}[dtype]
        else:
            self.ia = rng.integers(0, 256, size=init.arena_elems * es, dtype=np.uint8)
            self.fa = rng.integers(0, 256, size=fin.arena_elems * es, dtype=np.uint8)
            self.alpha = self.beta = None
        self.want = self.fa.copy()
        O.transform(init, self.ia.copy(), fin, self.want, dtype, trans, self.alpha, self.beta)
        self.ti = torch.from_numpy(self.ia).to("cuda")
        self.tf = torch.from_numpy(self.fa).to("cuda")
        self.plan = g2g.Plan(init, self.ti.data_ptr(), fin, self.tf.data_ptr(), dtype, trans)

    def reset(self):
        self.tf.copy_(torch.from_numpy(self.fa).to("cuda"))

    def check(self, what):
        np.testing.assert_array_equal(self.tf.cpu().numpy(), self.want, err_msg=what)


def _cases(n, seed, scaled_every=0):
    rng = np.random.default_rng(seed)
    reqs = _requests(n, seed)
    return [Case(i, f, d, t, rng, scaled=bool(scaled_every) and k % scaled_every == 0)
            for k, (i, f, d, t) in enumerate(reqs)]


def test_batch_mixed_equals_alone():
    cases = _cases(24, 1, scaled_every=5)
    batch = g2g.Batch([c.plan for c in cases], [c.alpha for c in cases], [c.beta for c in cases])
    assert batch.bytes == sum(c.plan.info.bytes for c in cases)
    batch.execute()
    torch.cuda.synchronize()
    for k, c in enumerate(cases):
        c.check(f"request {k} dtype {c.dtype} trans {c.trans}")
    batch.close()


@pytest.mark.parametrize("launch_mode", [1, 3], ids=["one_queue", "queue_per_xcd"])
def test_batch_launch_modes_and_autotune(launch_mode):
    """Both persistent launch modes, and whatever the batch autotuner picks,
    give every request's own result."""
    cases = _cases(20, 4, scaled_every=6)
    g2g.set_tuning(0, launch_mode)
    try:
        batch = g2g.Batch([c.plan for c in cases], [c.alpha for c in cases], [c.beta for c in cases])
    finally:
        g2g.set_tuning(0, -1)
    batch.execute()
    torch.cuda.synchronize()
    for k, c in enumerate(cases):
        c.check(f"mode {launch_mode} request {k}")
        c.reset()
    best = batch.autotune(reps=1)
    assert best["launch_mode"] in (1, 3)
    for c in cases:
        c.reset()
    torch.cuda.synchronize()
    batch.execute()
    torch.cuda.synchronize()
    for k, c in enumerate(cases):
        c.check(f"autotuned {best} request {k}")
    batch.close()


def test_batch_set_tuning_every_variant():
    """Every settable batch configuration (work-item size, workgroup size,
    bytes in flight, non-temporal hint, both persistent modes) gives every
    request's own result; settings outside the compiled variants, and stream
    fields other than 0, are refused."""
    cases = _cases(12, 9, scaled_every=4)
    batch = g2g.Batch([c.plan for c in cases], [c.alpha for c in cases], [c.beta for c in cases])
    for tb, th, ub, nt, mode in [(4096, 256, 64, 0, 1), (32768, 512, 128, 1, 3), (1 << 20, 512, 64, 0, 3),
                                 (65536, 256, 128, 1, 1)]:
        batch.set_tuning(tb, th, ub, nt, mode)
        t = batch.tuning
        assert (t["tile_bytes"], t["threads"], t["unroll_bytes"], t["nontemporal"], t["launch_mode"]) == \
            (tb, th, ub, nt, mode)
        for c in cases:
            c.reset()
        torch.cuda.synchronize()
        batch.execute()
        torch.cuda.synchronize()
        for k, c in enumerate(cases):
            c.check(f"tuning {t} request {k}")
    for bad in [dict(tile_bytes=1024), dict(tile_bytes=32768, threads=128), dict(tile_bytes=32768, launch_mode=0),
                dict(tile_bytes=32768, unroll_bytes=96)]:
        with pytest.raises(g2g.G2GError):
            batch.set_tuning(**bad)
    # the stream fields of a batch's own tuning must be 0 (each plan runs its
    # own stream, cut into the batch's work items)
    for k, field in enumerate(("column_chunks", "stream_order", "row_chunks", "tail_split")):
        t = _lib.G2GTuning(65536, 512, 128, 0, 1, 0, 0, 0, 0, 0.0)
        setattr(t, field, 2)
        assert _lib.load().g2g_batch_set_tuning(batch.handle, ctypes.byref(t), None) != 0, field
    batch.close()


def test_batch_graph_replays():
    """The captured graph gives the eager result on every replay, with eager
    work and synchronizations between the replays."""
    cases = _cases(16, 2)
    stream = torch.cuda.Stream()
    batch = g2g.Batch([c.plan for c in cases], stream=stream)
    for rep in range(4):
        for c in cases:
            c.reset()
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            batch.execute_graph(stream)
        stream.synchronize()
        for k, c in enumerate(cases):
            c.check(f"replay {rep} request {k}")
        # unrelated eager work between replays
        x = torch.randn(1 << 20, device="cuda")
        assert torch.isfinite(x.sum())
    batch.close()


def test_server_batched_results_equal_alone():
    """Requests submitted while others run are admitted as budget frees up;
    each result is identical to the transform run alone."""
    cases = _cases(40, 3, scaled_every=7)
    budget = max(2 * c.plan.info.bytes for c in cases) * 3  # forces waiting
    srv = g2g.Server(budget, n_streams=3, max_batch=8)
    tickets = [srv.submit(c.plan, c.alpha, c.beta) for c in cases]
    lat = [srv.wait(t) for t in tickets]
    srv.drain()
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(cases) and st["rejected"] == 0
    assert st["peak_in_flight_bytes"] <= budget
    assert all(x >= 0 for x in lat)
    for k, c in enumerate(cases):
        c.check(f"served request {k}")


def test_server_tuned_plans():
    """Plans on tuned streams (column and row chunks, destination order, tail
    splits, any work-item size), served batched: each result is the
    oracle's; the scaled ones run over the plans' pieces."""
    rng = np.random.default_rng(29)
    cases = []
    for k, (init, fin, dtype, trans, tu) in enumerate(shapes.random_stream_cases(seed=6161, count=24)):
        c = Case(init, fin, dtype, trans, rng, scaled=k % 6 == 0)
        tile, th, ub, nt, mode, w, order, h, sp = tu
        c.plan.set_tuning(tile, th, ub, nt, mode, column_chunks=w, stream_order=order, row_chunks=h, tail_split=sp)
        cases.append(c)
    srv = g2g.Server(1 << 40, n_streams=2, max_batch=6)
    tickets = [srv.submit(c.plan, c.alpha, c.beta) for c in cases]
    for t in tickets:
        srv.wait(t)
    srv.drain()
    srv.close()
    for k, c in enumerate(cases):
        c.check(f"served tuned request {k}")


def test_server_reports_each_ticket_once():
    """poll and wait report a finished ticket once, then forget it: asking
    again is an error, and an unknown ticket is refused."""
    cases = _cases(3, 6)
    srv = g2g.Server(1 << 40, n_streams=2, max_batch=4)
    t0 = srv.submit(cases[0].plan)
    t1 = srv.submit(cases[1].plan)
    assert srv.wait(t0) >= 0
    with pytest.raises(g2g.G2GError) as e:
        srv.wait(t0)
    assert e.value.code == 6
    srv.drain()
    assert srv.poll(t1) is not None
    with pytest.raises(g2g.G2GError) as e:
        srv.poll(t1)
    assert e.value.code == 6
    with pytest.raises(g2g.G2GError) as e:
        srv.poll(12345)
    assert e.value.code == 1
    srv.close()
    for k, c in enumerate(cases[:2]):
        c.check(f"request {k}")


def test_server_plain_mode_and_budget_reject():
    cases = _cases(6, 4)
    srv = g2g.Server(1 << 40, mode=g2g.SERVE_PLAIN)
    for c in cases:
        srv.wait(srv.submit(c.plan))
    for k, c in enumerate(cases):
        c.check(f"plain request {k}")
    srv.close()
    small = g2g.Server(16)
    with pytest.raises(g2g.G2GError) as e:
        small.submit(cases[0].plan)
    assert e.value.code == 4
    assert small.stats()["rejected"] == 1
    small.close()


def test_batch_refuses_after_a_plan_is_retuned():
    """A batch holds its plans' piece tables by address: once a plan is
    retuned to another stream (column chunks, stream order), launching the
    old batch fails cleanly, and a batch built again gives every request's
    own result."""
    cases = _cases(6, 9)
    untransposed = [c for c in cases if c.trans == "N"]
    assert untransposed
    batch = g2g.Batch([c.plan for c in cases])
    batch.execute()
    torch.cuda.synchronize()
    t = untransposed[0].plan.tuning
    untransposed[0].plan.set_tuning(t["tile_bytes"], t["threads"], t["unroll_bytes"], t["nontemporal"],
                                    t["launch_mode"], column_chunks=3)
    with pytest.raises(g2g.G2GError):
        batch.execute()
    s = torch.cuda.Stream()
    with pytest.raises(g2g.G2GError):
        batch.execute_graph(s)
    batch.close()
    for c in cases:
        c.reset()
    batch = g2g.Batch([c.plan for c in cases])
    batch.execute()
    torch.cuda.synchronize()
    for k, c in enumerate(cases):
        c.check(f"rebuilt batch request {k}")
    # the stream order is part of the stream too
    t = cases[-1].plan.tuning
    cases[-1].plan.set_tuning(t["tile_bytes"], t["threads"], t["unroll_bytes"], t["nontemporal"], t["launch_mode"],
                              column_chunks=t["column_chunks"], stream_order=1 - t["stream_order"])
    with pytest.raises(g2g.G2GError):
        batch.execute()
    batch.close()
    for c in cases:
        c.reset()
    batch = g2g.Batch([c.plan for c in cases])
    batch.execute()
    torch.cuda.synchronize()
    for k, c in enumerate(cases):
        c.check(f"batch rebuilt after the order change, request {k}")
    batch.close()


def test_batch_of_the_benchmark_distribution():
    """2048 requests drawn from the benchmark's mixed distribution at full
    size (m, n up to 4096, f32/f64/c64, blocks 16..256, grids up to 4x4, 30%
    transposed) in one hipGraph-replayed launch: after each of two replays
    every request's final layout equals the seeded closed form."""
    reqs = workloads.mixed_requests(n=2048, seed=77)
    rs = []
    for k, (init, fin, dtype, trans) in enumerate(reqs):
        es = g2g.DTYPE_SIZE[dtype]
        ti = torch.empty(init.arena_elems * es, dtype=torch.uint8, device="cuda")
        tf = torch.zeros(fin.arena_elems * es, dtype=torch.uint8, device="cuda")
        g2g.fill_layout(init, ti.data_ptr(), dtype, 300 + k)
        rs.append((init, fin, dtype, trans, ti, tf, g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans)))
    s = torch.cuda.Stream()
    batch = g2g.Batch([r[6] for r in rs], stream=s)
    for rep in range(2):
        for r in rs:
            r[5].zero_()
        torch.cuda.synchronize()
        batch.execute_graph(s)
        s.synchronize()
        bad = [k for k, (init, fin, dtype, trans, ti, tf, plan) in enumerate(rs)
               if g2g.verify_layout(fin, tf.data_ptr(), dtype, 300 + k, trans, init.n) != 0]
        assert not bad, f"replay {rep}: requests {bad[:10]} differ"
    batch.close()
    for r in rs:
        r[6].close()


def test_served_stream_of_the_benchmark_mix():
    """The served workload: a Poisson stream (2000 requests/s) of 300
    requests of the mixed distribution and 16 large 16384^2 f64 requests
    (128x128 on 4x4 -> 256x256 on 2x8; 5% of the stream) through the
    continuous-batching server under the 64 GB HBM budget. Every result
    equals the seeded closed form, the admitted footprint never exceeds the
    budget, and a sample of requests run again alone gives the same bytes as
    in the mix."""
    import time
    rng = np.random.default_rng(2005)
    small = [workloads.random_request(rng) for _ in range(300)]
    reqs = small[:]
    for k in range(16):  # the large requests at spread positions
        reqs.insert(int(k * len(reqs) / 16) + k, workloads.large_request())
    assert sum(1 for r in reqs if r[0].m == 16384) == 16
    rs = []
    for k, (init, fin, dtype, trans) in enumerate(reqs):
        es = g2g.DTYPE_SIZE[dtype]
        ti = torch.empty(init.arena_elems * es, dtype=torch.uint8, device="cuda")
        tf = torch.zeros(fin.arena_elems * es, dtype=torch.uint8, device="cuda")
        g2g.fill_layout(init, ti.data_ptr(), dtype, 500 + k)
        rs.append((init, fin, dtype, trans, ti, tf, g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans)))
    torch.cuda.synchronize()
    budget = 64 * 10**9
    srv = g2g.Server(budget, n_streams=3, max_batch=512)
    gaps = rng.exponential(1.0 / 2000.0, size=len(rs))
    tickets = []
    t_next = time.perf_counter()
    for k, r in enumerate(rs):
        t_next += gaps[k]
        while time.perf_counter() < t_next:
            pass
        tickets.append(srv.submit(r[6]))
    lat = [srv.wait(t) for t in tickets]
    srv.drain()
    st = srv.stats()
    srv.close()
    print(f"served {st['completed']} requests in {st['launches']} launches, peak in flight "
          f"{st['peak_in_flight_bytes'] / 1e9:.1f} GB, p50 {np.percentile(lat, 50) * 1e3:.2f} ms, "
          f"p99 {np.percentile(lat, 99) * 1e3:.2f} ms", flush=True)
    assert st["completed"] == len(rs) and st["rejected"] == 0
    assert st["peak_in_flight_bytes"] <= budget
    assert st["launches"] < len(rs)  # requests were batched together
    bad = [k for k, (init, fin, dtype, trans, ti, tf, plan) in enumerate(rs)
           if g2g.verify_layout(fin, tf.data_ptr(), dtype, 500 + k, trans, init.n) != 0]
    assert not bad, f"requests {bad[:10]} differ"
    # a sample again, alone: the same bytes as in the mix
    sample = [k for k in range(len(rs)) if rs[k][0].m == 16384][:3] + list(range(0, len(rs), 37))
    for k in sample:
        init, fin, dtype, trans, ti, tf, plan = rs[k]
        in_mix = tf.clone()
        tf.zero_()
        plan.transform()
        torch.cuda.synchronize()
        assert torch.equal(tf, in_mix), f"request {k}: alone differs from the mix"
        del in_mix
    for r in rs:
        r[6].close()
    del rs
    torch.cuda.empty_cache()
