#!/usr/bin/env python3
"""Benchmark of the MI355X grid2grid transform.

Default workload: the 32768^2 fp64 block-cyclic redistribution (128x128
blocks on a 4x4 rank grid -> 256x512 blocks on 2x8, 16 virtual ranks, 8 GiB),
run by the fused gather kernel. One step = one full transform of the matrix.
value = moved GB/s = 2*m*n*sizeof(T) / device time per step (one read and one
write per element; inputs resident in HBM). Output: one JSON line.

Beside it, in the same JSON line: the 4096^2 f64 and irregular 40000x25000
f32 legs (each with its own roofline), the drop-in call timed with the plan
cache on and off, the batched and served legs, and the reference CPU
pipeline on the same 32768^2 workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload bc32768]
                       [--mode fused|staged] [--no-cpu-baseline] [--no-extras]
       python bench.py --only-leg bc4096|irregular [--tuning T]   (one leg alone)
       python bench.py --only-batched [--batch-graph-only]       (the batched leg alone)
       python bench.py --sweep                                    (2048^2..32768^2, one line per size)
"""
import argparse
import json
import os
import re
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "moved GB/s vs HBM peak (bit-exact); p50/p99 latency, served-batched vs plain"


# single-workload legs beside the default workload: buffer pairs rotated, steps, drop-in calls timed
LEGS = {"bc4096": (8, 400, 80), "irregular": (1, 20, 0)}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10,
                    help="untimed steps first: back-to-back launches start slower for about 15 ms, then settle")
    ap.add_argument("--workload", default="bc32768",
                    help="bc4096 | bc32768 | irregular | copy | bc:m:n:imb:inb:ipr:ipc:fmb:fnb:fpr:fpc:dtype:trans")
    ap.add_argument("--no-autotune", action="store_true", help="skip the autotuner (default configuration)")
    ap.add_argument("--tuning", default=None,
                    help="explicit configuration tile_bytes:threads:unroll_bytes:nontemporal:launch_mode"
                         "[:column_chunks[:stream_order[:row_chunks[:tail_split]]]] (skips the autotuner)")
    ap.add_argument("--mode", default="fused", choices=["fused", "staged"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=6.0, help="budget of the CPU baseline's 8192^2 sample")
    ap.add_argument("--batch-requests", type=int, default=2048)
    ap.add_argument("--batch-ptrans", type=float, default=0.3, help="share of transposed batch requests")
    ap.add_argument("--batch-tile-bytes", type=int, default=0, help="work-item size of the batch plans (0: auto)")
    ap.add_argument("--only-batched", action="store_true", help="run only the batched leg")
    ap.add_argument("--sweep", action="store_true",
                    help="only the size sweep: the default workload's blocks and grids at 2048^2..32768^2 f64, "
                         "buffer pairs rotated past the Infinity Cache; one JSON line per size")
    ap.add_argument("--only-leg", default=None, choices=sorted(LEGS),
                    help="run only this single-workload leg (with its buffer rotation; --tuning applies)")
    ap.add_argument("--batch-graph-only", action="store_true",
                    help="batched leg without the per-request comparisons (for counter runs)")
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--qps", type=float, default=2000.0)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(name):
    from grid2grid_amd import workloads
    if name == "copy":
        # one 32768^2 f64 block on each side: a plain contiguous 8 GiB copy
        # (reference point for the HBM byte counters)
        from grid2grid_amd.layout import scalapack_layout
        lay = scalapack_layout(32768, 32768, 32768, 32768, 1, 1, "C", lld="rows")
        return lay, scalapack_layout(32768, 32768, 32768, 32768, 1, 1, "C", lld="rows"), 1, "N"
    if name.startswith("bc:"):
        # bc:m:n:imb:inb:ipr:ipc:fmb:fnb:fpr:fpc:dtype:trans -- any block-cyclic pair
        f = name.split(":")[1:]
        m, n, imb, inb, ipr, ipc, fmb, fnb, fpr, fpc = (int(x) for x in f[:10])
        from grid2grid_amd import DTYPE_NAMES
        from grid2grid_amd.layout import scalapack_layout
        dtype, trans = DTYPE_NAMES[f[10]], f[11]
        fm, fn = (n, m) if trans in "TC" else (m, n)
        return (scalapack_layout(m, n, imb, inb, ipr, ipc, "C", lld="rows"),
                scalapack_layout(fm, fn, fmb, fnb, fpr, fpc, "C", lld="rows"), dtype, trans)
    return {"bc4096": workloads.bc4096, "bc32768": workloads.bc32768, "irregular": workloads.irregular}[name]()


WORKLOAD_DESC = {
    "bc4096": "block-cyclic 4096^2 f64, 32x32 on 2x2 -> 64x128 on 2x2, 4 virtual ranks",
    "bc32768": "block-cyclic 32768^2 f64, 128x128 on 4x4 -> 256x512 on 2x8, 16 virtual ranks",
    "irregular": "irregular 40000x25000 f32, 64x48 -> 50x40 splits, 32 virtual ranks, padded ld",
    "copy": "single-block 32768^2 f64 -> single block: plain contiguous 8 GiB copy (calibration)",
}


def time_steps(torch, fn, steps, warmup, stream, per_launch=False, warm_s=0.0):
    """Average ms per step over `steps` back-to-back steps (events only at the
    two ends of the timed region), after `warmup` untimed ones, and, with
    warm_s, more untimed ones until the warm-up has lasted warm_s seconds (a
    burst of back-to-back launches starts slower for about 15 ms). With
    per_launch, a second, separate pass brackets every step with its own event
    pair on the launch stream and returns those durations too (their spread;
    each also holds the dispatch latency between the first event and the
    kernel's start)."""
    t0 = time.perf_counter()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    spent = time.perf_counter() - t0
    if warmup and spent < warm_s:
        for _ in range(int((warm_s - spent) / (spent / warmup)) + 1):
            fn()
        torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / steps
    per = []
    if per_launch:
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for a, b in evs:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        per = [a.elapsed_time(b) for a, b in evs]
    return ms, per


def read_traffic(workload, mode):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", f"gather_{workload}_pmc.txt" if mode == "fused" else
                        f"staged_{workload}_pmc.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9]+)", open(path).read())
    return int(m.group(1)) if m else None


def roofline(nbytes, per_launch_ms, workload, mode="fused"):
    """per_launch_ms: the average launch duration over the timed region (one
    launch per step, events at the region's two ends)."""
    achieved = nbytes / (per_launch_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(workload, mode),
            "algorithmic_bytes": int(nbytes)}


def launch_stats(per):
    return {"mean": round(statistics.mean(per), 4), "median": round(statistics.median(per), 4),
            "min": round(min(per), 4), "max": round(max(per), 4), "launches": len(per)}


def workload_leg(torch, g2g, workloads, name, stream, dev, steps, warmup, rotate=1, facade_calls=0, tuning=None):
    """One workload measured on its own: `rotate` buffer pairs, each with its
    own plan, run round-robin (one step = one transform of one pair), so that
    a working set below the 256 MB Infinity Cache is still read from HBM.
    The first plan is autotuned (or set to `tuning`, a tile:threads:unroll:
    nontemporal:mode[:chunks[:order[:rows[:tail_split]]]] string); the others start from the tuning it
    records for their shape. With facade_calls, the drop-in call
    (transform_layouts) is timed on the same pairs too."""
    init, fin, dtype, trans = build_workload(name)
    es = g2g.DTYPE_SIZE[dtype]
    pairs = []
    t0 = time.time()
    for k in range(rotate):
        ti = torch.empty(init.arena_elems * es, dtype=torch.uint8, device=dev)
        tf = torch.zeros(fin.arena_elems * es, dtype=torch.uint8, device=dev)
        g2g.fill_layout(init, ti.data_ptr(), dtype, 100 + k, stream)
        plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans, stream)
        if tuning is not None:
            f = [int(x) for x in tuning.split(":")]
            plan.set_tuning(*f[:5], column_chunks=f[5] if len(f) > 5 else 0,
                            stream_order=f[6] if len(f) > 6 else 0, row_chunks=f[7] if len(f) > 7 else 0,
                            tail_split=f[8] if len(f) > 8 else 0)
        elif k == 0:
            plan.autotune(stream=stream, reps=3)
        pairs.append((ti, tf, plan))
    tuning = pairs[0][2].tuning
    later = pairs[-1][2].tuning
    log(f"leg {name}: {rotate} buffer pair(s), planned and tuned in {time.time() - t0:.1f}s: {tuning}")
    it = [0]

    def step():
        pairs[it[0] % rotate][2].transform(stream=stream)
        it[0] += 1

    n = max(steps, rotate)
    n -= n % rotate  # every pair the same number of times
    ms, per = time_steps(torch, step, n, max(warmup, rotate), stream, per_launch=True, warm_s=0.03)
    bad = sum(g2g.verify_layout(fin, tf.data_ptr(), dtype, 100 + k, trans, init.n, stream)
              for k, (_, tf, _) in enumerate(pairs))
    nbytes = workloads.algorithmic_bytes(fin, dtype)
    info = pairs[0][2].info
    out = {
        "workload": WORKLOAD_DESC.get(name, name),
        "dtype": {g2g.FLOAT32: "f32", g2g.FLOAT64: "f64", g2g.COMPLEX64: "c64", g2g.COMPLEX128: "c128"}[dtype],
        "buffer_pairs": rotate,
        "working_set_bytes": int(rotate * (init.arena_elems + fin.arena_elems) * es),
        "steps": n,
        "ms_per_step": round(ms, 4),
        "value": round(nbytes / (ms * 1e-3) / 1e9, 2),
        "unit": "GB/s",
        "bit_exact": bad == 0,
        "mismatches": int(bad),
        "pieces": int(info.n_pieces),
        "work_items": int(info.n_tiles),
        "tuning": {k: tuning[k] for k in ("tile_bytes", "threads", "unroll_bytes", "nontemporal", "launch_mode",
                                          "column_chunks", "stream_order", "row_chunks", "tail_split")},
        "later_pairs_start_tuned": all(later[k] == tuning[k] for k in ("tile_bytes", "threads", "unroll_bytes",
                                                                         "nontemporal", "launch_mode",
                                                                         "stream_order", "tail_split")),
        "launch_ms": launch_stats(per),
        "roofline": roofline(nbytes, ms, name),
    }
    log(f"leg {name}: {ms:.4f} ms/step, {out['value']} GB/s, frac {out['roofline']['frac']}, mismatches {bad}")
    if facade_calls:
        out["facade"] = facade_leg(torch, g2g, name, init, fin, dtype, trans, pairs,
                                   [100 + k for k in range(rotate)], stream, facade_calls)
    for _, _, plan in pairs:
        plan.close()
    del pairs
    torch.cuda.empty_cache()
    return out


def facade_leg(torch, g2g, name, init, fin, dtype, trans, pairs, seeds, stream, calls):
    """The drop-in call g2g_transform_layouts (what the C++ facade's
    transform() runs), one call per step over the rotating buffer pairs,
    each call waited for as the reference's transform() returns after its
    exchange: with the plan cache off (every call plans, launches and frees
    its plan, as the reference plans on every call) and on (every call after
    the first per pair replays the cached plan). Host wall time per call."""
    def call(k):
        ti, tf, _ = pairs[k % len(pairs)]
        g2g.transform_layouts(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans, stream=stream)
        stream.synchronize()

    def timed(n):
        ts = []
        for k in range(n):
            t0 = time.perf_counter()
            call(k)
            ts.append((time.perf_counter() - t0) * 1e3)
        return ts

    out = {}
    for tf in (p[1] for p in pairs):
        tf.zero_()
    g2g.plan_cache_clear()
    g2g.plan_cache_set_capacity(0)
    try:
        call(0)  # first-call costs of the process (module load, pools) outside the sample
        unc = timed(calls)
    finally:
        g2g.plan_cache_set_capacity(64)
    g2g.plan_cache_clear()
    for k in range(len(pairs)):
        call(k)
    hits0, misses0, _ = g2g.plan_cache_stats()
    cac = timed(calls)
    hits, misses, entries = g2g.plan_cache_stats()
    bad = sum(g2g.verify_layout(fin, p[1].data_ptr(), dtype, seed, trans, init.n, stream)
              for p, seed in zip(pairs, seeds))
    g2g.plan_cache_clear()
    for key, ts in (("uncached", unc), ("cached", cac)):
        out[key] = {"ms_per_call_median": round(statistics.median(ts), 4), "ms_per_call_mean": round(statistics.mean(ts), 4),
                    "calls": len(ts)}
    out["cache_hits_in_sample"] = int(hits - hits0)
    out["cache_misses_in_sample"] = int(misses - misses0)
    out["uncached_over_cached"] = round(out["uncached"]["ms_per_call_median"] / out["cached"]["ms_per_call_median"], 2)
    out["bit_exact"] = bad == 0
    log(f"facade {name}: {out}")
    return out


def _cpu_sample(O, m, warmups, runs, budget_s, kind):
    """Median seconds of the CPU transform of the default workload's blocks and
    grids at size m x m: `warmups` untimed runs, then timed runs until both
    `runs` runs and `budget_s` seconds are done (at most 200)."""
    import numpy as np
    from grid2grid_amd import workloads
    init, fin, dtype, _ = workloads.bc32768(m=m)
    es = O.ES[dtype]
    ia = np.zeros(init.arena_elems * es, dtype=np.uint8)
    fa = np.zeros(fin.arena_elems * es, dtype=np.uint8)
    ia[::4096] = 1
    times = []
    t_end = time.time() + budget_s
    n = 0
    while True:
        t0 = time.perf_counter()
        if kind == "reference":
            O.ref_transform([(init, ia, fin, fa)], dtype)
        else:
            O.transform(init, ia, fin, fa, dtype)
        dt = time.perf_counter() - t0
        n += 1
        if n > warmups:
            times.append(dt)
        if (time.time() > t_end and len(times) >= runs) or len(times) >= 200:
            break
    del ia, fa
    return statistics.median(times), len(times), 2 * m * m * es


def cpu_baseline(budget_s):
    """The reference's own pipeline (oracle/_ref, compiled from the reference
    sources) on the default workload itself, 32768^2 f64 (128x128 on 4x4 ->
    256x512 on 2x8, 16 virtual ranks in one process, OpenMP over all host
    threads; 16 GiB of host arrays): median of 5 runs after one warm-up; and,
    as a second sample, 8192^2 with the same blocks and grids. Without the
    reference build, the C restatement (single thread) on the 8192^2 sample
    only."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    kind = "reference" if O.ref() is not None else "port"
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)) if kind == "reference" else 1
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            cpu_model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "")
    except OSError:
        pass
    host = {"cpu_model": cpu_model, "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}
    med_s, n_s, bytes_s = _cpu_sample(O, 8192, 2, 3, budget_s, kind)
    small = {"value": round(bytes_s / med_s / 1e9, 3), "unit": "GB/s",
             "sample": f"8192^2 f64, same blocks/grids, median of {n_s} runs after 2 warm-ups",
             "ms_per_transform": round(med_s * 1e3, 3)}
    if kind != "reference":
        return {"value": small["value"], "unit": "GB/s", "cores": 1, "kind": kind,
                "sample": small["sample"] + " (C restatement, one thread); time includes planning",
                "ms_per_transform": small["ms_per_transform"], "host": host}
    med, n, nbytes = _cpu_sample(O, 32768, 1, 5, 0.0, kind)
    return {
        "value": round(nbytes / med / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": kind,
        "sample": f"the full workload: 32768^2 f64, 128x128 on 4x4 -> 256x512 on 2x8, 16 virtual ranks in one "
                  f"process, median of {n} runs after 1 warm-up, {threads} OpenMP threads; time includes the "
                  f"reference's per-call planning",
        "ms_per_transform": round(med * 1e3, 3),
        "second_sample": small,
        "host": host,
    }


def _percentile(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs), q)) if xs else None


def make_requests(torch, g2g, reqs, seed0, dev):
    """Resident buffers and plans for a list of (init, final, dtype, trans)."""
    out = []
    for k, (init, fin, dtype, trans) in enumerate(reqs):
        es = g2g.DTYPE_SIZE[dtype]
        ti = torch.empty(init.arena_elems * es, dtype=torch.uint8, device=dev)
        tf = torch.zeros(fin.arena_elems * es, dtype=torch.uint8, device=dev)
        g2g.fill_layout(init, ti.data_ptr(), dtype, seed0 + k)
        plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans)
        out.append(dict(init=init, fin=fin, dtype=dtype, trans=trans, ti=ti, tf=tf, plan=plan, seed=seed0 + k,
                        bytes=2 * fin.m * fin.n * es))
    return out


def verify_requests(g2g, rs):
    bad = 0
    for r in rs:
        bad += g2g.verify_layout(r["fin"], r["tf"].data_ptr(), r["dtype"], r["seed"], r["trans"], r["init"].n)
    return bad


def batched_leg(torch, g2g, workloads, n_req, steps, stream, dev, p_trans=0.3, tile_bytes=0, per_request=True):
    """Mixed requests: n_req transforms in one persistent launch (hipGraph
    replay) versus plain per-request launches (the staged form per request)
    and fused per-request launches."""
    t0 = time.time()
    if tile_bytes:
        g2g.set_tuning(tile_bytes, -1)
    rs = make_requests(torch, g2g, workloads.mixed_requests(n=n_req, p_trans=p_trans), 40000, dev)
    g2g.set_tuning(0, -1)
    nbytes = sum(r["bytes"] for r in rs)
    log(f"mixed requests: {n_req} requests, {nbytes / 1e9:.2f} GB moved per pass, planned in {time.time() - t0:.1f}s")
    batch = g2g.Batch([r["plan"] for r in rs], stream=stream)
    t0 = time.time()
    btune = batch.autotune(stream, reps=2)
    log(f"batch autotune ({time.time() - t0:.1f}s): {btune}")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ms_b, _ = time_steps(torch, lambda: batch.execute_graph(s), steps, 2, s)
    bad_b = verify_requests(g2g, rs)
    if not per_request:
        batch.close()
        out = {"bytes_per_pass": int(nbytes),
               "batched_graph": {"ms": round(ms_b, 3), "GB/s": round(nbytes / ms_b / 1e6, 1), "bit_exact": bad_b == 0},
               "batch_tuning": btune}
        for r in rs:
            r["plan"].close()
        return out
    for r in rs:
        r["tf"].zero_()
    ms_f, _ = time_steps(torch, lambda: [r["plan"].transform(stream=stream) for r in rs], max(2, steps // 2), 1,
                         stream)
    bad_f = verify_requests(g2g, rs)
    for r in rs:
        r["tf"].zero_()
    for r in rs:  # allocate the staging buffers outside the timed region
        r["plan"].transform_staged(stream=stream)
    ms_p, _ = time_steps(torch, lambda: [r["plan"].transform_staged(stream=stream) for r in rs],
                         max(2, steps // 2), 1, stream)
    bad_p = verify_requests(g2g, rs)
    batch.close()
    out = {
        "workload": f"{n_req} mixed transforms (m,n log-uniform [64,4096], f32/f64/c64, blocks 16..256, "
                    f"grids 1x1..4x4, 30% transposed)",
        "bytes_per_pass": int(nbytes),
        "batched_graph": {"ms": round(ms_b, 3), "GB/s": round(nbytes / ms_b / 1e6, 1), "bit_exact": bad_b == 0},
        "plain_per_request": {"ms": round(ms_p, 3), "GB/s": round(nbytes / ms_p / 1e6, 1), "bit_exact": bad_p == 0},
        "fused_per_request": {"ms": round(ms_f, 3), "GB/s": round(nbytes / ms_f / 1e6, 1), "bit_exact": bad_f == 0},
        "speedup_vs_plain": round(ms_p / ms_b, 2),
        "speedup_vs_fused_per_request": round(ms_f / ms_b, 2),
        "batch_tuning": {k: btune[k] for k in ("tile_bytes", "threads", "unroll_bytes", "nontemporal")},
    }
    for r in rs:
        r["plan"].close()
    del rs
    torch.cuda.empty_cache()
    return out


def served_leg(torch, g2g, workloads, n_small, n_large, n_req, qps, budget, dev):
    """Served stream: Poisson arrivals at `qps` of mixed requests plus 5% large
    16384^2 f64 requests, served with continuous batching under an HBM budget,
    versus plain per-request launches; then the saturated throughput of both."""
    import numpy as np
    rng = np.random.default_rng(1005)
    t0 = time.time()
    small = make_requests(torch, g2g, [workloads.random_request(rng) for _ in range(n_small)], 50000, dev)
    large = make_requests(torch, g2g, [workloads.large_request() for _ in range(n_large)], 60000, dev)
    log(f"served stream: pool of {n_small} small + {n_large} large requests ready in {time.time() - t0:.1f}s")
    for r in small + large:  # staging buffers of the plain form, outside the timed region
        r["plan"].transform_staged()
    torch.cuda.synchronize()
    # arrival sequence: 5% large
    kinds = rng.random(n_req) < 0.05
    gaps = rng.exponential(1.0 / qps, size=n_req)

    def run(server, open_loop):
        busy = {}  # pool index -> ticket in flight
        pools = {True: list(range(len(large))), False: list(range(len(small)))}
        nxt = {True: 0, False: 0}
        lat_of = {}  # ticket -> latency, once the server reported it (each ticket is reported once)
        done_bytes = 0
        t_start = time.perf_counter()
        t_next = t_start
        tickets = []

        def finished(t):
            if t not in lat_of:
                x = server.poll(t)
                if x is None:
                    return False
                lat_of[t] = x
            return True

        for i in range(n_req):
            big = bool(kinds[i])
            pool = large if big else small
            # next pool entry not in flight
            for _ in range(len(pool)):
                j = pools[big][nxt[big] % len(pool)]
                nxt[big] += 1
                key = (big, j)
                t = busy.get(key)
                if t is None or finished(t):
                    break
            else:
                if busy[key] not in lat_of:
                    lat_of[busy[key]] = server.wait(busy[key])
            if open_loop:
                t_next += gaps[i]
                while time.perf_counter() < t_next:
                    pass
            t = server.submit(pool[j]["plan"])
            busy[key] = t
            tickets.append((t, pool[j]["bytes"]))
        for t, b in tickets:
            if t not in lat_of:
                lat_of[t] = server.wait(t)
            done_bytes += b
        wall = time.perf_counter() - t_start
        lat = [lat_of[t] for t, _ in tickets]
        return lat, done_bytes, wall

    res = {}
    for name, mode, streams, mb in (("served_batched", g2g.SERVE_BATCHED, 3, 512), ("plain_per_request", g2g.SERVE_PLAIN, 1, 1)):
        srv = g2g.Server(budget, n_streams=streams, max_batch=mb, mode=mode)
        lat, nbytes, wall = run(srv, True)
        st = srv.stats()
        srv.close()
        srv = g2g.Server(budget, n_streams=streams, max_batch=mb, mode=mode)
        _, sat_bytes, sat_wall = run(srv, False)
        srv.close()
        res[name] = {
            "p50_ms": round(_percentile(lat, 50) * 1e3, 3),
            "p99_ms": round(_percentile(lat, 99) * 1e3, 3),
            "sustained_GB/s_at_qps": round(nbytes / wall / 1e9, 1),
            "saturated_GB/s": round(sat_bytes / sat_wall / 1e9, 1),
            "launches": int(st["launches"]),
            "peak_in_flight_GB": round(st["peak_in_flight_bytes"] / 1e9, 2),
        }
        log(f"{name}: {res[name]}")
    bad = verify_requests(g2g, small + large)
    out = {
        "workload": f"Poisson {qps} req/s, {n_req} requests (mixed-request distribution + 5% 16384^2 f64), "
                    f"HBM budget {budget / 1e9:.0f} GB",
        **res,
        "throughput_ratio_saturated": round(res["served_batched"]["saturated_GB/s"] /
                                            max(1e-9, res["plain_per_request"]["saturated_GB/s"]), 2),
        "bit_exact": bad == 0,
    }
    for r in small + large:
        r["plan"].close()
    torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU: this build maps the P ranks of a layout to virtual ranks "
              f"on one MI355X; --gpus {args.gpus} is not supported.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import grid2grid_amd as g2g
    from grid2grid_amd import workloads

    g2g.load()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()

    if args.only_batched:
        out = batched_leg(torch, g2g, workloads, args.batch_requests, args.steps if args.batch_graph_only else 5,
                          stream, dev, args.batch_ptrans, args.batch_tile_bytes,
                          per_request=not args.batch_graph_only)
        print(json.dumps({"batched": out}), flush=True)
        return

    if args.sweep:
        # the default workload's geometry (128x128 blocks on 4x4 -> 256x512 on
        # 2x8, f64) at growing sizes; every size uses enough buffer pairs in
        # turn that its working set is at least 2 GiB
        for m in (2048, 4096, 8192, 16384, 32768):
            pair_bytes = 2 * m * m * 8
            rotate = max(1, -(-(2 << 30) // pair_bytes))
            steps = max(rotate, min(400, int(50e-3 * 6e12 / pair_bytes)))
            out = workload_leg(torch, g2g, workloads, f"bc:{m}:{m}:128:128:4:4:256:512:2:8:f64:N", stream, dev,
                               steps, args.warmup, rotate=rotate)
            print(json.dumps({"sweep_m": m, **out}), flush=True)
        return

    if args.only_leg:
        rotate, steps, _ = LEGS[args.only_leg]
        out = workload_leg(torch, g2g, workloads, args.only_leg, stream, dev, max(args.steps, rotate), args.warmup,
                           rotate=rotate, tuning=args.tuning)
        print(json.dumps({"leg": args.only_leg, **out}), flush=True)
        return

    init, fin, dtype, trans = build_workload(args.workload)
    es = g2g.DTYPE_SIZE[dtype]
    log(f"workload {args.workload}: {WORKLOAD_DESC.get(args.workload, args.workload)}")
    ti = torch.empty(init.arena_elems * es, dtype=torch.uint8, device=dev)
    tf = torch.zeros(fin.arena_elems * es, dtype=torch.uint8, device=dev)
    g2g.fill_layout(init, ti.data_ptr(), dtype, 2, stream)
    t0 = time.time()
    plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans, stream)
    info = plan.info
    log(f"plan: {info.n_pieces} pieces, {info.n_tiles} work items of {info.tile_bytes} B, "
        f"planned in {time.time() - t0:.2f}s")
    tuning = plan.tuning
    if args.tuning:
        f = [int(x) for x in args.tuning.split(":")]
        plan.set_tuning(*f[:5], column_chunks=f[5] if len(f) > 5 else 0,
                            stream_order=f[6] if len(f) > 6 else 0, row_chunks=f[7] if len(f) > 7 else 0,
                            tail_split=f[8] if len(f) > 8 else 0)
        tuning = plan.tuning
        info = plan.info
    elif args.mode == "fused" and not args.no_autotune:
        t0 = time.time()
        tuning = plan.autotune(stream=stream, reps=3)
        info = plan.info
        log(f"autotune ({time.time() - t0:.1f}s): {tuning}")

    run = {"fused": lambda: plan.transform(stream=stream),
           "staged": lambda: plan.transform_staged(stream=stream)}[args.mode]
    ms, per = time_steps(torch, run, args.steps, args.warmup, stream, per_launch=True)
    a_n_cols = init.n
    bad = g2g.verify_layout(fin, tf.data_ptr(), dtype, 2, trans, a_n_cols, stream)
    nbytes = workloads.algorithmic_bytes(fin, dtype)
    gbs = nbytes / (ms * 1e-3) / 1e9
    achieved = gbs  # one launch per step: the average launch duration over the timed region is ms
    log(f"{args.mode}: {ms:.3f} ms/step, {gbs:.1f} GB/s, mismatches {bad}")

    result = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {g2g.FLOAT32: "f32", g2g.FLOAT64: "f64", g2g.COMPLEX64: "c64", g2g.COMPLEX128: "c128"}[dtype],
        "data": "synthetic (seeded splitmix64 values in [-1,1))",
        "config": {
            "workload": WORKLOAD_DESC.get(args.workload, args.workload),
            "mode": args.mode,
            "m": fin.m, "n": fin.n,
            "virtual_ranks": int(max(init.n_ranks, fin.n_ranks)),
            "pieces": int(info.n_pieces),
            "work_items": int(info.n_tiles),
            "tile_bytes": int(info.tile_bytes),
            "kernel_variant": {k: tuning[k] for k in ("threads", "unroll_bytes", "nontemporal", "launch_mode",
                                                      "column_chunks", "stream_order", "row_chunks",
                                                      "tail_split")},
            "parallelism": "1 GPU, P virtual ranks",
        },
        "bit_exact": bad == 0,
        "mismatches": int(bad),
        # every launch of a second pass timed on its own (an event pair on the
        # launch stream around each): the spread of the launch durations.
        # roofline.achieved uses the timed region's average (events at its two
        # ends): an event pair around one launch adds the dispatch latency,
        # about 2-3 us, which is 5 % of a 50 us launch
        "launch_ms": launch_stats(per),
        "roofline": roofline(nbytes, ms, args.workload, args.mode),
    }

    # streaming reference rates of this device on the same buffers (16-byte
    # vectors, 16 workgroups per CU): context for the roofline fraction
    nst = (min(ti.numel(), tf.numel()) // 16) * 16
    st_ms = {k: g2g.stream_probe(k, ti.data_ptr(), tf.data_ptr(), nst, reps=10, stream=stream) for k in (0, 1, 2)}
    result["device_stream"] = {
        "bytes": int(nst),
        "copy_GB/s": round(2 * nst / (st_ms[0] * 1e-3) / 1e9, 1),
        "read_GB/s": round(nst / (st_ms[1] * 1e-3) / 1e9, 1),
        "write_GB/s": round(nst / (st_ms[2] * 1e-3) / 1e9, 1),
    }
    result["device_stream"]["gather_vs_read"] = round(achieved / result["device_stream"]["read_GB/s"], 3)
    log(f"device streaming rates: {result['device_stream']}")

    if not args.no_extras and args.mode == "fused":
        # plain two-stage form on the same plan, for comparison
        tf.zero_()
        ms_s, _ = time_steps(torch, lambda: plan.transform_staged(stream=stream), max(3, args.steps // 4), 1,
                             stream)
        bad_s = g2g.verify_layout(fin, tf.data_ptr(), dtype, 2, trans, a_n_cols, stream)
        result["plain_staged"] = {"ms_per_step": round(ms_s, 4), "value": round(nbytes / (ms_s * 1e-3) / 1e9, 2),
                                  "unit": "GB/s", "bit_exact": bad_s == 0}
        log(f"staged: {ms_s:.3f} ms/step, mismatches {bad_s}")
    if not args.no_extras and args.mode == "fused":
        tf.zero_()
        result["facade"] = facade_leg(torch, g2g, args.workload, init, fin, dtype, trans, [(ti, tf, plan)], [2],
                                      stream, 10)
    plan.close()
    del ti, tf
    torch.cuda.empty_cache()

    if not args.no_extras:
        # the other single-transform workloads, each with its own roofline;
        # 4096^2 (2 x 128 MiB) rotates 8 buffer pairs (2 GiB) past the 256 MB Infinity Cache
        result["legs"] = {}
        for name, (rotate, steps, calls) in LEGS.items():
            if name != args.workload:
                result["legs"][name] = workload_leg(torch, g2g, workloads, name, stream, dev, steps, args.warmup,
                                                    rotate=rotate, facade_calls=calls)
        result["batched"] = batched_leg(torch, g2g, workloads, args.batch_requests, 5, stream, dev,
                                        args.batch_ptrans, args.batch_tile_bytes)
        log(f"batched: {result['batched']}")
        result["served"] = served_leg(torch, g2g, workloads, 512, 6, args.serve_requests, args.qps,
                                      int(args.budget_gb * 1e9), dev)

    if not args.no_cpu_baseline:
        try:
            result["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
            log(f"cpu baseline: {result['cpu_baseline']}")
        except Exception as e:  # the baseline is reported, never the target
            result["cpu_baseline"] = {"value": None, "error": str(e)}
        cb, fa = result["cpu_baseline"], result.get("facade")
        if cb.get("value") and cb.get("kind") == "reference" and fa and args.workload == "bc32768":
            # the same basis on both sides: one call that plans, moves and
            # returns when done (the reference plans on every call)
            gpu_ms = fa["uncached"]["ms_per_call_median"]
            cb["same_basis"] = {"gpu_ms_per_call_planning_included": gpu_ms,
                                "cpu_ms_per_transform": cb["ms_per_transform"],
                                "gpu_over_cpu": round(cb["ms_per_transform"] / gpu_ms, 1)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
