#!/usr/bin/env python3
"""Benchmark of the MI355X grid2grid transform.

Default workload: the 32768^2 fp64 block-cyclic redistribution (128x128
blocks on a 4x4 rank grid -> 256x512 blocks on 2x8, 16 virtual ranks, 8 GiB),
run by the fused gather kernel. One step = one full transform of the matrix.
value = moved GB/s = 2*m*n*sizeof(T) / device time per step (one read and one
write per element; inputs resident in HBM). Output: one JSON line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cfg2]
                       [--mode fused|staged] [--no-cpu-baseline] [--no-extras]
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cfg2", choices=["cfg1", "cfg2", "cfg3"])
    ap.add_argument("--mode", default="fused", choices=["fused", "staged"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline leg")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(name):
    from grid2grid_amd import workloads
    return {"cfg1": workloads.cfg1, "cfg2": workloads.cfg2, "cfg3": workloads.cfg3}[name]()


WORKLOAD_DESC = {
    "cfg1": "block-cyclic 4096^2 f64, 32x32 on 2x2 -> 64x128 on 2x2, 4 virtual ranks",
    "cfg2": "block-cyclic 32768^2 f64, 128x128 on 4x4 -> 256x512 on 2x8, 16 virtual ranks",
    "cfg3": "irregular 40000x25000 f32, 64x48 -> 50x40 splits, 32 virtual ranks, padded ld",
}


def time_steps(torch, fn, steps, warmup, stream):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    per = []
    # one event pair per step on the launch stream: average launch duration
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    start.record(stream)
    for a, b in evs:
        a.record(stream)
        fn()
        b.record(stream)
    end.record(stream)
    torch.cuda.synchronize()
    for a, b in evs:
        per.append(a.elapsed_time(b))
    return start.elapsed_time(end) / steps, per


def read_traffic(workload, mode):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", f"gather_{workload}_pmc.txt" if mode == "fused" else
                        f"staged_{workload}_pmc.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9]+)", open(path).read())
    return int(m.group(1)) if m else None


def cpu_baseline(budget_s):
    """The reference's own pipeline (oracle/_ref, compiled from the reference
    sources) on a bounded sample of the default workload: 8192^2 f64 with the
    same block sizes and rank grids (128x128 on 4x4 -> 256x512 on 2x8), 16
    virtual ranks in one process, OpenMP over all host threads. Falls back to
    the C restatement (single thread) when the reference build is absent."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from grid2grid_amd import workloads
    init, fin, dtype, _ = workloads.cfg2(m=8192)
    es = O.ES[dtype]
    ia = np.zeros(init.arena_elems * es, dtype=np.uint8)
    fa = np.zeros(fin.arena_elems * es, dtype=np.uint8)
    ia[::4096] = 1
    nbytes = 2 * init.m * init.n * es
    kind = "reference" if O.ref() is not None else "port"
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)) if kind == "reference" else 1
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
        if n > 2:  # two warm-ups
            times.append(dt)
        if (time.time() > t_end and len(times) >= 3) or len(times) >= 50:
            break
    med = statistics.median(times)
    return {
        "value": round(nbytes / med / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": kind,
        "sample": f"8192^2 f64, config-2 blocks/grids (128x128 on 4x4 -> 256x512 on 2x8), 16 virtual ranks "
                  f"in one process, median of {len(times)} runs after 2 warm-ups, {threads} OpenMP threads; "
                  f"time includes the reference's per-call planning",
        "ms_per_transform": round(med * 1e3, 3),
    }


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

    init, fin, dtype, trans = build_workload(args.workload)
    es = g2g.DTYPE_SIZE[dtype]
    log(f"workload {args.workload}: {WORKLOAD_DESC[args.workload]}")
    ti = torch.empty(init.arena_elems * es, dtype=torch.uint8, device=dev)
    tf = torch.zeros(fin.arena_elems * es, dtype=torch.uint8, device=dev)
    g2g.fill_layout(init, ti.data_ptr(), dtype, 2, stream)
    t0 = time.time()
    plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans, stream)
    info = plan.info
    log(f"plan: {info.n_pieces} pieces, {info.n_tiles} work items of {info.tile_bytes} B, "
        f"planned in {time.time() - t0:.2f}s")

    run = {"fused": lambda: plan.transform(stream=stream),
           "staged": lambda: plan.transform_staged(stream=stream)}[args.mode]
    ms, per = time_steps(torch, run, args.steps, args.warmup, stream)
    a_n_cols = init.n
    bad = g2g.verify_layout(fin, tf.data_ptr(), dtype, 2, trans, a_n_cols, stream)
    nbytes = workloads.algorithmic_bytes(fin, dtype)
    gbs = nbytes / (ms * 1e-3) / 1e9
    per_launch_ms = statistics.mean(per)
    achieved = nbytes / (per_launch_ms * 1e-3) / 1e9
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
            "workload": WORKLOAD_DESC[args.workload],
            "mode": args.mode,
            "m": fin.m, "n": fin.n,
            "virtual_ranks": int(max(init.n_ranks, fin.n_ranks)),
            "pieces": int(info.n_pieces),
            "work_items": int(info.n_tiles),
            "tile_bytes": int(info.tile_bytes),
            "parallelism": "1 GPU, P virtual ranks",
        },
        "bit_exact": bad == 0,
        "mismatches": int(bad),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": read_traffic(args.workload, args.mode),
            "algorithmic_bytes": int(nbytes),
        },
    }

    if not args.no_extras and args.mode == "fused":
        # plain two-stage form on the same plan, for comparison
        tf.zero_()
        ms_s, _ = time_steps(torch, lambda: plan.transform_staged(stream=stream), max(3, args.steps // 4), 1,
                             stream)
        bad_s = g2g.verify_layout(fin, tf.data_ptr(), dtype, 2, trans, a_n_cols, stream)
        result["plain_staged"] = {"ms_per_step": round(ms_s, 4), "value": round(nbytes / (ms_s * 1e-3) / 1e9, 2),
                                  "unit": "GB/s", "bit_exact": bad_s == 0}
        log(f"staged: {ms_s:.3f} ms/step, mismatches {bad_s}")
    plan.close()
    del ti, tf
    torch.cuda.empty_cache()

    if not args.no_cpu_baseline:
        try:
            result["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
            log(f"cpu baseline: {result['cpu_baseline']}")
        except Exception as e:  # the baseline is reported, never the target
            result["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
