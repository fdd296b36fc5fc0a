"""Python front end of the C ABI: plans, batches and the test-support kernels.

Device memory is whatever the caller passes (an integer device address, e.g.
torch.Tensor.data_ptr()); streams are hipStream_t handles as integers
(torch.cuda.Stream.cuda_stream) or None for the null stream.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, G2GPlanInfo, G2GEdge, DTYPE_SIZE


def _scalar_bytes(dtype, value):
    """One element of dtype as a ctypes buffer (None stays None)."""
    if value is None:
        return None
    if dtype == _lib.FLOAT32:
        arr = np.array([value], dtype=np.float32)
    elif dtype == _lib.FLOAT64:
        arr = np.array([value], dtype=np.float64)
    elif dtype == _lib.COMPLEX64:
        arr = np.array([value], dtype=np.complex64)
    else:
        arr = np.array([value], dtype=np.complex128)
    return ctypes.create_string_buffer(arr.tobytes(), arr.nbytes)


def _tuning_dict(t):
    return {k: getattr(t, k) for k, _ in t._fields_}


def _stream(s):
    if s is None:
        return None
    if hasattr(s, "cuda_stream"):
        return ctypes.c_void_p(s.cuda_stream)
    return ctypes.c_void_p(int(s))


class Plan:
    """grid2grid::transform(initial, final, trans, alpha, beta) planned once
    for a layout pair (see g2g_plan_create)."""

    def __init__(self, initial, init_base, final, final_base, dtype, trans="N", stream=None):
        lib = _lib.load()
        self.dtype = dtype
        self.trans = trans
        es = DTYPE_SIZE.get(dtype, 8)  # an unknown dtype is rejected by the library
        self._di = initial.descriptor(init_base, es)
        self._df = final.descriptor(final_base, es)
        h = ctypes.c_void_p()
        check(lib.g2g_plan_create(self._di.ref(), self._df.ref(), dtype, trans.encode(), _stream(stream),
                                  ctypes.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().g2g_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def info(self):
        i = G2GPlanInfo()
        check(_lib.load().g2g_plan_get_info(self.handle, ctypes.byref(i)))
        return i

    def transform(self, alpha=None, beta=None, stream=None):
        a = _scalar_bytes(self.dtype, alpha)
        b = _scalar_bytes(self.dtype, beta)
        check(_lib.load().g2g_transform(self.handle, a, b, _stream(stream)))

    def transform_graph(self, stream, alpha=None, beta=None):
        """The same launch replayed from the plan's hipGraph (g2g_transform_graph)."""
        a = _scalar_bytes(self.dtype, alpha)
        b = _scalar_bytes(self.dtype, beta)
        check(_lib.load().g2g_transform_graph(self.handle, a, b, _stream(stream)))

    def transform_staged(self, alpha=None, beta=None, stream=None):
        a = _scalar_bytes(self.dtype, alpha)
        b = _scalar_bytes(self.dtype, beta)
        check(_lib.load().g2g_transform_staged(self.handle, a, b, _stream(stream)))

    def counts(self):
        P = self.info.n_ranks
        out = np.zeros(P * P, dtype=np.int64)
        check(_lib.load().g2g_plan_counts(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                          P * P))
        return out.reshape(P, P)

    def wire(self):
        n = self.info.n_pieces
        i32 = lambda: np.zeros(n, dtype=np.int32)
        i64 = lambda: np.zeros(n, dtype=np.int64)
        out = dict(row0=i32(), col0=i32(), rows=i32(), cols=i32(), src=i32(), dst=i32(),
                   send_off=i64(), recv_off=i64())
        P32 = ctypes.POINTER(ctypes.c_int32)
        P64 = ctypes.POINTER(ctypes.c_int64)
        check(_lib.load().g2g_plan_wire(
            self.handle, *(out[k].ctypes.data_as(P32) for k in ("row0", "col0", "rows", "cols", "src", "dst")),
            out["send_off"].ctypes.data_as(P64), out["recv_off"].ctypes.data_as(P64), n))
        return out

    @property
    def tuning(self):
        t = _lib.G2GTuning()
        check(_lib.load().g2g_plan_get_tuning(self.handle, ctypes.byref(t)))
        return _tuning_dict(t)

    def set_tuning(self, tile_bytes, threads=256, unroll_bytes=64, nontemporal=0, launch_mode=0,
                   column_chunks=0, stream_order=0, row_chunks=0, tail_split=0):
        """launch_mode: 0 one workgroup per work item, 1 persistent grid, 2 one
        workgroup per work item with a contiguous range of work items per XCD,
        3 persistent with one counter per XCD. column_chunks: 0 runs the
        pieces, w > 0 chunks of at most w columns ordered down each column
        interval. stream_order: 0 union-cell order, 1 destination order.
        row_chunks: 0 whole columns, h > 0 (untransposed plans) chunks of at
        most h rows. tail_split: 0, or 2/4/8: the stream's last round of work
        items cut into items of tile_bytes / tail_split."""
        t = _lib.G2GTuning(int(tile_bytes), threads, unroll_bytes, int(nontemporal), int(launch_mode),
                           int(column_chunks), int(stream_order), int(row_chunks), int(tail_split), 0.0)
        check(_lib.load().g2g_plan_set_tuning(self.handle, ctypes.byref(t)))

    def autotune(self, stream=None, reps=3):
        t = _lib.G2GTuning()
        check(_lib.load().g2g_plan_autotune(self.handle, _stream(stream), reps, ctypes.byref(t)))
        return _tuning_dict(t)

    def send_buffer(self, rank):
        p = ctypes.c_void_p()
        n = ctypes.c_int64()
        check(_lib.load().g2g_plan_send_buffer(self.handle, rank, ctypes.byref(p), ctypes.byref(n)))
        return p.value or 0, n.value


class Batch:
    """Several plans run by one persistent launch (g2g_batch_*)."""

    def __init__(self, plans, alphas=None, betas=None, stream=None):
        lib = _lib.load()
        n = len(plans)
        self.plans = list(plans)
        arr = (ctypes.c_void_p * n)(*[p.handle for p in plans])
        self._scal = []
        a_arr = b_arr = None
        if alphas is not None:
            a_bufs = [_scalar_bytes(p.dtype, a) for p, a in zip(plans, alphas)]
            b_bufs = [_scalar_bytes(p.dtype, b) for p, b in zip(plans, betas)]
            self._scal = a_bufs + b_bufs
            a_arr = (ctypes.c_void_p * n)(*[ctypes.cast(x, ctypes.c_void_p) if x is not None else None
                                            for x in a_bufs])
            b_arr = (ctypes.c_void_p * n)(*[ctypes.cast(x, ctypes.c_void_p) if x is not None else None
                                            for x in b_bufs])
        h = ctypes.c_void_p()
        check(lib.g2g_batch_create(n, arr, a_arr, b_arr, _stream(stream), ctypes.byref(h)))
        self.handle = h

    def execute(self, stream=None):
        check(_lib.load().g2g_batch_execute(self.handle, _stream(stream)))

    def execute_graph(self, stream):
        check(_lib.load().g2g_batch_execute_graph(self.handle, _stream(stream)))

    @property
    def bytes(self):
        v = ctypes.c_int64()
        check(_lib.load().g2g_batch_bytes(self.handle, ctypes.byref(v)))
        return v.value

    def autotune(self, stream=None, reps=2):
        t = _lib.G2GTuning()
        check(_lib.load().g2g_batch_autotune(self.handle, _stream(stream), reps, ctypes.byref(t)))
        return _tuning_dict(t)

    @property
    def tuning(self):
        t = _lib.G2GTuning()
        check(_lib.load().g2g_batch_get_tuning(self.handle, ctypes.byref(t)))
        return _tuning_dict(t)

    def set_tuning(self, tile_bytes, threads=512, unroll_bytes=128, nontemporal=0, launch_mode=1, stream=None):
        """launch_mode 1 (one work counter) or 3 (one per XCD); batches always launch persistent."""
        t = _lib.G2GTuning(int(tile_bytes), threads, unroll_bytes, int(nontemporal), int(launch_mode), 0, 0, 0, 0,
                           0.0)
        check(_lib.load().g2g_batch_set_tuning(self.handle, ctypes.byref(t), _stream(stream)))

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().g2g_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def transform_plans_noalloc(plans, alphas=None, betas=None, stream=None):
    """Several plans in one launch with no device allocation
    (g2g_transform_plans_noalloc); calls that overlap on one GPU race."""
    n = len(plans)
    arr = (ctypes.c_void_p * max(1, n))(*[p.handle for p in plans])
    keep = []
    a_arr = b_arr = None
    if alphas is not None:
        a_bufs = [_scalar_bytes(p.dtype, a) for p, a in zip(plans, alphas)]
        b_bufs = [_scalar_bytes(p.dtype, b) for p, b in zip(plans, betas)]
        keep = a_bufs + b_bufs
        a_arr = (ctypes.c_void_p * n)(*[ctypes.cast(x, ctypes.c_void_p) if x is not None else None for x in a_bufs])
        b_arr = (ctypes.c_void_p * n)(*[ctypes.cast(x, ctypes.c_void_p) if x is not None else None for x in b_bufs])
    check(_lib.load().g2g_transform_plans_noalloc(n, arr, a_arr, b_arr, _stream(stream)))
    del keep


SERVE_BATCHED, SERVE_PLAIN = 0, 1


class Server:
    """Request loop with continuous batching under an HBM budget (g2g_server_*)."""

    def __init__(self, budget_bytes, n_streams=2, max_batch=256, mode=SERVE_BATCHED):
        h = ctypes.c_void_p()
        check(_lib.load().g2g_server_create(int(budget_bytes), n_streams, max_batch, mode, ctypes.byref(h)))
        self.handle = h
        self._keep = {}

    def submit(self, plan, alpha=None, beta=None):
        a = _scalar_bytes(plan.dtype, alpha)
        b = _scalar_bytes(plan.dtype, beta)
        t = ctypes.c_int64()
        check(_lib.load().g2g_server_submit(self.handle, plan.handle, a, b, ctypes.byref(t)))
        return t.value

    def poll(self, ticket):
        done = ctypes.c_int32()
        lat = ctypes.c_double()
        check(_lib.load().g2g_server_poll(self.handle, ticket, ctypes.byref(done), ctypes.byref(lat)))
        return (lat.value if done.value else None)

    def wait(self, ticket):
        lat = ctypes.c_double()
        check(_lib.load().g2g_server_wait(self.handle, ticket, ctypes.byref(lat)))
        return lat.value

    def drain(self):
        check(_lib.load().g2g_server_drain(self.handle))

    def stats(self):
        s = _lib.G2GServerStats()
        check(_lib.load().g2g_server_get_stats(self.handle, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().g2g_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_tuning(tile_bytes=0, launch_mode=-1):
    """Process-wide override of the work-item size (0 = automatic) and launch
    mode (-1 = automatic) for plans and launches made afterwards."""
    check(_lib.load().g2g_set_tuning(int(tile_bytes), int(launch_mode)))


def transform_layouts(initial, init_base, final, final_base, dtype, trans="N", alpha=None, beta=None,
                      stream=None):
    """The drop-in call (g2g_transform_layouts): plans the layout pair on its
    first call, replays the cached plan afterwards."""
    es = DTYPE_SIZE.get(dtype, 8)
    di = initial.descriptor(init_base, es)
    df = final.descriptor(final_base, es)
    a = _scalar_bytes(dtype, alpha)
    b = _scalar_bytes(dtype, beta)
    check(_lib.load().g2g_transform_layouts(di.ref(), df.ref(), dtype, trans.encode(), a, b, _stream(stream)))


def plan_cache_stats():
    """(hits, misses, entries) of the plan cache behind transform_layouts."""
    h, m, e = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    check(_lib.load().g2g_plan_cache_stats(ctypes.byref(h), ctypes.byref(m), ctypes.byref(e)))
    return h.value, m.value, e.value


def plan_cache_clear():
    check(_lib.load().g2g_plan_cache_clear())


def plan_cache_set_capacity(entries):
    check(_lib.load().g2g_plan_cache_set_capacity(int(entries)))


def scale_layout(layout, base, dtype, beta, stream=None):
    """grid_layout::scale_by on device memory (g2g_scale_layout)."""
    d = layout.descriptor(base, DTYPE_SIZE[dtype])
    check(_lib.load().g2g_scale_layout(d.ref(), dtype, _scalar_bytes(dtype, beta), _stream(stream)))


def fill_layout(layout, base, dtype, seed, stream=None):
    d = layout.descriptor(base, DTYPE_SIZE[dtype])
    check(_lib.load().g2g_fill_layout(d.ref(), dtype, ctypes.c_uint64(seed), _stream(stream)))


def verify_layout(layout, base, dtype, seed, trans, a_n_cols, stream=None):
    d = layout.descriptor(base, DTYPE_SIZE[dtype])
    bad = ctypes.c_int64()
    check(_lib.load().g2g_verify_layout(d.ref(), dtype, ctypes.c_uint64(seed), trans.encode(), a_n_cols,
                                        _stream(stream), ctypes.byref(bad)))
    return bad.value


def comm_volume(initial, final, stream=None):
    """Elements exchanged between each unordered pair of ranks {a, b}, a <= b
    (a == b: elements a rank keeps), as {(a, b): count}, computed on the
    device (reference: transform.cpp:9-39, utils.hpp:91-139)."""
    lib = _lib.load()
    di, df = initial.descriptor(0, 1), final.descriptor(0, 1)
    n = ctypes.c_int64()
    check(lib.g2g_comm_volume(di.ref(), df.ref(), None, 0, ctypes.byref(n), _stream(stream)))
    edges = (G2GEdge * max(1, n.value))()
    check(lib.g2g_comm_volume(di.ref(), df.ref(), edges, n.value, ctypes.byref(n), _stream(stream)))
    return {(int(e.src), int(e.dest)): int(e.weight) for e in edges[:n.value]}


def optimal_reordering(volume, n_ranks):
    """Greedy rank relabelling on a volume {(a, b): count}
    (reference: ranks_reordering.cpp:4-61). Returns (permutation, reordered)."""
    lib = _lib.load()
    items = list(volume.items())
    edges = (G2GEdge * max(1, len(items)))()
    for k, ((a, b), w) in enumerate(items):
        edges[k].src, edges[k].dest, edges[k].weight = int(a), int(b), int(w)
    perm = (ctypes.c_int32 * n_ranks)()
    reordered = ctypes.c_int32()
    check(lib.g2g_optimal_reordering(edges, len(items), n_ranks, perm, ctypes.byref(reordered)))
    return [int(x) for x in perm], bool(reordered.value)


def stream_probe(kind, src, dst, nbytes, reps=10, stream=None):
    """Streaming reference rate of the device: kind 0 copy src -> dst, 1 read
    src, 2 write dst; returns milliseconds per pass over nbytes (see
    g2g_stream_probe)."""
    ms = ctypes.c_double()
    check(_lib.load().g2g_stream_probe(int(kind), ctypes.c_void_p(src or None), ctypes.c_void_p(dst or None),
                                       int(nbytes), int(reps), _stream(stream), ctypes.byref(ms)))
    return ms.value
