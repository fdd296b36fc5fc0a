"""ctypes access to the test oracle (oracle/liboracle.so) and to the reference
harness (oracle/_ref/libg2g_ref.so, the reference's own sources compiled in
place). Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libg2g_ref.so")

F32, F64, C64, C128 = 0, 1, 2, 3
ES = {F32: 4, F64: 8, C64: 8, C128: 16}
NP_DTYPE = {F32: np.float32, F64: np.float64, C64: np.complex64, C128: np.complex128}


class OrcMsg(ctypes.Structure):
    _fields_ = [("rank", ctypes.c_int32), ("tag", ctypes.c_int32), ("r0", ctypes.c_int32),
                ("r1", ctypes.c_int32), ("c0", ctypes.c_int32), ("c1", ctypes.c_int32),
                ("offset", ctypes.c_int64), ("stride", ctypes.c_int64), ("data_off", ctypes.c_int64),
                ("block_index", ctypes.c_int32), ("transposed", ctypes.c_int32),
                ("conjugated", ctypes.c_int32), ("pad", ctypes.c_int32)]


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: make -C oracle")
        lib = ctypes.CDLL(ORACLE_SO)
        vp, i, i64, p32, p64 = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64))
        lib.orc_decomp_cover.argtypes = [p32, i, p32, i, p32]
        lib.orc_n_ranks.argtypes = [vp, vp]
        lib.orc_plan_rank.restype = i64
        lib.orc_plan_rank.argtypes = [vp, vp, i, ctypes.c_char, i, i, ctypes.POINTER(OrcMsg), i64, p64, p64]
        lib.orc_pack_rank.restype = i64
        lib.orc_pack_rank.argtypes = [vp, vp, i, ctypes.c_char, i, vp, i64]
        lib.orc_transform.argtypes = [vp, vp, i, ctypes.c_char, vp, vp]
        lib.orc_transform_batched.argtypes = [i, vp, vp, i, ctypes.c_char_p, vp, vp]
        lib.orc_line_split.argtypes = [i, i, i, p32, i]
        lib.orc_scalapack_grid.argtypes = [i] * 15 + [p32, p32, p32, p32, p32, p32, p32, p64, i]
        lib.orc_splitmix64.restype = ctypes.c_uint64
        lib.orc_splitmix64.argtypes = [ctypes.c_uint64]
        lib.orc_value.argtypes = [i, ctypes.c_uint64, i64, i64, i64, vp]
        lib.orc_fill_layout.argtypes = [vp, i, ctypes.c_uint64]
        lib.orc_check_layout.restype = i64
        lib.orc_check_layout.argtypes = [vp, i, ctypes.c_uint64, ctypes.c_char, i64]
        lib.orc_comm_volume.argtypes = [vp, vp, p64]
        lib.orc_optimal_reordering.argtypes = [p64, i, p32, ctypes.POINTER(ctypes.c_int)]
        lib.orc_cmul.argtypes = [i, vp, vp, i, vp]
        lib.orc_libgcc_mulxc3.argtypes = [i, vp, vp, vp]
        lib.orc_debug_order_xor.argtypes = [i]
        _oracle = lib
    return _oracle


def ref():
    """The reference harness, or None when it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        lib = ctypes.CDLL(REF_SO)
        vp, i, i64, p32, p64 = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64))
        lib.ref_decomp_cover.argtypes = [p32, i, p32, i, p32]
        lib.ref_transform.argtypes = [i, vp, vp, i, ctypes.c_char_p, vp, vp, ctypes.POINTER(ctypes.c_double)]
        lib.ref_plan_rank.argtypes = [vp, vp, i, ctypes.c_char, i, i, p64, p64, vp, i64, p64]
        lib.ref_comm_volume.argtypes = [vp, vp, p64]
        lib.ref_optimal_reordering.argtypes = [p64, i, p32, ctypes.POINTER(ctypes.c_int)]
        _ref = lib
    return _ref


def i32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def i64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def decomp_cover(lib_fn, blue, red):
    blue = np.ascontiguousarray(blue, dtype=np.int32)
    red = np.ascontiguousarray(red, dtype=np.int32)
    out = np.zeros(2 * (len(blue) - 1), dtype=np.int32)
    rc = lib_fn(i32p(blue), len(blue), i32p(red), len(red), i32p(out))
    return rc, out.reshape(-1, 2)


def scalar_buf(dtype, value):
    if value is None:
        return None
    a = np.array([value], dtype=NP_DTYPE[dtype])
    return ctypes.create_string_buffer(a.tobytes(), a.nbytes)


def transform(init_layout, init_arena, fin_layout, fin_arena, dtype, trans="N", alpha=None, beta=None):
    """Oracle pipeline on host arenas (numpy uint8); fin_arena is updated."""
    es = ES[dtype]
    di = init_layout.descriptor(init_arena.ctypes.data, es)
    df = fin_layout.descriptor(fin_arena.ctypes.data, es)
    a = scalar_buf(dtype, alpha)
    b = scalar_buf(dtype, beta)
    rc = oracle().orc_transform(ctypes.addressof(di.struct), ctypes.addressof(df.struct), dtype,
                                trans.encode(), a, b)
    assert rc == 0, "oracle transform failed"


def ref_transform(pairs, dtype, trans=None, alphas=None, betas=None):
    """Reference pipeline. pairs = [(init_layout, init_arena, fin_layout, fin_arena), ...]."""
    lib = ref()
    es = ES[dtype]
    n = len(pairs)
    descs = []
    inits = (ctypes.c_byte * (48 * n))()
    fins = (ctypes.c_byte * (48 * n))()
    for k, (il, ia, fl, fa) in enumerate(pairs):
        di = il.descriptor(ia.ctypes.data, es)
        df = fl.descriptor(fa.ctypes.data, es)
        descs += [di, df]
        ctypes.memmove(ctypes.addressof(inits) + 48 * k, ctypes.addressof(di.struct), 48)
        ctypes.memmove(ctypes.addressof(fins) + 48 * k, ctypes.addressof(df.struct), 48)
    t = (trans or "N" * n).encode()
    ab = bb = None
    if alphas is not None:
        ab = np.array(alphas, dtype=NP_DTYPE[dtype])
        bb = np.array(betas, dtype=NP_DTYPE[dtype])
    secs = ctypes.c_double()
    rc = lib.ref_transform(n, ctypes.addressof(inits), ctypes.addressof(fins), dtype, t,
                           ab.ctypes.data if ab is not None else None,
                           bb.ctypes.data if bb is not None else None, ctypes.byref(secs))
    assert rc == 0, f"reference transform failed ({rc})"
    return secs.value


def batched(pairs, dtype, trans=None, alphas=None, betas=None):
    """Oracle batched pipeline (one exchange for all pairs)."""
    lib = oracle()
    es = ES[dtype]
    n = len(pairs)
    descs = []
    inits = (ctypes.c_byte * (48 * n))()
    fins = (ctypes.c_byte * (48 * n))()
    for k, (il, ia, fl, fa) in enumerate(pairs):
        di = il.descriptor(ia.ctypes.data, es)
        df = fl.descriptor(fa.ctypes.data, es)
        descs += [di, df]
        ctypes.memmove(ctypes.addressof(inits) + 48 * k, ctypes.addressof(di.struct), 48)
        ctypes.memmove(ctypes.addressof(fins) + 48 * k, ctypes.addressof(df.struct), 48)
    t = (trans or "N" * n).encode()
    ab = bb = None
    if alphas is not None:
        ab = np.array(alphas, dtype=NP_DTYPE[dtype])
        bb = np.array(betas, dtype=NP_DTYPE[dtype])
    rc = lib.orc_transform_batched(n, ctypes.addressof(inits), ctypes.addressof(fins), dtype, t,
                                   ab.ctypes.data if ab is not None else None,
                                   bb.ctypes.data if bb is not None else None)
    assert rc == 0


def plan_rank(init_layout, init_arena, fin_layout, fin_arena, dtype, trans, rank, side):
    """Oracle messages (sorted), counts and dspls of one rank's send/recv side."""
    lib = oracle()
    es = ES[dtype]
    di = init_layout.descriptor(init_arena.ctypes.data, es)
    df = fin_layout.descriptor(fin_arena.ctypes.data, es)
    P = max(init_layout.n_ranks, fin_layout.n_ranks)
    counts = np.zeros(P, dtype=np.int64)
    dspls = np.zeros(P, dtype=np.int64)
    cap = (len(init_layout.rows_split) + len(fin_layout.rows_split)) * \
          (len(init_layout.cols_split) + len(fin_layout.cols_split)) * 2 + 16
    msgs = (OrcMsg * cap)()
    n = lib.orc_plan_rank(ctypes.addressof(di.struct), ctypes.addressof(df.struct), dtype, trans.encode(),
                          rank, side, msgs, cap, i64p(counts), i64p(dspls))
    assert n >= 0
    return list(msgs[:n]), counts, dspls


def pack_rank(init_layout, init_arena, fin_layout, fin_arena, dtype, trans, rank):
    lib = oracle()
    es = ES[dtype]
    di = init_layout.descriptor(init_arena.ctypes.data, es)
    df = fin_layout.descriptor(fin_arena.ctypes.data, es)
    cap = init_layout.m * init_layout.n + 1
    buf = np.zeros(cap * es, dtype=np.uint8)
    n = lib.orc_pack_rank(ctypes.addressof(di.struct), ctypes.addressof(df.struct), dtype, trans.encode(),
                          rank, buf.ctypes.data, cap)
    assert n >= 0
    return buf[:n * es]


def ref_plan_rank(init_layout, init_arena, fin_layout, fin_arena, dtype, trans, rank, side):
    lib = ref()
    es = ES[dtype]
    di = init_layout.descriptor(init_arena.ctypes.data, es)
    df = fin_layout.descriptor(fin_arena.ctypes.data, es)
    P = max(init_layout.n_ranks, fin_layout.n_ranks)
    counts = np.zeros(P, dtype=np.int64)
    dspls = np.zeros(P, dtype=np.int64)
    cap = init_layout.m * init_layout.n + 1
    buf = np.zeros(cap * es, dtype=np.uint8)
    total = ctypes.c_int64()
    rc = lib.ref_plan_rank(ctypes.addressof(di.struct), ctypes.addressof(df.struct), dtype, trans.encode(),
                           rank, side, i64p(counts), i64p(dspls), buf.ctypes.data if side == 0 else None, cap,
                           ctypes.byref(total))
    assert rc == 0
    return counts, dspls, buf[:total.value * es] if side == 0 else None


def fill(layout, arena, dtype, seed):
    d = layout.descriptor(arena.ctypes.data, ES[dtype])
    assert oracle().orc_fill_layout(ctypes.addressof(d.struct), dtype, seed) == 0


def check(layout, arena, dtype, seed, trans, a_n_cols):
    d = layout.descriptor(arena.ctypes.data, ES[dtype])
    return oracle().orc_check_layout(ctypes.addressof(d.struct), dtype, seed, trans.encode(), a_n_cols)


def value(dtype, seed, i, j, n_cols):
    buf = np.zeros(16, dtype=np.uint8)
    oracle().orc_value(dtype, seed, i, j, n_cols, buf.ctypes.data)
    return buf[:ES[dtype]].copy()


def comm_volume(lib_fn, init_layout, fin_layout):
    """Dense upper-triangular volume matrix from orc_/ref_comm_volume."""
    n = max(init_layout.n_ranks, fin_layout.n_ranks)
    di, df = init_layout.descriptor(0, 1), fin_layout.descriptor(0, 1)
    vol = np.zeros((n, n), dtype=np.int64)
    rc = lib_fn(ctypes.addressof(di.struct), ctypes.addressof(df.struct), i64p(vol))
    assert rc == 0, "comm_volume failed"
    return vol


def volume_dict(vol):
    a, b = np.nonzero(vol)
    return {(int(x), int(y)): int(vol[x, y]) for x, y in zip(a, b)}


def optimal_reordering(lib_fn, vol):
    """(permutation, reordered) from orc_/ref_optimal_reordering on a dense volume."""
    vol = np.ascontiguousarray(vol, dtype=np.int64)
    n = vol.shape[0]
    perm = np.zeros(n, dtype=np.int32)
    r = ctypes.c_int()
    rc = lib_fn(i64p(vol), n, i32p(perm), ctypes.byref(r))
    assert rc == 0, "optimal_reordering failed"
    return [int(x) for x in perm], bool(r.value)


# Bit patterns the scaled epilogue must treat exactly as the reference's x86
# build does: quiet and signalling NaNs with payloads and both signs, the
# infinities, signed zeros, the smallest and largest denormals, 1, the largest
# finite value, the smallest normal and values whose products overflow.
SPECIAL32 = np.array([0x7fc00000, 0xffc00000, 0x7f800001, 0x7fa12345, 0xffb00001, 0x7fc0beef,
                      0x7f800000, 0xff800000, 0x00000000, 0x80000000, 0x00000001, 0x807fffff,
                      0x3f800000, 0xbf800000, 0x7f7fffff, 0x00800000, 0x5f000000, 0xdf000000],
                     dtype=np.uint32)
SPECIAL64 = np.array([0x7ff8000000000000, 0xfff8000000000000, 0x7ff0000000000001, 0x7ff4123456789abc,
                      0xfff0000000000005, 0x7ff8dead0000beef, 0x7ff0000000000000, 0xfff0000000000000,
                      0, 0x8000000000000000, 1, 0x800fffffffffffff, 0x3ff0000000000000,
                      0xbff0000000000000, 0x7fefffffffffffff, 0x0010000000000000, 0x5ff0000000000000,
                      0xdff0000000000000], dtype=np.uint64)


def special_arena(dtype, n_elems, rng, frac=0.6):
    """n_elems elements of dtype as bytes: random bit patterns, `frac` of the
    components replaced by SPECIAL32/SPECIAL64 entries."""
    es = ES[dtype]
    if dtype in (F32, C64):
        k = n_elems * es // 4
        v = rng.integers(0, 2**32, size=k, dtype=np.uint64).astype(np.uint32)
        pool = SPECIAL32
    else:
        k = n_elems * es // 8
        v = rng.integers(0, 2**63, size=k, dtype=np.uint64) * 2 + rng.integers(0, 2, size=k, dtype=np.uint64)
        pool = SPECIAL64
    pick = rng.random(k) < frac
    v[pick] = pool[rng.integers(0, len(pool), size=int(pick.sum()))]
    return v.view(np.uint8)


def special_scalar(dtype, rng, p_special=0.7):
    """One alpha or beta of dtype: special components with probability p_special."""
    pool = SPECIAL32 if dtype in (F32, C64) else SPECIAL64
    n = 1 if dtype in (F32, F64) else 2
    if rng.random() < p_special:
        v = pool[rng.integers(0, len(pool), size=n)]
    else:
        v = rng.integers(0, 2**31, size=n).astype(pool.dtype) << (0 if pool.dtype == np.uint32 else 32)
    return v.view(NP_DTYPE[dtype])[0]


def cmul(dtype, l, r, swap=0):
    """The oracle's complex product of two complex scalars as bytes."""
    es = ES[dtype]
    a = np.array([l], dtype=NP_DTYPE[dtype])
    b = np.array([r], dtype=NP_DTYPE[dtype])
    out = np.zeros(es, dtype=np.uint8)
    oracle().orc_cmul(dtype, a.ctypes.data, b.ctypes.data, swap, out.ctypes.data)
    return out


def libgcc_mulxc3(dtype, l, r):
    es = ES[dtype]
    a = np.array([l], dtype=NP_DTYPE[dtype])
    b = np.array([r], dtype=NP_DTYPE[dtype])
    out = np.zeros(es, dtype=np.uint8)
    oracle().orc_libgcc_mulxc3(dtype, a.ctypes.data, b.ctypes.data, out.ctypes.data)
    return out


def scale_layout(layout, arena, dtype, beta):
    """Oracle grid_layout::scale_by on a host arena."""
    d = layout.descriptor(arena.ctypes.data, ES[dtype])
    b = np.array([beta], dtype=NP_DTYPE[dtype])
    lib = oracle()
    lib.orc_scale_layout.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    assert lib.orc_scale_layout(ctypes.addressof(d.struct), dtype, b.ctypes.data) == 0


def ref_scale_layout(layout, arena, dtype, beta):
    """The reference's block::scale_by over every block of a host arena."""
    d = layout.descriptor(arena.ctypes.data, ES[dtype])
    b = np.array([beta], dtype=NP_DTYPE[dtype])
    lib = ref()
    lib.ref_scale_layout.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    assert lib.ref_scale_layout(ctypes.addressof(d.struct), dtype, b.ctypes.data) == 0
