"""The scaled epilogue dst = beta*dst + alpha*op(src) of the oracle, pinned on
non-finite and denormal values (host only).

The reference computes it with std::complex / float arithmetic compiled for
x86-64 (memory_utils.hpp:37, 237, 251): NaN operands are resolved by the SSE
rule (the first source operand if it is a NaN), invalid operations give the
x86 default NaN, and complex products fall back to libgcc's __mulsc3 /
__muldc3 (C11 Annex G.5.1) when both parts are NaN. Which operand is the
first source is a property of the compiled reference, per call site and
element position; the oracle states it, and these tests hold that statement
against the reference's own build (oracle/_ref) on arenas full of NaN
payloads, infinities, signed zeros and denormals, and the complex product
against libgcc itself."""
import itertools

import numpy as np
import pytest

import oracle_lib as O
from grid2grid_amd.layout import scalapack_layout

needs_ref = pytest.mark.skipif(O.ref() is None, reason="reference harness (oracle/_ref) not built")

ORD_M1S, ORD_M2S, ORD_ARE, ORD_AIM = 1, 2, 4, 8


@pytest.mark.parametrize("dtype", [O.C64, O.C128])
def test_complex_product_matches_libgcc(dtype):
    """The oracle's complex product (GCC's inline (ac - bd, ad + bc), then the
    Annex G.5.1 recomputation when both parts are NaN) gives libgcc's
    __mul?c3 bits for every combination of special parts, in both argument
    orders."""
    pool = (O.SPECIAL32 if dtype == O.C64 else O.SPECIAL64)[::2]
    comp = pool.view(np.float32 if dtype == O.C64 else np.float64)
    vals = [complex(a, b) for a, b in itertools.product(comp, comp)]
    n = 0
    for x, y in itertools.product(vals, vals):
        got = O.cmul(dtype, x, y, swap=0)
        want = O.libgcc_mulxc3(dtype, x, y)
        assert got.tobytes() == want.tobytes(), (x, y, got.tobytes().hex(), want.tobytes().hex())
        got_swapped = O.cmul(dtype, y, x, swap=1)
        assert got_swapped.tobytes() == want.tobytes(), (x, y)
        n += 1
    assert n == len(vals) ** 2


def test_default_nan_and_operand_rule_known_answers():
    """Closed-form cases of the x86 rules: inf*0 gives the default NaN
    (0xffc00000: sign and quiet bit set), a signalling NaN comes out quieted
    with its payload, and a NaN in the first source operand beats the other."""
    # (inf + 0i) * (0 + 1i): re = inf*0 - 0*1 = NaN (default), im = inf*1 + 0*0 = inf
    got = O.cmul(O.C64, complex(np.inf, 0), complex(0, 1))
    assert got.view(np.uint32).tolist() == [0xffc00000, 0x7f800000]
    # signalling NaN payload quieted: (sNaN(0x12345) + 1i) * (1 + 0i) -> both parts NaN -> libgcc
    l = np.array([0x7f812345, 0x3f800000], dtype=np.uint32).view(np.complex64)[0]
    got = O.cmul(O.C64, l, complex(1, 0))
    want = O.libgcc_mulxc3(O.C64, l, complex(1, 0))
    assert got.view(np.uint32)[0] == 0x7fc12345 and got.tobytes() == want.tobytes()


def _geometries(rng, trans):
    out = []
    for _ in range(8):
        m, n = int(rng.integers(1, 90)), int(rng.integers(1, 90))
        fm, fn = (n, m) if trans != "N" else (m, n)
        lay = lambda a, b: scalapack_layout(a, b, int(rng.integers(1, 40)), int(rng.integers(1, 40)),
                                            int(rng.integers(1, 3)), int(rng.integers(1, 3)), "C",
                                            lld=str(rng.choice(["rows", "padded"])))
        out.append((lay(m, n), lay(fm, fn)))
    # pieces past one tile of copy_transpose_and_scale (128 f32, 64 f64/c64,
    # 32 c128 elements) in both dimensions, contiguous and strided, local and remote
    m, n = 300, 170
    fm, fn = (n, m) if trans != "N" else (m, n)
    out.append((scalapack_layout(m, n, 300, 170, 1, 1, "C", lld="rows"),
                scalapack_layout(fm, fn, fm, fn, 1, 1, "C", lld="rows")))
    out.append((scalapack_layout(m, n, 150, 85, 2, 2, "C", lld="rows"),
                scalapack_layout(fm, fn, 100, 60, 2, 1, "C")))
    return out


def _nan_scalar(dtype, rng):
    pool = O.SPECIAL32 if dtype in (O.F32, O.C64) else O.SPECIAL64
    nans = pool[[0, 1, 2, 3, 4, 5]]
    n = 1 if dtype in (O.F32, O.F64) else 2
    return nans[rng.integers(0, len(nans), size=n)].view(O.NP_DTYPE[dtype])[0]


def _mismatches(seed):
    """Elements where the oracle and the reference differ, over scaled
    transforms of every type and operation on special-value arenas."""
    rng = np.random.default_rng(seed)
    bad = {}
    for dtype in (O.F32, O.F64, O.C64, O.C128):
        for trans in ("N", "T", "C") if dtype in (O.C64, O.C128) else ("N", "T"):
            for k, (init, fin) in enumerate(_geometries(rng, trans)):
                es = O.ES[dtype]
                ia = O.special_arena(dtype, init.arena_elems, rng)
                fa = O.special_arena(dtype, fin.arena_elems, rng)
                # alpha and beta: NaNs with payloads in every combination, else
                # any special or random value
                alpha, beta = O.special_scalar(dtype, rng), O.special_scalar(dtype, rng)
                if k % 2 == 0:
                    alpha = _nan_scalar(dtype, rng)
                if k % 4 < 2:
                    beta = _nan_scalar(dtype, rng)
                want = fa.copy()
                O.ref_transform([(init, ia.copy(), fin, want)], dtype, trans, [alpha], [beta])
                got = fa.copy()
                O.transform(init, ia.copy(), fin, got, dtype, trans, alpha, beta)
                n = int((want.reshape(-1, es) != got.reshape(-1, es)).any(axis=1).sum())
                bad[(dtype, trans)] = bad.get((dtype, trans), 0) + n
    return bad


@needs_ref
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_scaled_special_values_match_reference(seed):
    """Every element equals the reference's, bit for bit."""
    bad = _mismatches(seed)
    assert not any(bad.values()), bad


@needs_ref
@pytest.mark.parametrize("planted", [ORD_M1S, ORD_M2S, ORD_ARE, ORD_AIM])
def test_planted_operand_order_is_caught(planted):
    """A wrong operand order planted in the oracle (one order bit flipped
    everywhere) makes the comparison with the reference fail, for the types
    the bit applies to."""
    O.oracle().orc_debug_order_xor(planted)
    try:
        bad = _mismatches(0)
    finally:
        O.oracle().orc_debug_order_xor(0)
    complex_only = planted == ORD_AIM
    for (dtype, trans), n in bad.items():
        if complex_only and dtype in (O.F32, O.F64):
            assert n == 0
        else:
            assert n > 0, (planted, dtype, trans)


@needs_ref
@pytest.mark.parametrize("dtype", [O.F32, O.F64, O.C64, O.C128])
def test_scale_by_matches_reference(dtype):
    """grid_layout::scale_by of the oracle equals the reference's
    block::scale_by on special-value blocks with padding and gaps, for special
    and plain betas (beta == 1 leaves everything untouched); a planted wrong
    operand order is caught."""
    from grid2grid_amd.layout import arbitrary_layout
    rng = np.random.default_rng(7 + dtype)
    lay = arbitrary_layout(rng, 70, 50, 5, 4, 3, lo=3, hi=30, pad=(0, 4), gap=(0, 6))
    betas = [_nan_scalar(dtype, rng) for _ in range(3)] + [O.special_scalar(dtype, rng) for _ in range(6)]
    betas.append(np.array([1], dtype=O.NP_DTYPE[dtype])[0])
    for beta in betas:
        a = O.special_arena(dtype, lay.arena_elems, rng)
        want, got = a.copy(), a.copy()
        O.ref_scale_layout(lay, want, dtype, beta)
        O.scale_layout(lay, got, dtype, beta)
        np.testing.assert_array_equal(got, want, err_msg=repr(beta))
    # a planted mistake: the operand order swapped in the oracle
    a = O.special_arena(dtype, lay.arena_elems, rng, frac=1.0)
    want, got = a.copy(), a.copy()
    O.ref_scale_layout(lay, want, dtype, betas[0])
    O.oracle().orc_debug_order_xor(ORD_M1S)
    try:
        O.scale_layout(lay, got, dtype, betas[0])
    finally:
        O.oracle().orc_debug_order_xor(0)
    assert (got != want).any()
