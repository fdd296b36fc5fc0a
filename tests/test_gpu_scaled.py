"""GPU parity of the scaled epilogue on non-finite and denormal values, and of
grid_layout::scale_by: the HIP path, called through the C ABI, against the
oracle (itself pinned to the reference's build, test_scaled_semantics.py) on
arenas full of NaN payloads, infinities, signed zeros and denormals, with
NaN, infinite, zero and random alphas and betas. Every byte of the final
arena is compared."""
import numpy as np
import pytest

import oracle_lib as O
import grid2grid_amd as g2g
from grid2grid_amd.layout import scalapack_layout, arbitrary_layout

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DTYPES = [g2g.FLOAT32, g2g.FLOAT64, g2g.COMPLEX64, g2g.COMPLEX128]


def _dev(host):
    return torch.from_numpy(host).to("cuda")


def _nan_scalar(dtype, rng):
    pool = O.SPECIAL32 if dtype in (O.F32, O.C64) else O.SPECIAL64
    n = 1 if dtype in (O.F32, O.F64) else 2
    return pool[rng.integers(0, 6, size=n)].view(O.NP_DTYPE[dtype])[0]


def _pairs(trans):
    """Layout pairs covering every call site of the reference's scaled loops:
    strided and contiguous pieces, local and remote, one-row columns, odd
    element counts, and transposed local pieces past one tile of
    copy_transpose_and_scale in both dimensions."""
    rng = np.random.default_rng(404)
    def fin_shape(m, n):
        return (n, m) if trans != "N" else (m, n)
    out = []
    m, n = 37, 29
    out.append((scalapack_layout(m, n, 4, 3, 2, 2, "C"), scalapack_layout(*fin_shape(m, n), 5, 7, 2, 2, "C")))
    out.append((scalapack_layout(m, n, 37, 29, 1, 1, "C", lld="rows"),
                scalapack_layout(*fin_shape(m, n), *fin_shape(m, n), 1, 1, "C", lld="rows")))
    m, n = 300, 170
    out.append((scalapack_layout(m, n, 300, 170, 1, 1, "C", lld="rows"),
                scalapack_layout(*fin_shape(m, n), *fin_shape(m, n), 1, 1, "C", lld="rows")))
    out.append((scalapack_layout(m, n, 150, 85, 2, 2, "C", lld="rows"),
                scalapack_layout(*fin_shape(m, n), 100, 60, 2, 1, "C")))
    m, n = 130, 210
    a = arbitrary_layout(rng, m, n, 4, 5, 4, lo=1, hi=80, pad=(0, 6), gap=(0, 17))
    fm, fn = fin_shape(m, n)
    b = arbitrary_layout(rng, fm, fn, 6, 3, 3, lo=1, hi=80, pad=(0, 6), gap=(0, 17))
    out.append((a, b))
    return out


def _scalars(dtype, k, rng):
    alpha, beta = O.special_scalar(dtype, rng), O.special_scalar(dtype, rng)
    if k % 2 == 0:
        alpha = _nan_scalar(dtype, rng)
    if k % 4 < 2:
        beta = _nan_scalar(dtype, rng)
    return alpha, beta


@pytest.mark.parametrize("form", ["fused", "staged", "graph", "chunked"])
@pytest.mark.parametrize("dtype", DTYPES)
def test_scaled_special_values_bit_exact(dtype, form):
    trans_set = ["N", "T", "C"] if dtype in (g2g.COMPLEX64, g2g.COMPLEX128) else ["N", "T"]
    rng = np.random.default_rng(1000 + dtype)
    stream = torch.cuda.Stream() if form == "graph" else None
    k = 0
    for trans in trans_set:
        for init, fin in _pairs(trans):
            alpha, beta = _scalars(dtype, k, rng)
            k += 1
            ia = O.special_arena(dtype, init.arena_elems, rng)
            fa = O.special_arena(dtype, fin.arena_elems, rng)
            want = fa.copy()
            O.transform(init, ia.copy(), fin, want, dtype, trans, alpha, beta)
            print(f"scaled {form} dtype {dtype} trans {trans} {init.m}x{init.n} alpha {alpha!r} beta {beta!r}",
                  flush=True)
            ti, tf = _dev(ia), _dev(fa)
            plan = g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans)
            try:
                if form == "chunked":  # the plan's gather runs column chunks; scaled requests run the pieces
                    t = plan.tuning
                    plan.set_tuning(t["tile_bytes"], 256, 64, 0, 1, column_chunks=3)
                if form == "staged":
                    plan.transform_staged(alpha, beta)
                elif form == "graph":
                    with torch.cuda.stream(stream):
                        plan.transform_graph(stream, alpha, beta)
                    stream.synchronize()
                else:
                    plan.transform(alpha, beta)
                torch.cuda.synchronize()
                np.testing.assert_array_equal(tf.cpu().numpy(), want, err_msg=f"{trans} {init.m}x{init.n}")
            finally:
                plan.close()


def test_scaled_special_values_batched():
    """Scaled requests of every type with special operands next to plain ones
    in one persistent launch."""
    rng = np.random.default_rng(77)
    cases = []
    for k, dtype in enumerate(DTYPES * 2):
        trans = ["N", "T"][k % 2]
        init, fin = _pairs(trans)[k % 5]
        alpha, beta = _scalars(dtype, k, rng) if k % 3 else (None, None)
        ia = O.special_arena(dtype, init.arena_elems, rng)
        fa = O.special_arena(dtype, fin.arena_elems, rng)
        want = fa.copy()
        O.transform(init, ia.copy(), fin, want, dtype, trans, alpha, beta)
        ti, tf = _dev(ia), _dev(fa)
        cases.append((g2g.Plan(init, ti.data_ptr(), fin, tf.data_ptr(), dtype, trans), ti, tf, want, alpha, beta))
    batch = g2g.Batch([c[0] for c in cases], [c[4] for c in cases], [c[5] for c in cases])
    batch.execute()
    torch.cuda.synchronize()
    for k, (plan, ti, tf, want, alpha, beta) in enumerate(cases):
        np.testing.assert_array_equal(tf.cpu().numpy(), want, err_msg=f"request {k}")
    batch.close()
    for c in cases:
        c[0].close()


@pytest.mark.parametrize("dtype", DTYPES)
def test_scale_layout_bit_exact(dtype):
    """grid_layout::scale_by on the device equals the oracle (and so the
    reference's block::scale_by) for special betas; beta == 1 and padding and
    gaps leave every byte alone."""
    rng = np.random.default_rng(9 + dtype)
    lay = arbitrary_layout(rng, 300, 200, 7, 5, 3, lo=3, hi=80, pad=(0, 5), gap=(0, 9))
    betas = [_nan_scalar(dtype, rng) for _ in range(2)] + [O.special_scalar(dtype, rng) for _ in range(4)]
    betas.append(np.array([1], dtype=O.NP_DTYPE[dtype])[0])
    for beta in betas:
        a = O.special_arena(dtype, lay.arena_elems, rng)
        want = a.copy()
        O.scale_layout(lay, want, dtype, beta)
        t = _dev(a)
        g2g.scale_layout(lay, t.data_ptr(), dtype, beta)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(t.cpu().numpy(), want, err_msg=repr(beta))
