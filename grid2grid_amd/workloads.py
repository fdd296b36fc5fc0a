"""The benchmark workloads (layout pairs).

bc4096: 4096^2 f64, 32x32 blocks on a 2x2 column-major rank grid -> 64x128
   blocks on 2x2; lld = local rows; 4 virtual ranks.
bc32768: 32768^2 f64 (8 GiB), 128x128 blocks on 4x4 -> 256x512 blocks on 2x8;
   lld = local rows; 16 virtual ranks.
irregular: 40000x25000 f32, irregular 64x48 -> 50x40 grids (block sizes
   U[200,1200] rescaled), owners uniform over 32 ranks, ld = rows + U[0,63],
   blocks placed in shuffled order with gaps.
mixed_requests: 2048 independent transforms: m, n log-uniform in [64, 4096],
   dtype U{f32, f64, c64}, block sizes from {16..256} per dim, rank grids
   from {1..4}x{1..4} for each side, 30% transposed.
served stream: Poisson arrivals of mixed requests plus 5% of large_request
   (16384^2 f64, 128x128 on 4x4 -> 256x256 on 2x8).
Value seeds 1, 2, 3 (bc4096, bc32768, irregular); geometry seeds 1003
(irregular) and 1004 (mixed_requests).
"""
import numpy as np

from . import _lib
from .layout import scalapack_layout, arbitrary_layout


def bc4096():
    init = scalapack_layout(4096, 4096, 32, 32, 2, 2, "C", lld="rows")
    fin = scalapack_layout(4096, 4096, 64, 128, 2, 2, "C", lld="rows")
    return init, fin, _lib.FLOAT64, "N"


def bc32768(m=32768):
    init = scalapack_layout(m, m, 128, 128, 4, 4, "C", lld="rows")
    fin = scalapack_layout(m, m, 256, 512, 2, 8, "C", lld="rows")
    return init, fin, _lib.FLOAT64, "N"


def irregular(m=40000, n=25000, seed=1003):
    rng = np.random.default_rng(seed)
    init = arbitrary_layout(rng, m, n, 64, 48, 32, lo=200, hi=1200, pad=(0, 63), gap=(0, 64))
    fin = arbitrary_layout(rng, m, n, 50, 40, 32, lo=200, hi=1200, pad=(0, 63), gap=(0, 64))
    return init, fin, _lib.FLOAT32, "N"


BLOCKS = (16, 32, 64, 128, 256)


def random_request(rng, dtypes=(_lib.FLOAT32, _lib.FLOAT64, _lib.COMPLEX64), lo=64, hi=4096,
                   p_trans=0.3):
    """One transform of the mixed-request distribution: (init, final, dtype, trans)."""
    m = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
    n = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
    dtype = int(rng.choice(dtypes))
    trans = "T" if rng.random() < p_trans else "N"
    ib = [int(rng.choice(BLOCKS)) for _ in range(2)]
    fb = [int(rng.choice(BLOCKS)) for _ in range(2)]
    ig = [int(rng.integers(1, 5)) for _ in range(2)]
    fg = [int(rng.integers(1, 5)) for _ in range(2)]
    init = scalapack_layout(m, n, ib[0], ib[1], ig[0], ig[1], "C", lld="rows")
    fm, fn = (n, m) if trans == "T" else (m, n)
    fin = scalapack_layout(fm, fn, fb[0], fb[1], fg[0], fg[1], "C", lld="rows")
    return init, fin, dtype, trans


def mixed_requests(n=2048, seed=1004, p_trans=0.3):
    rng = np.random.default_rng(seed)
    return [random_request(rng, p_trans=p_trans) for _ in range(n)]


def large_request():
    """The 5% large requests of the served stream: 16384^2 f64."""
    init = scalapack_layout(16384, 16384, 128, 128, 4, 4, "C", lld="rows")
    fin = scalapack_layout(16384, 16384, 256, 256, 2, 8, "C", lld="rows")
    return init, fin, _lib.FLOAT64, "N"


def algorithmic_bytes(layout, dtype):
    """2 * m * n * sizeof(T): one read and one write per element."""
    return 2 * layout.m * layout.n * _lib.DTYPE_SIZE[dtype]
