"""Layout pairs shared by the GPU parity tests and the host replay of the
gather's indexing (tests/test_gpu_transform.py, tests/test_gather_replay.py)."""
import numpy as np

from grid2grid_amd.layout import scalapack_layout, arbitrary_layout


def layout_pairs():
    rng = np.random.default_rng(101)
    return {
        "bc_small": (scalapack_layout(37, 29, 4, 3, 2, 2, "C"), scalapack_layout(37, 29, 5, 7, 2, 2, "C")),
        "bc_grid_change": (scalapack_layout(100, 77, 8, 8, 2, 3, "R"), scalapack_layout(100, 77, 3, 5, 3, 1, "C")),
        "bc_4096_blocks": (scalapack_layout(256, 256, 32, 32, 2, 2, "C", lld="rows"),
                          scalapack_layout(256, 256, 64, 128, 2, 2, "C", lld="rows")),
        "arbitrary": (arbitrary_layout(rng, 300, 260, 7, 6, 5, lo=10, hi=80, pad=(0, 9), gap=(0, 33)),
                      arbitrary_layout(rng, 300, 260, 5, 8, 6, lo=10, hi=80, pad=(0, 9), gap=(0, 33))),
        "single_rank": (scalapack_layout(64, 48, 64, 48, 1, 1), scalapack_layout(64, 48, 16, 16, 1, 1)),
        "tall_thin": (scalapack_layout(1000, 3, 17, 1, 4, 1), scalapack_layout(1000, 3, 300, 2, 2, 2)),
        # more ranks than the LDS pair table holds: the plain form's wire
        # order comes from a radix sort by rank pair
        "many_ranks": (arbitrary_layout(rng, 240, 180, 16, 12, 130, lo=5, hi=25, pad=(0, 5), gap=(0, 9)),
                       arbitrary_layout(rng, 240, 180, 13, 15, 100, lo=5, hi=25, pad=(0, 5), gap=(0, 9))),
    }


def transposed_pairs():
    rng = np.random.default_rng(202)
    pairs = {
        "bcT": (scalapack_layout(70, 93, 8, 5, 2, 2, "C"), scalapack_layout(93, 70, 6, 9, 2, 2, "C")),
        "arbT": (arbitrary_layout(rng, 130, 210, 4, 5, 4, lo=10, hi=80, pad=(0, 6), gap=(0, 17)),
                 arbitrary_layout(rng, 210, 130, 6, 3, 3, lo=10, hi=80, pad=(0, 6), gap=(0, 17))),
        "bcT_big_tiles": (scalapack_layout(160, 96, 64, 48, 2, 1, "C"), scalapack_layout(96, 160, 40, 64, 1, 2, "C")),
        # pieces of 100 and 60 final rows by 128 and 44 final columns: 8- and
        # 16-byte elements take LDS tiles for the 100-row pieces and the
        # element queue for the 60-row ones; f32 pieces (under 128 rows) all
        # take the element queue
        "bcT_mixed_paths": (scalapack_layout(300, 260, 128, 100, 2, 2, "C"),
                            scalapack_layout(260, 300, 100, 128, 2, 2, "C")),
        # pieces of 8-24 final rows by up to 96 final columns: the element
        # queue (register tiles for 4- and 8-byte elements, rows for 16-byte ones)
        "bcT_short_tiles": (scalapack_layout(400, 300, 96, 40, 2, 2, "C"),
                            scalapack_layout(300, 400, 24, 160, 2, 2, "C")),
        # pieces of 16-40 final rows by up to 200 final columns: the element queue
        "bcT_mid_tiles": (scalapack_layout(400, 300, 200, 56, 2, 2, "C"),
                          scalapack_layout(300, 400, 40, 256, 2, 2, "C")),
    }
    # the cut between the element queue and LDS tiles: final row blocks of
    # 127, 128, 63 and 64 rows (f32 127/128 rows = 508/512 bytes, 8-byte
    # elements 63/64 rows) by final column blocks of 63, 64 and 100 columns
    # (f32 column ranges of 252/256 bytes), one piece per block
    frs = [0, 127, 255, 318, 382]
    fcs = [0, 63, 127, 227]
    init = arbitrary_layout(rng, 227, 382, 3, 4, 3, pad=(0, 3), gap=(0, 5), rows_split=fcs, cols_split=frs)
    fin = arbitrary_layout(rng, 382, 227, 4, 3, 4, pad=(0, 3), gap=(0, 5), rows_split=frs, cols_split=fcs)
    pairs["bcT_cut_edges"] = (init, fin)
    return pairs


def ragged_pair(dtype):
    """Pieces of 1-11 rows (4 to 88 bytes a column) with ld padding and gaps."""
    rng = np.random.default_rng(515 + dtype)
    init = arbitrary_layout(rng, 150, 70, 27, 5, 3, lo=1, hi=11, pad=(0, 3), gap=(0, 5))
    fin = arbitrary_layout(rng, 150, 70, 23, 6, 4, lo=1, hi=11, pad=(0, 3), gap=(0, 5))
    return init, fin, rng


def irregular_small():
    """The irregular benchmark pair's construction at 6000 x 400: row blocks
    of 40-520 rows (f32 columns of 160 B to 2 KB), ld = rows + U[0, 63],
    blocks shuffled with gaps, so columns start at every dword offset of a
    128-byte line."""
    rng = np.random.default_rng(303)
    init = arbitrary_layout(rng, 6000, 400, 16, 5, 6, lo=40, hi=520, pad=(0, 63), gap=(0, 64))
    fin = arbitrary_layout(rng, 6000, 400, 13, 4, 5, lo=40, hi=520, pad=(0, 63), gap=(0, 64))
    return init, fin


def random_stream_cases(seed=4242, count=36):
    """Random requests (block-cyclic from the mixed-request distribution and
    irregular arbitrary layouts, every type, 'N'/'T'/'C') with random
    tunings: (init, final, dtype, trans, tuning), tuning = (tile_bytes,
    threads, unroll_bytes, nontemporal, launch_mode, column_chunks,
    stream_order, row_chunks, tail_split)."""
    from grid2grid_amd import workloads
    from grid2grid_amd import _lib
    dtypes = [_lib.FLOAT32, _lib.FLOAT64, _lib.COMPLEX64, _lib.COMPLEX128]
    rng = np.random.default_rng(seed)
    cases = []
    for k in range(count):
        if k % 3 == 2:
            m, n = (int(x) for x in rng.integers(20, 400, size=2))
            p1, p2 = (int(x) for x in rng.integers(1, 9, size=2))
            init = arbitrary_layout(rng, m, n, int(rng.integers(1, 8)), int(rng.integers(1, 8)), p1, lo=3, hi=90,
                                    pad=(0, 9), gap=(0, 17))
            trans = "N" if rng.random() < 0.6 else "T"
            fm, fn = (n, m) if trans == "T" else (m, n)
            fin = arbitrary_layout(rng, fm, fn, int(rng.integers(1, 8)), int(rng.integers(1, 8)), p2, lo=3, hi=90,
                                   pad=(0, 9), gap=(0, 17))
            dtype = dtypes[int(rng.integers(0, 4))]
        else:
            init, fin, dtype, trans = workloads.random_request(rng, dtypes=dtypes, lo=8, hi=400, p_trans=0.4)
        if trans == "T" and dtype in (_lib.COMPLEX64, _lib.COMPLEX128) and rng.random() < 0.5:
            trans = "C"
        tile = int(rng.choice([4096, 12288, 40960, 49152, 65536, 98304, 1 << 20]))
        th, ub, nt, mode = (int(rng.choice([256, 512])), int(rng.choice([64, 128])), int(rng.integers(0, 2)),
                            int(rng.integers(0, 4)))
        w = int(rng.choice([0, 1, 2, 3, 8, 16, 64]))
        order = int(rng.integers(0, 2))
        h = int(rng.choice([0, 0, 1, 5, 16, 128])) if trans == "N" else 0
        sp = int(rng.choice([0, 0, 2, 4, 8]))
        cases.append((init, fin, dtype, trans, (tile, th, ub, nt, mode, w, order, h, sp)))
    return cases
