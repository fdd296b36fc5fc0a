"""Communication volume between ranks and the greedy rank relabelling.

The oracle (orc_comm_volume, orc_optimal_reordering) is pinned by a brute
force over elements, by the reference's own optimal_reordering and cover
(oracle/_ref), and by matching properties; the product (device
g2g_comm_volume, host g2g_optimal_reordering) is then checked against the
oracle."""
import itertools

import numpy as np
import pytest

import oracle_lib as O
import grid2grid_amd as g2g
from grid2grid_amd.layout import Layout, scalapack_layout, arbitrary_layout

needs_ref = pytest.mark.skipif(O.ref() is None, reason="reference harness (oracle/_ref) not built")


def _pairs():
    rng = np.random.default_rng(31)
    out = [
        ("bc_2x2", scalapack_layout(37, 29, 4, 3, 2, 2, "C"), scalapack_layout(37, 29, 5, 7, 2, 2, "C")),
        ("bc_grid_change", scalapack_layout(40, 33, 8, 8, 2, 3, "R"), scalapack_layout(40, 33, 3, 5, 3, 1, "C")),
        ("bc_ranks_differ", scalapack_layout(64, 48, 8, 8, 4, 2, "C"), scalapack_layout(64, 48, 16, 12, 1, 3, "R")),
        ("arbitrary", arbitrary_layout(rng, 45, 38, 5, 4, 5, lo=3, hi=14),
         arbitrary_layout(rng, 45, 38, 4, 6, 3, lo=3, hi=14)),
        ("arbitrary_many", arbitrary_layout(rng, 90, 70, 9, 7, 11, lo=3, hi=14),
         arbitrary_layout(rng, 90, 70, 6, 8, 7, lo=3, hi=14)),
        ("single_rank", scalapack_layout(20, 10, 20, 10, 1, 1), scalapack_layout(20, 10, 4, 4, 1, 1)),
    ]
    return out


def brute_volume(init, fin):
    """Element-by-element count: for every element, the pair {initial owner,
    final owner} gains one."""
    n = max(init.n_ranks, fin.n_ranks)
    vol = np.zeros((n, n), dtype=np.int64)
    ri = np.searchsorted(init.rows_split, np.arange(init.m), side="right") - 1
    ci = np.searchsorted(init.cols_split, np.arange(init.n), side="right") - 1
    rf = np.searchsorted(fin.rows_split, np.arange(fin.m), side="right") - 1
    cf = np.searchsorted(fin.cols_split, np.arange(fin.n), side="right") - 1
    a = init.owners[ri[:, None], ci[None, :]]
    b = fin.owners[rf[:, None], cf[None, :]]
    lo, hi = np.minimum(a, b).ravel(), np.maximum(a, b).ravel()
    np.add.at(vol, (lo, hi), 1)
    return vol


@pytest.mark.parametrize("idx", range(len(_pairs())))
def test_volume_oracle_brute_force(idx):
    name, init, fin = _pairs()[idx]
    vol = O.comm_volume(O.oracle().orc_comm_volume, init, fin)
    np.testing.assert_array_equal(vol, brute_volume(init, fin))
    assert vol.sum() == init.m * init.n
    assert np.all(np.tril(vol, -1) == 0)


@needs_ref
@pytest.mark.parametrize("idx", range(len(_pairs())))
def test_volume_oracle_matches_reference(idx):
    name, init, fin = _pairs()[idx]
    np.testing.assert_array_equal(O.comm_volume(O.oracle().orc_comm_volume, init, fin),
                                  O.comm_volume(O.ref().ref_comm_volume, init, fin))


def test_volume_brute_force_catches_a_swapped_owner():
    """A plausible mistake (one owner relabelled) changes the brute-force
    volume, so the check above would see it."""
    name, init, fin = _pairs()[0]
    owners = init.owners.copy()
    owners[0, 0] = (owners[0, 0] + 1) % init.n_ranks
    bad = Layout(init.rows_split, init.cols_split, owners, init.n_ranks, init.blk_row, init.blk_col,
                 init.blk_stride, init.blk_offset, init.arena_elems)
    assert not np.array_equal(O.comm_volume(O.oracle().orc_comm_volume, bad, fin), brute_volume(init, fin))


# ------------------------------------------------------------ reordering

def _determined_volume(rng, n):
    """A volume on which the reference's result does not depend on its hash
    map: every rank keeps elements, adjusted weights above 1 are distinct and
    none equals 1."""
    while True:
        vol = np.zeros((n, n), dtype=np.int64)
        for a in range(n):
            vol[a, a] = rng.integers(1, 60)
            for b in range(a + 1, n):
                if rng.random() < 0.6:
                    vol[a, b] = rng.integers(1, 400)
        adj = [int(vol[a, b] - vol[a, a] - vol[b, b]) for a in range(n) for b in range(a + 1, n) if vol[a, b]]
        big = [w for w in adj if w > 1]
        if 1 not in adj and len(big) == len(set(big)):
            return vol


def _check_matching(vol, perm):
    n = vol.shape[0]
    assert sorted(perm) == list(range(n))
    assert all(perm[perm[a]] == a for a in range(n))  # an involution: pairs swapped


@needs_ref
@pytest.mark.parametrize("seed", range(12))
def test_reordering_oracle_matches_reference(seed):
    rng = np.random.default_rng(500 + seed)
    vol = _determined_volume(rng, int(rng.integers(2, 13)))
    got = O.optimal_reordering(O.oracle().orc_optimal_reordering, vol)
    want = O.optimal_reordering(O.ref().ref_optimal_reordering, vol)
    assert got == want
    _check_matching(vol, got[0])


def test_reordering_oracle_greedy_by_hand():
    # ranks 0..3; adjusted: {0,1}: 100-5-5 = 90, {2,3}: 50-5-5 = 40,
    # {1,2}: 300-5-5 = 290 (taken first, blocks {0,1} and {2,3})
    vol = np.zeros((4, 4), dtype=np.int64)
    for a in range(4):
        vol[a, a] = 5
    vol[0, 1], vol[2, 3], vol[1, 2] = 100, 50, 300
    perm, reordered = O.optimal_reordering(O.oracle().orc_optimal_reordering, vol)
    assert perm == [0, 2, 1, 3] and reordered
    # with no profitable pair the permutation is the identity
    vol2 = np.diag([7, 7, 7]).astype(np.int64)
    vol2[0, 1] = 3
    assert O.optimal_reordering(O.oracle().orc_optimal_reordering, vol2) == ([0, 1, 2], False)


def test_reordering_oracle_is_greedy_maximal():
    """No edge of adjusted weight above 1 is left with both ends free."""
    rng = np.random.default_rng(77)
    for _ in range(20):
        n = int(rng.integers(2, 10))
        vol = _determined_volume(rng, n)
        perm, _ = O.optimal_reordering(O.oracle().orc_optimal_reordering, vol)
        _check_matching(vol, perm)
        free = [a for a in range(n) if perm[a] == a]
        for a, b in itertools.combinations(free, 2):
            w = vol[a, b] - vol[a, a] - vol[b, b] if vol[a, b] else 0
            # a free pair can only be left out when a self edge (weight 1)
            # claimed an end first, i.e. its own weight is at most 1
            assert w <= 1


@pytest.mark.parametrize("seed", range(8))
def test_reordering_product_matches_oracle(seed):
    """Host-only entry point g2g_optimal_reordering (no device work)."""
    rng = np.random.default_rng(900 + seed)
    n = int(rng.integers(1, 16))
    vol = np.zeros((n, n), dtype=np.int64)
    for a in range(n):
        if rng.random() < 0.8:
            vol[a, a] = rng.integers(0, 40)
        for b in range(a + 1, n):
            if rng.random() < 0.5:
                vol[a, b] = rng.integers(1, 120)   # ties and negative weights allowed here
    want = O.optimal_reordering(O.oracle().orc_optimal_reordering, vol)
    d = O.volume_dict(vol)
    assert g2g.optimal_reordering(d, n) == want
    # edges given as (dest, src) and split in two repeats sum to the same volume
    d2 = {}
    for (a, b), w in d.items():
        d2[(b, a)] = w // 2
    items = list(d2.items()) + [((a, b), w - w // 2) for (a, b), w in d.items()]
    lib = g2g.load()
    import ctypes
    edges = (g2g._lib.G2GEdge * max(1, len(items)))()
    for k, ((a, b), w) in enumerate(items):
        edges[k].src, edges[k].dest, edges[k].weight = a, b, w
    perm = (ctypes.c_int32 * n)()
    r = ctypes.c_int32()
    assert lib.g2g_optimal_reordering(edges, len(items), n, perm, ctypes.byref(r)) == 0
    assert ([int(x) for x in perm], bool(r.value)) == want


def test_reordering_product_rejects_bad_edges():
    with pytest.raises(g2g.G2GError):
        g2g.optimal_reordering({(0, 5): 3}, 4)
    with pytest.raises(g2g.G2GError):
        g2g.optimal_reordering({}, 0)


# ------------------------------------------------------------ device volume

@pytest.mark.gpu
@pytest.mark.parametrize("idx", range(len(_pairs())))
def test_volume_device_matches_oracle(idx):
    name, init, fin = _pairs()[idx]
    want = O.volume_dict(O.comm_volume(O.oracle().orc_comm_volume, init, fin))
    got = g2g.comm_volume(init, fin)
    assert got == want
    assert list(got) == sorted(got)


@pytest.mark.gpu
def test_volume_device_large_grids():
    """Block-cyclic 32768^2 grids (128x128 on 4x4 -> 256x512 on 2x8): 16
    ranks; the oracle's dense walk is fast enough at this size."""
    init = scalapack_layout(32768, 32768, 128, 128, 4, 4, "C", lld="rows")
    fin = scalapack_layout(32768, 32768, 256, 512, 2, 8, "C", lld="rows")
    want = O.volume_dict(O.comm_volume(O.oracle().orc_comm_volume, init, fin))
    got = g2g.comm_volume(init, fin)
    assert got == want
    assert sum(got.values()) == 32768 * 32768
    perm, _ = g2g.optimal_reordering(got, 16)
    assert (perm, _) == O.optimal_reordering(O.oracle().orc_optimal_reordering,
                                             O.comm_volume(O.oracle().orc_comm_volume, init, fin))


@pytest.mark.gpu
def test_volume_device_rejects_mismatched_grids():
    init = scalapack_layout(40, 30, 8, 8, 2, 2)
    fin = scalapack_layout(40, 31, 8, 8, 2, 2)
    with pytest.raises(g2g.G2GError):
        g2g.comm_volume(init, fin)
