"""Pins the oracle's locality relabelling (the reference's -l/-L/-O modes,
README.md:362-411), CPU only.

The implementation lives in the absent OpenGraph reorder module, so the exact
labels it produced are not available: the oracle restates the published
algorithms (degree sort; DBG, Faldu et al. 2019; HUBSort, Zhang et al. 2017;
HUBCluster, Balaji and Lucia 2018) with conventions stated in oracle.h, and
parity with OpenGraph's exact labels is unpinned.  What is pinned here:

* every mode yields a permutation, and relabelling is a graph isomorphism:
  pull PageRank of the relabelled graph, read back through the labels, equals
  that of the original (catches a label / inverse-label mix-up);
* the orders against independent numpy constructions (np.lexsort / stable
  argsort over np.bincount degrees, np.select groups): a wrong degree
  direction, an unstable tie order or an off-by-one group boundary fails;
* a hand-worked 4-vertex graph for every family.
"""
import numpy as np
import pytest

import oracle


def _graph(seed, n, m):
    rng = np.random.RandomState(seed)
    # skewed on both ends so that hubs and several DBG groups exist
    src = (rng.zipf(1.6, m) % n).astype(np.uint32)
    dst = ((rng.zipf(1.6, m) * 7919) % n).astype(np.uint32)
    return src, dst


def _deg(src, dst, n, kind):
    d = np.zeros(n, np.int64)
    if kind in ("out", "both"):
        d += np.bincount(src, minlength=n)
    if kind in ("in", "both"):
        d += np.bincount(dst, minlength=n)
    return d


def _labels_from_order(order):
    lab = np.empty(order.size, np.uint32)
    lab[order] = np.arange(order.size, dtype=np.uint32)
    return lab


KIND = {1: "out", 2: "in", 3: "both", 4: "out", 5: "in", 6: "out", 7: "in", 8: "out", 9: "in"}


@pytest.mark.parametrize("mode", list(range(11)))
def test_permutation_and_isomorphism(mode):
    src, dst = _graph(1, 3000, 40000)
    n = 3000
    lab = oracle.reorder_labels(src, dst, n, mode, seed=99)
    assert np.array_equal(np.sort(lab), np.arange(n, dtype=np.uint32))
    # FP64, so that the check sees the labels and not the FP32 summation order
    # of hub rows (relabelling reorders each row's in-edges)
    pr0, *_ = oracle.pagerank_pull_f64(oracle.build_csr(src, dst, n), 20, 0.0)
    pr1, *_ = oracle.pagerank_pull_f64(oracle.build_csr(lab[src], lab[dst], n), 20, 0.0)
    np.testing.assert_allclose(pr1[lab], pr0, rtol=1e-12)


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_degree_sort_matches_lexsort(mode):
    src, dst = _graph(2, 2000, 30000)
    n = 2000
    d = _deg(src, dst, n, KIND[mode])
    order = np.lexsort((np.arange(n), -d))  # degree descending, then id
    np.testing.assert_array_equal(oracle.reorder_labels(src, dst, n, mode), _labels_from_order(order))


@pytest.mark.parametrize("mode", [4, 5])
def test_dbg_groups(mode):
    src, dst = _graph(3, 5000, 80000)
    n = 5000
    d = _deg(src, dst, n, KIND[mode])
    avg = len(src) // n
    group = np.select([d >= 32 * avg, d >= 16 * avg, d >= 8 * avg, d >= 4 * avg, d >= 2 * avg,
                       d >= avg, 2 * d >= avg], [0, 1, 2, 3, 4, 5, 6], 7)
    assert len(np.unique(group)) >= 5
    order = np.argsort(group, kind="stable")
    np.testing.assert_array_equal(oracle.reorder_labels(src, dst, n, mode), _labels_from_order(order))


@pytest.mark.parametrize("mode", [6, 7, 8, 9])
def test_hub_orders(mode):
    src, dst = _graph(4, 4000, 50000)
    n = 4000
    d = _deg(src, dst, n, KIND[mode])
    avg = len(src) // n
    hub = d > avg
    ids = np.arange(n)
    if mode in (6, 7):
        hubs = ids[hub][np.lexsort((ids[hub], -d[hub]))]
    else:
        hubs = ids[hub]
    order = np.concatenate([hubs, ids[~hub]])
    np.testing.assert_array_equal(oracle.reorder_labels(src, dst, n, mode), _labels_from_order(order))


def test_hand_worked_graph():
    # out-degrees 3, 1, 1, 2 (avg 7 // 4 = 1); in-degrees 1, 1, 3, 2
    src = np.array([0, 0, 0, 1, 2, 3, 3], np.uint32)
    dst = np.array([1, 2, 3, 2, 3, 0, 2], np.uint32)
    for mode in (1, 4, 6, 8):  # order 0, 3, 1, 2 by out-degree / hubs {0, 3}
        assert list(oracle.reorder_labels(src, dst, 4, mode)) == [0, 2, 3, 1]
    for mode in (2, 5, 7, 9):  # by in-degree: 2, 3, 0, 1 (hubs {2, 3})
        assert list(oracle.reorder_labels(src, dst, 4, mode)) == [2, 3, 0, 1]
    # in+out: 4, 2, 4, 4 -> 0, 2, 3 then 1
    assert list(oracle.reorder_labels(src, dst, 4, 3)) == [0, 3, 1, 2]
    assert list(oracle.reorder_labels(src, dst, 4, 0)) == [0, 1, 2, 3]


def test_random_mode_is_seeded():
    src, dst = _graph(5, 1000, 5000)
    a = oracle.reorder_labels(src, dst, 1000, 10, seed=1)
    b = oracle.reorder_labels(src, dst, 1000, 10, seed=1)
    c = oracle.reorder_labels(src, dst, 1000, 10, seed=2)
    np.testing.assert_array_equal(a, b)
    assert (a != c).mean() > 0.9
    with pytest.raises(ValueError):
        oracle.reorder_labels(src, dst, 1000, 11)
