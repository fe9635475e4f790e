"""Pins the oracle's CSR builder (CPU only).

The reference's own builder (OpenGraph) is absent, so the oracle is pinned by
checks that do not restate its code:
  * numpy's stable argsort (an independent sort) for the forward order;
  * lexicographic (dst, src) order for the inverse built from the
    forward-sorted list (SSSP.c:518-520 + a stable sort), and stable-by-dst for
    the inverse built from the input order;
  * brute force over every edge list of <= 4 edges on 3 vertices;
  * the closed forms degree = histogram, edges_idx = segment start (0 if empty).
A plausible mistake -- an unstable scatter, swapped endpoints, off-by-one
offsets, a forward degree used for the inverse -- fails one of these.
"""
import itertools

import numpy as np
import pytest

import oracle


def _expect(src, dst, n, mode=0, from_input=False):
    src = np.asarray(src, np.uint32)
    dst = np.asarray(dst, np.uint32)
    if mode in (1, 3):
        p = np.lexsort((dst, src))
    else:
        p = np.argsort(src, kind="stable")
    fs, fd = src[p], dst[p]
    if from_input:
        if mode in (1, 3):
            q = np.lexsort((src, dst))
        else:
            q = np.argsort(dst, kind="stable")
        is_, id_ = dst[q], src[q]
    else:
        q = np.lexsort((fs, fd))  # stable by new src over a list sorted by old src
        is_, id_ = fd[q], fs[q]
    out = {
        "sorted_edges_array.edges_array_src": fs,
        "sorted_edges_array.edges_array_dest": fd,
        "inverse_sorted_edges_array.edges_array_src": is_,
        "inverse_sorted_edges_array.edges_array_dest": id_,
    }
    for side, key in (("vertices", fs), ("inverse_vertices", is_)):
        deg = np.bincount(key, minlength=n).astype(np.uint32)[:n]
        idx = np.zeros(n, np.uint32)
        starts = np.concatenate([[0], np.cumsum(deg)[:-1]]).astype(np.uint32)
        idx[deg > 0] = starts[deg > 0]
        out[f"{side}.out_degree"] = deg
        out[f"{side}.edges_idx"] = idx
    out["vertices.in_degree"] = out["inverse_vertices.out_degree"]
    out["inverse_vertices.in_degree"] = out["vertices.out_degree"]
    return out


def _check(src, dst, n, mode=0, from_input=False):
    got = oracle.build_csr(src, dst, n, mode, from_input)
    want = _expect(src, dst, n, mode, from_input)
    for k in oracle.CSR_KEYS:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)


def test_count_sort_matches_stable_argsort():
    rng = np.random.RandomState(1)
    for n, m in [(1, 5), (7, 100), (1000, 20000), (65536, 100000)]:
        key = rng.randint(0, n, m).astype(np.uint32)
        val = np.arange(m, dtype=np.uint32)  # value = input position: exposes instability
        ko, vo = oracle.count_sort(n, key, val)
        p = np.argsort(key, kind="stable")
        np.testing.assert_array_equal(ko, key[p])
        np.testing.assert_array_equal(vo, p.astype(np.uint32))


def test_brute_force_tiny_graphs():
    n = 3
    edges = [(u, v) for u in range(n) for v in range(n)]
    count = 0
    for m in range(0, 5):
        for combo in itertools.product(edges, repeat=m):
            src = [e[0] for e in combo]
            dst = [e[1] for e in combo]
            for mode in (0, 1):
                _check(src, dst, n, mode)
            _check(src, dst, n, 0, from_input=True)
            count += 1
    assert count == 1 + 9 + 81 + 729 + 6561


def test_random_small_graphs_all_modes():
    """4,000 random edge lists on at most 6 vertices with at most 10 edges
    (duplicates and self-loops included), every sort mode and both inverse
    conventions, against the closed forms."""
    rng = np.random.RandomState(77)
    for _ in range(500):
        n = rng.randint(1, 7)
        m = rng.randint(0, 11)
        src = rng.randint(0, n, m)
        dst = rng.randint(0, n, m)
        for mode in range(4):
            for from_input in (False, True):
                _check(src, dst, n, mode, from_input)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("from_input", [False, True])
def test_random_graphs_all_modes(mode, from_input):
    rng = np.random.RandomState(100 + mode)
    for n, m in [(50, 1000), (3000, 40000)]:
        src = rng.randint(0, n, m)
        dst = rng.randint(0, n, m)
        _check(src, dst, n, mode, from_input)


def test_invariants_uniform_c1_shape():
    rng = np.random.RandomState(3)
    n, m = 65536, 1 << 20
    src = rng.randint(0, n, m).astype(np.uint32)
    dst = rng.randint(0, n, m).astype(np.uint32)
    c = oracle.build_csr(src, dst, n)
    for side, arr in (("vertices", "sorted_edges_array"),
                      ("inverse_vertices", "inverse_sorted_edges_array")):
        deg = c[f"{side}.out_degree"].astype(np.int64)
        idx = c[f"{side}.edges_idx"].astype(np.int64)
        assert deg.sum() == m
        nz = deg > 0
        assert np.all(np.diff(idx[nz]) > 0)  # segments in vertex order
        key = c[f"{arr}.edges_array_src"]
        assert np.all(np.diff(key.astype(np.int64)) >= 0)
        np.testing.assert_array_equal(key[idx[nz]], np.nonzero(nz)[0])
    # multiset of edges preserved in both directions
    fw = np.sort(c["sorted_edges_array.edges_array_src"].astype(np.uint64) << 32 |
                 c["sorted_edges_array.edges_array_dest"])
    iv = np.sort(c["inverse_sorted_edges_array.edges_array_dest"].astype(np.uint64) << 32 |
                 c["inverse_sorted_edges_array.edges_array_src"])
    inp = np.sort(src.astype(np.uint64) << 32 | dst)
    np.testing.assert_array_equal(fw, inp)
    np.testing.assert_array_equal(iv, inp)


def test_num_vertices_rule():
    assert oracle.num_vertices([], []) == 0
    assert oracle.num_vertices([3, 1], [0, 9]) == 10


def test_planted_mistake_is_caught():
    """An oracle that sorted unstably would fail _check: plant one on the CPU."""
    src = np.array([1, 0, 1, 0], np.uint32)
    dst = np.array([5, 6, 7, 8], np.uint32)
    want = _expect(src, dst, 9)
    bad = dict(oracle.build_csr(src, dst, 9))
    bad["sorted_edges_array.edges_array_dest"] = np.array([8, 6, 5, 7], np.uint32)
    assert not np.array_equal(bad["sorted_edges_array.edges_array_dest"],
                              want["sorted_edges_array.edges_array_dest"])
