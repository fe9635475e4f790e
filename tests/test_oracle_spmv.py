"""Pins the oracle's SPMV pull (SPMVPullGraphCSR, SPMV.c:622-750, unweighted),
CPU only.

* Against an independent FP64 scatter over the raw edge list (numpy
  bincount, no CSR): a transposed operand (forward instead of inverse CSR), a
  dropped or repeated edge, or an index slip fails it.
* Unit input vectors: y = e_u gives, at every v, the multiplicity of u -> v
  (exact small integers in FP32).
* Linearity in x, to FP32 rounding.
* The benchmark input (SPMV.c:699-706) on a hand-built graph, and the printed
  checksum's two-decimal rounding (:735-739) on hand-picked values.
"""
import numpy as np
import pytest

import oracle


def _graph(seed, n, m):
    rng = np.random.RandomState(seed)
    src = rng.randint(0, n, m).astype(np.uint32)
    dst = rng.randint(0, n, m).astype(np.uint32)
    return src, dst, oracle.build_csr(src, dst, n)


@pytest.mark.parametrize("seed,n,m", [(1, 300, 5000), (2, 2000, 3000), (3, 17, 400)])
def test_against_edge_list_scatter(seed, n, m):
    src, dst, csr = _graph(seed, n, m)
    x = np.random.RandomState(seed + 10).rand(n).astype(np.float32)
    y = oracle.spmv_pull_f32(csr, x)
    y64 = np.bincount(dst, weights=x[src].astype(np.float64), minlength=n)
    indeg = np.bincount(dst, minlength=n)
    # sequential FP32 sum of k positive terms: |error| <= k * 2^-24 * sum
    assert np.all(np.abs(y - y64) <= (indeg + 1) * 2.0 ** -24 * y64 + 1e-30)
    # the transposed product is a different vector on these graphs
    ysrc = np.bincount(src, weights=x[dst].astype(np.float64), minlength=n)
    assert np.abs(ysrc - y64).max() > 1e-3


def test_unit_vectors_count_multiplicity():
    src = np.array([0, 0, 0, 1, 2, 2, 3], np.uint32)
    dst = np.array([1, 1, 2, 2, 0, 2, 3], np.uint32)
    csr = oracle.build_csr(src, dst, 4)
    for u in range(4):
        e = np.zeros(4, np.float32)
        e[u] = 1.0
        y = oracle.spmv_pull_f32(csr, e)
        want = np.array([np.sum((src == u) & (dst == v)) for v in range(4)], np.float32)
        np.testing.assert_array_equal(y, want)


def test_linearity():
    src, dst, csr = _graph(4, 1000, 20000)
    rng = np.random.RandomState(5)
    a, b = rng.rand(1000).astype(np.float32), rng.rand(1000).astype(np.float32)
    ya, yb = oracle.spmv_pull_f32(csr, a), oracle.spmv_pull_f32(csr, b)
    yab = oracle.spmv_pull_f32(csr, 2 * a + 3 * b)
    np.testing.assert_allclose(yab, 2 * ya + 3 * yb, rtol=1e-5)


def test_benchmark_input_and_checksum():
    # out-degrees 3, 1, 0, 2
    src = np.array([0, 0, 0, 1, 3, 3], np.uint32)
    dst = np.array([1, 2, 3, 0, 0, 1], np.uint32)
    csr = oracle.build_csr(src, dst, 4)
    x = oracle.spmv_input(csr)
    np.testing.assert_array_equal(
        x, np.array([np.float32(1) / np.float32(3), 1.0, 0.001, 0.5], np.float32))
    y = oracle.spmv_pull_f32(csr, x)
    # y0 = x1 + x3, y1 = x0 + x3, y2 = x0, y3 = x0
    np.testing.assert_allclose(y, [1.5, 1 / 3 + 0.5, 1 / 3, 1 / 3], rtol=1e-7)
    # 1.234 -> 1.23, 2.996 -> 3.00, 0.004 -> 0.00, 7.0 -> 7.00
    assert oracle.spmv_sum(np.array([1.234, 2.996, 0.004, 7.0], np.float32)) == \
        pytest.approx(11.23, abs=1e-12)
