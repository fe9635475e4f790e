"""Pins the oracle's warm-start variants (CPU only): the pull, data-driven
pull and STANDARD functions started from a caller's rank vector instead of the
reference's start (pageRank.c:942-1108 always starts at base).

* A start equal to the reference's own (base / N, exact for N a power of two)
  reproduces the cold run bit for bit.
* k sweeps, then a warm start from the returned ranks with m more, equal
  k + m sweeps (REFERENCE FP32 bit for bit at N = 2^k; STANDARD FP64 exactly).
* One sweep from a random start equals an independent FP64 sweep of the same
  operator (scipy.sparse), so a start that is ignored, not scaled by N, or
  scaled twice fails.
"""
import numpy as np
import scipy.sparse as sp

import oracle

D32 = np.float32(0.85)
B32 = np.float32(1.0) - D32


def _graph(seed, n, m):
    rng = np.random.RandomState(seed)
    src = rng.randint(0, n, m).astype(np.uint32)
    dst = rng.randint(0, n, m).astype(np.uint32)
    return src, dst, oracle.build_csr(src, dst, n)


def _pull64(src, dst, n, pr):
    """One REFERENCE sweep in FP64 on unnormalised ranks pr (base 1-d, dangling dropped)."""
    od = np.bincount(src, minlength=n).astype(np.float64)
    P = sp.csr_matrix((1.0 / od[src.astype(np.int64)], (dst.astype(np.int64), src.astype(np.int64))),
                      shape=(n, n))
    return float(B32) + float(D32) * (P @ pr)


def test_reference_start_reproduces_cold_run():
    n = 1 << 12
    src, dst, csr = _graph(5, n, 6 * n)
    cold, it, et, ps, act, err = oracle.pagerank_pull_f32(csr, 15, 0.0)
    x0 = np.full(n, B32 / np.float32(n), np.float32)
    assert np.all(x0 * np.float32(n) == B32)
    warm, it2, et2, ps2, act2, err2 = oracle.pagerank_pull_f32(csr, 15, 0.0, init=x0)
    np.testing.assert_array_equal(warm, cold)
    assert (it, et, ps) == (it2, et2, ps2)
    np.testing.assert_array_equal(act, act2)
    dcold = oracle.pagerank_dd_pull_f32(csr, 15, 1e-4)
    dwarm = oracle.pagerank_dd_pull_f32(csr, 15, 1e-4, init=x0)
    for a, b in zip(dcold, dwarm):
        np.testing.assert_array_equal(a, b)


def test_reference_resume_equals_uninterrupted():
    n = 1 << 11
    src, dst, csr = _graph(6, n, 9 * n)
    full, *_ = oracle.pagerank_pull_f32(csr, 20, 0.0)
    part, *_ = oracle.pagerank_pull_f32(csr, 8, 0.0)
    resumed, *_ = oracle.pagerank_pull_f32(csr, 12, 0.0, init=part)
    np.testing.assert_array_equal(resumed, full)


def test_reference_one_sweep_from_random_start():
    n = 3000  # not a power of two: the start is scaled by N in FP32
    src, dst, csr = _graph(7, n, 5 * n)
    x0 = np.random.RandomState(8).rand(n).astype(np.float32) / np.float32(n)
    got, *_ = oracle.pagerank_pull_f32(csr, 1, 0.0, init=x0)
    want = _pull64(src, dst, n, x0.astype(np.float64) * n) / n
    assert np.abs(got - want).max() <= 1e-6 * want.max()


def test_data_driven_one_sweep_from_random_start():
    n = 2000
    src, dst, csr = _graph(9, n, 4 * n)
    x0 = np.random.RandomState(10).rand(n).astype(np.float32) / np.float32(n)
    got, it, et, ps, act, err, work = oracle.pagerank_dd_pull_f32(csr, 1, 0.0, init=x0)
    want = _pull64(src, dst, n, x0.astype(np.float64) * n) / n
    assert work[0] == n  # every vertex on the first worklist
    assert np.abs(got - want).max() <= 1e-6 * want.max()


def test_standard_warm_start():
    n = 2500
    src, dst, csr = _graph(11, n, 5 * n)
    full, _, _ = oracle.pagerank_standard_f64(csr, 30)
    part, _, _ = oracle.pagerank_standard_f64(csr, 10)
    resumed, _, _ = oracle.pagerank_standard_f64(csr, 20, init=part)
    np.testing.assert_array_equal(resumed, full)
    # one sweep from a random distribution against an independent FP64 sweep
    x0 = np.random.RandomState(12).rand(n)
    x0 /= x0.sum()
    got, _, res = oracle.pagerank_standard_f64(csr, 1, init=x0)
    od = np.bincount(src, minlength=n).astype(np.float64)
    P = sp.csr_matrix((1.0 / od[src.astype(np.int64)], (dst.astype(np.int64), src.astype(np.int64))),
                      shape=(n, n))
    want = 0.15 / n + 0.85 * (P @ x0 + x0[od == 0].sum() / n)
    np.testing.assert_allclose(got, want, rtol=1e-12)
    assert abs(res - np.abs(want - x0).sum()) <= 1e-12
    # the fixed point is a fixed point
    fx, _, _ = oracle.pagerank_standard_f64(csr, 500, tol=1e-15)
    again, _, res = oracle.pagerank_standard_f64(csr, 1, init=fx)
    assert res <= 1e-14
