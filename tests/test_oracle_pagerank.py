"""Pins the oracle's PageRank functions (CPU only).

* FP32 restatement of pageRankPullGraphCSR (pageRank.c:942-1108):
  closed forms (directed cycle, two-vertex path, star, self-loop), the
  iteration-count semantics of the epsilon break (pageRank.c:1073), and
  agreement with the FP64 oracle to FP32 rounding.
* FP64 oracle: an independent scipy.sparse power iteration of the same
  operator (reference formulation: base 1-d, dangling mass dropped).
* STANDARD mode and PPR: scipy.sparse iteration / dense linear solve of the
  fixed point, and mass conservation.
A plausible mistake -- inverse out-degree used for contributions, damping on
the wrong term, the break counted as an iteration, missing dangling mass --
fails one of these.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle

D32 = np.float32(0.85)
B32 = np.float32(1.0) - D32


def _csr(src, dst, n):
    return oracle.build_csr(np.asarray(src, np.uint32), np.asarray(dst, np.uint32), n)


def _matrix(src, dst, n):
    """P[v, u] = (#edges u->v) / outdeg(u): the pull operator, built independently."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    outdeg = np.bincount(src, minlength=n).astype(np.float64)
    w = 1.0 / outdeg[src]
    return sp.csr_matrix((w, (dst, src)), shape=(n, n)), outdeg


def test_cycle_closed_form():
    n, k = 10, 7
    src = np.arange(n)
    dst = (src + 1) % n
    pr, it, err, psum, act, errs = oracle.pagerank_pull_f32(_csr(src, dst, n), k, 0.0)
    x = np.float64(1.0)
    # x_k = 1 - 0.85^(k+1) (FP64 closed form), all vertices equal, then /N
    closed = 1.0 - float(D32) ** (k + 1)
    assert it == k
    assert np.allclose(pr, closed / n, rtol=2e-6)
    assert np.all(pr == pr[0])
    assert abs(psum - closed) < 1e-5
    # per-sweep mean |delta| = 0.15 * 0.85^i
    for i, e in enumerate(errs):
        assert abs(e - float(B32) * float(D32) ** (i + 1)) < 1e-6
    del x


def test_two_vertex_and_star():
    # u -> v: v = base + d*base; u keeps base (no in-edges); v's mass is dropped
    pr, it, *_ = oracle.pagerank_pull_f32(_csr([0], [1], 2), 5, 0.0)
    assert pr[0] == np.float32(B32 / np.float32(2))
    assert pr[1] == np.float32((B32 + D32 * B32) / np.float32(2))
    # star: leaves 1..n-1 -> hub 0
    n = 9
    pr, it, *_ = oracle.pagerank_pull_f32(_csr(np.arange(1, n), np.zeros(n - 1), n), 4, 0.0)
    hub = 0.15 + 0.85 * 0.15 * (n - 1)
    assert abs(pr[0] * n - hub) < 1e-5
    assert np.allclose(pr[1:] * n, 0.15, rtol=1e-6)


def test_forward_outdegree_divides_contribution():
    # 0 -> 1, 0 -> 2, 1 -> 2: vertex 0 splits its rank over 2 out-edges
    pr, *_ = oracle.pagerank_pull_f32(_csr([0, 0, 1], [1, 2, 2], 3), 1, 0.0)
    c0 = B32 / np.float32(2)
    assert pr[1] * 3 == pytest.approx(float(B32 + D32 * c0), rel=1e-7)
    assert pr[2] * 3 == pytest.approx(float(B32 + D32 * (c0 + B32)), rel=1e-7)


def test_epsilon_break_not_counted():
    # cycle: sweep i changes every vertex by 0.15 * 0.85^(i+1); the first
    # sweep with change < eps has active == 0 and breaks without ++iterations.
    n = 4
    src = np.arange(n)
    dst = (src + 1) % n
    eps = 1e-3
    pr, it, err, psum, act, errs = oracle.pagerank_pull_f32(_csr(src, dst, n), 100, eps)
    first_quiet = next(i for i in range(100) if 0.15 * 0.85 ** (i + 1) < eps)
    assert it == first_quiet
    assert act[-1] == 0 and np.all(act[:-1] == n)


def test_fp64_oracle_matches_scipy_power_iteration():
    rng = np.random.RandomState(5)
    n, m = 500, 4000
    src = rng.randint(0, n, m)
    dst = rng.randint(0, n, m)
    P, outdeg = _matrix(src, dst, n)
    d, base = float(D32), float(B32)
    x = np.full(n, base)
    for _ in range(30):
        x = base + d * (P @ x)
    got, it, _ = oracle.pagerank_pull_f64(_csr(src, dst, n), 30, 0.0)
    assert it == 30
    np.testing.assert_allclose(got, x / n, rtol=1e-12, atol=0)


def test_fp32_restatement_tracks_fp64():
    rng = np.random.RandomState(6)
    n, m = 2000, 30000
    src = rng.randint(0, n, m)
    dst = rng.randint(0, n, m)
    c = _csr(src, dst, n)
    f32, it32, *_ = oracle.pagerank_pull_f32(c, 20, 0.0)
    f64, it64, _ = oracle.pagerank_pull_f64(c, 20, 0.0)
    assert it32 == it64 == 20
    rel = np.abs(f32 - f64) / f64
    assert rel.max() < 1e-5


def test_standard_mode_matches_scipy_and_conserves_mass():
    rng = np.random.RandomState(8)
    n, m = 300, 1500
    src = rng.randint(0, n, m)
    dst = rng.randint(0, n, m)
    P, outdeg = _matrix(src, dst, n)
    dang = outdeg == 0
    d = 0.85
    x = np.full(n, 1.0 / n)
    for _ in range(25):
        x = (1 - d) / n + d * (P @ x + x[dang].sum() / n)
    got, it, res = oracle.pagerank_standard_f64(_csr(src, dst, n), 25, 0.0, d)
    np.testing.assert_allclose(got, x, rtol=1e-12)
    assert abs(got.sum() - 1.0) < 1e-12
    # all-dangling graph stays uniform
    g2, _, _ = oracle.pagerank_standard_f64(_csr([], [], 5), 10, 0.0, d)
    np.testing.assert_allclose(g2, 0.2, rtol=1e-15)


def test_ppr_fixed_point_dense_solve():
    rng = np.random.RandomState(9)
    n, m = 200, 900
    src = rng.randint(0, n, m)
    dst = rng.randint(0, n, m)
    P, outdeg = _matrix(src, dst, n)
    seeds = np.array([3, 17, 17, 150], np.uint32)
    s = np.zeros(n)
    for q in seeds:
        s[q] += 1.0 / len(seeds)
    d = 0.85
    dang = (outdeg == 0).astype(np.float64)
    # x = (1-d) s + d (P x + (dang . x) s)  =>  (I - d P - d s dang^T) x = (1-d) s
    A = np.eye(n) - d * P.toarray() - d * np.outer(s, dang)
    want = np.linalg.solve(A, (1 - d) * s)
    got, it, res = oracle.ppr_f64(_csr(src, dst, n), seeds, 1e-13, 5000, d)
    assert res <= 1e-13
    np.testing.assert_allclose(got, want, atol=1e-11)
    assert abs(got.sum() - 1.0) < 1e-10


def test_omp_baseline_matches_serial():
    rng = np.random.RandomState(10)
    n, m = 5000, 80000
    src = rng.randint(0, n, m).astype(np.uint32)
    dst = rng.randint(0, n, m).astype(np.uint32)
    c = _csr(src, dst, n)
    secs, o = oracle.csr_build_omp(src, dst, n, threads=4)
    for k in oracle.CSR_KEYS:
        np.testing.assert_array_equal(o[k], c[k], err_msg=k)
    s, pr, it, err = oracle.pagerank_pull_omp(c, 20, 0.0, threads=4)
    ref, it2, err2, *_ = oracle.pagerank_pull_f32(c, 20, 0.0)
    assert it == it2 == 20
    np.testing.assert_array_equal(pr, ref)  # same per-vertex order, same FP32 ops
    assert abs(err - err2) <= 1e-12 * max(err2, 1e-300) + 1e-18
