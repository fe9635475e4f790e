"""Pins the oracle's data-driven pull PageRank (pageRankDataDrivenPullGraphCSR,
pageRank.c:2810-2968), CPU only.

* Two-vertex graph, hand-computed: below epsilon a vertex keeps its old rank
  (pull would write the new one), and the break sweep is not counted.
* Directed path 0->1->...->L-1, closed form: updates travel one hop per sweep
  along forward edges, so the worklist after sweep t is {t+2..L-1}, and the
  loop stops at the first sweep whose change 0.15*0.85^(t+1) is below
  epsilon; vertex k ends at 0.15 * sum_{i<=min(t*,k)} 0.85^i.  Marking
  in-neighbours instead of out-neighbours, or updating below epsilon, breaks it.
* epsilon = 0 reduces to the plain pull oracle (itself pinned by closed forms
  -- cycle, two-vertex and star graphs -- and by an independent scipy power
  iteration in FP64, in test_oracle_pagerank.py): every evaluated vertex
  updates, and a vertex off the worklist has a constant gathered sum, so ranks
  and error totals agree bit for bit.  A Gauss-Seidel slip (in-place
  contributions) breaks this.
* Stopping property on random graphs (FP64 recomputation): when the loop stops
  before its cap, every vertex's pull residual against the final ranks is below
  epsilon.
"""
import numpy as np
import pytest

import oracle

D = 0.85
B = 0.15


def _csr(src, dst, n):
    return oracle.build_csr(np.asarray(src, np.uint32), np.asarray(dst, np.uint32), n)


def test_two_vertex_hand_computed():
    c = _csr([0], [1], 2)
    # sweep 0: v1 changes by d*base = 0.1275 (FP32)
    pr, it, err, psum, act, errs, work = oracle.pagerank_dd_pull_f32(c, 20, 0.2)
    assert it == 0 and list(act) == [0] and list(work) == [2]
    base32 = np.float32(1.0) - np.float32(D)  # base_pr = 1.0f - 0.85f
    assert np.all(pr == base32 / np.float32(2))
    pull = oracle.pagerank_pull_f32(c, 20, 0.2)[0]
    assert pull[1] > pr[1]  # pull writes the sub-epsilon change, data-driven does not
    pr, it, err, psum, act, errs, work = oracle.pagerank_dd_pull_f32(c, 20, 0.1)
    # v1 updates in sweep 0 and marks nothing (no out-edges): sweep 1 is empty
    assert it == 1 and list(act) == [1, 0] and list(work) == [2, 0]
    assert err == 0.0 and errs[0] == pytest.approx(0.1275 / 2, rel=1e-6)
    assert pr[1] == pytest.approx((B + D * B) / 2, rel=1e-6)


@pytest.mark.parametrize("eps", [1e-2, 1e-3, 3e-5])
def test_path_closed_form(eps):
    L = 200
    s = np.arange(L - 1)
    c = _csr(s, s + 1, L)
    pr, it, err, psum, act, errs, work = oracle.pagerank_dd_pull_f32(c, 500, eps)
    # first sweep whose change 0.15*0.85^(t+1) falls below epsilon
    tstar = next(t for t in range(500) if B * D ** (t + 1) < eps)
    assert it == tstar
    want = np.array([B * sum(D ** i for i in range(min(tstar, k) + 1)) for k in range(L)])
    np.testing.assert_allclose(pr.astype(np.float64) * L, want, rtol=3e-6)
    # worklist and update counts per sweep
    assert work[0] == L and list(work[1:]) == [L - 1 - t for t in range(1, tstar + 1)]
    assert list(act[:tstar]) == [L - 1 - t for t in range(tstar)] and act[tstar] == 0
    # |delta| is an FP32 difference of values near 1: absolute error ~2^-24 each
    for t in range(1, tstar + 1):
        assert errs[t] == pytest.approx((L - 1 - t) * B * D ** (t + 1) / L, abs=2e-7)


@pytest.mark.parametrize("seed,n,m", [(1, 500, 4000), (2, 3000, 9000), (3, 64, 2000)])
def test_epsilon_zero_equals_pull(seed, n, m):
    rng = np.random.RandomState(seed)
    src = rng.randint(0, n, m).astype(np.uint32)
    dst = rng.randint(0, n, m).astype(np.uint32)
    c = _csr(src, dst, n)
    pr, it, err, psum, act, errs, work = oracle.pagerank_dd_pull_f32(c, 15, 0.0)
    pp, pit, perr, ppsum, pact, perrs = oracle.pagerank_pull_f32(c, 15, 0.0)
    assert it == pit == 15
    np.testing.assert_array_equal(pr, pp)
    np.testing.assert_array_equal(errs, perrs)
    # every evaluated vertex is updated at epsilon 0
    np.testing.assert_array_equal(act, work)
    indeg = c["inverse_vertices.out_degree"]
    assert work[1] == np.count_nonzero(indeg)


@pytest.mark.parametrize("seed,eps", [(4, 1e-3), (5, 1e-4), (6, 3e-3)])
def test_stop_means_epsilon_fixpoint(seed, eps):
    rng = np.random.RandomState(seed)
    n, m = 2000, 16000
    src = rng.randint(0, n, m).astype(np.uint32)
    dst = rng.randint(0, n, m).astype(np.uint32)
    c = _csr(src, dst, n)
    pr, it, err, psum, act, errs, work = oracle.pagerank_dd_pull_f32(c, 1000, eps)
    assert it < 1000 and act[-1] == 0
    x = pr.astype(np.float64) * n  # un-normalised ranks
    od = c["vertices.out_degree"].astype(np.float64)
    contrib = np.where(od > 0, x / np.maximum(od, 1), 0.0)
    nxt = np.zeros(n)
    np.add.at(nxt, dst.astype(np.int64), contrib[src.astype(np.int64)])
    resid = np.abs(B + D * nxt - x)
    assert resid.max() < eps * (1 + 1e-3) + 1e-6
    # and the loop did not stop early: the previous sweep still updated vertices
    assert len(act) == 1 or act[-2] > 0
