"""Extrapolated personalized PageRank (ag_ppr_options.extrapolate) on the GPU.

Once a column's residual falls by a steady ratio rho, one sweep stores
x + rho / (1 - rho) (T(x) - x) instead of T(x) (Aitken-type extrapolation,
Kamvar et al., WWW 2003).  That changes the path to the fixed point, not the
fixed point: a column still stops only on a plain sweep's residual R <= tol, so
||x - x*||_1 <= d / (1 - d) * tol against the exact fixed point.

Bars: that bound plus FP32 rounding (2e-5) against the FP64 oracle's converged
fixed point, for the batch (FP32 and BF16 storage), the single-vector path and
the serving loop; fewer sweeps than power iteration on average; a column alone
equals its run in a mixed batch (and in the serving mix) bit for bit; graph
replay equals eager launches."""
import ctypes as C

import numpy as np
import pytest

import accelgraph_express_amd as ag
from accelgraph_express_amd import _lib as L
import oracle

pytestmark = pytest.mark.gpu
D = 0.85
FP32_L1 = 2e-5


def _graph(ctx, scale=14, seed=31):
    src, dst = ag.gen_rmat(seed, scale, 16)
    n = 1 << scale
    return ag.Graph.build(ctx, src, dst, n), oracle.build_csr(src, dst, n), n


def _queries(n, k, seed, tol=1e-6):
    rng = np.random.RandomState(seed)
    return [(rng.randint(0, n, rng.randint(1, 17)).astype(np.uint32), tol, 1000) for _ in range(k)]


def _fixed_point(csr, seeds):
    x, _, _ = oracle.ppr_f64(csr, seeds, 1e-13, 5000)
    return x


@pytest.mark.parametrize("prec,cpw", [("fp32", 16), ("bf16", 64)])
def test_batch_within_bound_and_fewer_sweeps(ctx, prec, cpw):
    g, csr, n = _graph(ctx)
    qs = _queries(n, 20, 5)
    print(f"case: extrapolated batch {prec} cpw={cpw} nq=20", flush=True)
    X, st, _ = ag.ppr_batch(g, qs, precision=prec, cols_per_wave=cpw, extrapolate=True)
    _, sp, _ = ag.ppr_batch(g, qs, precision=prec, cols_per_wave=cpw, return_x=False)
    worst = 0.0
    for i, (seeds, tol, _) in enumerate(qs):
        assert st[i].status == L.AG_OK
        assert st[i].residual <= tol, i
        l1 = np.abs(X[i].astype(np.float64) - _fixed_point(csr, seeds)).sum()
        worst = max(worst, l1)
        assert l1 <= D / (1 - D) * tol + FP32_L1, i
    ext = np.mean([s.iterations for s in st])
    plain = np.mean([s.iterations for s in sp])
    print(f"{prec}: mean sweeps {ext:.2f} extrapolated vs {plain:.2f} plain; "
          f"worst L1 to the fixed point {worst:.2e}")
    assert ext < plain


def test_single_within_bound_and_fewer_sweeps(ctx):
    g, csr, n = _graph(ctx)
    for seeds in ([3], [3, 17, 17, 900], list(range(0, n, n // 16))):
        print(f"case: extrapolated single seeds={len(seeds)}", flush=True)
        x, st = ag.ppr_single(g, seeds, tol=1e-6, extrapolate=True)
        _, sp = ag.ppr_single(g, seeds, tol=1e-6, return_x=False)
        assert st.residual <= 1e-6
        l1 = np.abs(x.astype(np.float64) - _fixed_point(csr, np.array(seeds, np.uint32))).sum()
        print(f"{st.iterations} sweeps extrapolated vs {sp.iterations} plain, L1 {l1:.2e}")
        assert l1 <= D / (1 - D) * 1e-6 + FP32_L1
        assert st.iterations <= sp.iterations
        assert abs(x.sum() - 1.0) < 1e-4


def test_extrapolation_is_refused_outside_personalized_mode(ctx):
    g, _, _ = _graph(ctx, scale=10)
    lib = L.load()
    o = L.PrOptions()
    lib.ag_pr_options_default(o)
    o.flags = 0x4  # AG_PR_FLAG_EXTRAPOLATE
    out = np.zeros(g.num_vertices, np.float32)
    rc = lib.ag_pagerank_pull(g.ctx.h, g.h, o, out.ctypes.data_as(C.POINTER(C.c_float)), None,
                              None, None)
    assert rc == L.AG_ERR_INVALID


@pytest.mark.parametrize("prec,cpw", [("fp32", 8), ("bf16", 64)])
def test_alone_equals_mixed(ctx, prec, cpw):
    g, csr, n = _graph(ctx)
    qs = _queries(n, 24, 9, tol=1e-5)
    X, st, _ = ag.ppr_batch(g, qs, precision=prec, cols_per_wave=cpw, extrapolate=True)
    for i in range(0, 24, 3):
        Xa, sa, _ = ag.ppr_batch(g, [qs[i]], precision=prec, cols_per_wave=cpw, extrapolate=True)
        assert sa[0].iterations == st[i].iterations, i
        np.testing.assert_array_equal(Xa[0], X[i])


def test_graph_replay_matches_eager(ctx):
    g, csr, n = _graph(ctx)
    qs = _queries(n, 12, 11)
    Xr, sr, _ = ag.ppr_batch(g, qs, precision="fp32", cols_per_wave=8, replay=True, extrapolate=True)
    Xe, se, _ = ag.ppr_batch(g, qs, precision="fp32", cols_per_wave=8, replay=False, extrapolate=True)
    np.testing.assert_array_equal(Xr, Xe)
    assert [s.iterations for s in sr] == [s.iterations for s in se]


def test_serving_alone_equals_in_mix(ctx):
    g, csr, n = _graph(ctx, scale=13, seed=2)
    rng = np.random.RandomState(7)
    srv = ag.Server(ctx, precision="bf16", cols_per_wave=64, extrapolate=True)
    gid = srv.add_graph(g)
    tickets, meta = [], {}
    for i in range(24):
        seeds = rng.randint(0, n, rng.randint(1, 17)).astype(np.uint32)
        tol = float(10 ** rng.uniform(-6, -4))
        t = srv.submit(gid, seeds, tol)
        tickets.append(t)
        meta[t] = (seeds, tol)
        if i % 6 == 5:
            srv.step()
    pending, res = list(tickets), {}
    for _ in range(5000):
        active = srv.step()
        for t in list(pending):
            code, x, st = srv.poll(t, n)
            if code != L.AG_ERR_PENDING:
                res[t] = (code, x, st)
                pending.remove(t)
        if not pending and active == 0:
            break
    assert not pending, "queries did not finish"
    for t in tickets[:8]:
        code, x, st = res[t]
        assert code == L.AG_OK
        seeds, tol = meta[t]
        Xa, sa, _ = ag.ppr_batch(g, [(seeds, tol, 1000)], precision="bf16", cols_per_wave=64,
                                 extrapolate=True)
        np.testing.assert_array_equal(x, Xa[0])
        l1 = np.abs(x.astype(np.float64) - _fixed_point(csr, seeds)).sum()
        assert l1 <= D / (1 - D) * tol + FP32_L1
    srv.close()


def test_max_iter_and_null_options(ctx):
    """A column that reaches max_iter stops there whatever the extrapolation
    state (its last sweep may have stepped); its result stays finite with mass
    1, alone == mixed.  ag_ppr_single_opts refuses a null options pointer."""
    g, csr, n = _graph(ctx)
    qs = [(s, 1e-12, 6) for s, _, _ in _queries(n, 5, 17)]
    X, st, _ = ag.ppr_batch(g, qs, precision="fp32", cols_per_wave=8, extrapolate=True)
    for i in range(5):
        assert st[i].iterations == 6 and st[i].status == L.AG_OK
        assert np.isfinite(X[i]).all() and abs(X[i].astype(np.float64).sum() - 1.0) < 1e-4
        Xa, sa, _ = ag.ppr_batch(g, [qs[i]], precision="fp32", cols_per_wave=8, extrapolate=True)
        np.testing.assert_array_equal(Xa[0], X[i])
    x, s1 = ag.ppr_single(g, qs[0][0], 1e-12, max_iter=6, extrapolate=True)
    assert s1.iterations == 6 and np.isfinite(x).all()
    q, keep = ag.graph._query(qs[0][0], 1e-4, 100)
    st1 = L.PprStat()
    rc = L.load().ag_ppr_single_opts(g.ctx.h, g.h, C.byref(q), None, None, C.byref(st1))
    assert rc == L.AG_ERR_INVALID
