"""Full-size checks on the configurations' own graphs, through properties that
do not need the oracle to run to convergence at these sizes.

* C3, R-MAT scale 24 (N = 16,777,216, E = 268,435,456): every CSR array equals
  the oracle's build bit for bit.  STANDARD PageRank to L1 residual 1e-6 in
  FP32, BF16 and scaled FP16 (both finishing in FP32), plain and with
  extrapolated steps: total mass 1, and an
  FP64 error bound against the exact fixed point of 2e-5.  The map T(x) = (1-d)/N + d (A^T D^-1 x
  + dangling(x)/N) is a d-contraction in L1, so ||x - x*||_1 <= ||x - T(x)||_1
  / (1 - d), with T applied once in FP64 (numpy bincount over the edge list).
* C4, R-MAT scale 22: a 64-query BF16 batch at residual 1.5e-5; the same bound
  for the personalized map gives each query's L1 error against the FP64 fixed
  point, which must be <= 1e-4.
* R-MAT scale 22: data-driven pull stops at an epsilon-fixpoint; SPMV pull
  against an FP64 bincount; DBG relabelling labels against the oracle.
"""
import numpy as np
import pytest

import accelgraph_express_amd as ag
import oracle

pytestmark = pytest.mark.gpu
D = 0.85


@pytest.fixture(scope="module")
def s24():
    src, dst = ag.gen_rmat(2024, 24, 16)
    return src, dst, 1 << 24


@pytest.fixture(scope="module")
def s22():
    src, dst = ag.gen_rmat(2022, 22, 16)
    return src, dst, 1 << 22


def _standard_T(src, dst, n, x):
    od = np.bincount(src, minlength=n).astype(np.float64)
    x = x.astype(np.float64)
    contrib = np.divide(x, od, out=np.zeros(n), where=od > 0)
    dang = x[od == 0].sum()
    y = np.bincount(dst, weights=contrib[src], minlength=n)
    return (1 - D) / n + D * (y + dang / n)


def _ppr_T(src, dst, n, x, seeds, od):
    s = np.zeros(n)
    np.add.at(s, seeds.astype(np.int64), 1.0 / len(seeds))
    x = x.astype(np.float64)
    contrib = np.divide(x, od, out=np.zeros(n), where=od > 0)
    dang = x[od == 0].sum()
    y = np.bincount(dst, weights=contrib[src], minlength=n)
    return (1 - D) * s + D * (y + dang * s)


def test_c3_s24_csr_bit_exact(ctx, s24):
    src, dst, n = s24
    print(f"case: C3 CSR build N={n} E={src.size}", flush=True)
    g = ag.Graph.build(ctx, src, dst, n)
    got = g.download()
    ref = oracle.build_csr(src, dst, n)
    for k in oracle.CSR_KEYS:
        assert np.array_equal(got[k], ref[k]), k
    g.free()


def test_c3_s24_standard_precisions(ctx, s24):
    src, dst, n = s24
    print(f"case: C3 STANDARD PageRank to 1e-6 N={n}", flush=True)
    g = ag.Graph.build(ctx, src, dst, n)
    xs, sweeps = {}, {}
    # plain power iteration, then with extrapolated steps: both stop on a
    # plain sweep's residual, so the same bound holds
    for ex in (False, True):
        for prec, finish in (("fp32", False), ("bf16", True), ("fp16", True)):
            st = g.pagerank_pull(iterations=200, epsilon=1e-6, mode="standard", precision=prec,
                                 finish_fp32=finish, extrapolate=ex)
            assert st.iterations < 200
            x = st.pageRanks.astype(np.float64)
            r = np.abs(x - _standard_T(src, dst, n, x)).sum()
            bound = r / (1 - D)
            name = prec + ("_extrap" if ex else "")
            print(f"{name}: {st.iterations} sweeps, sum {x.sum():.9f}, ||x - T(x)||_1 {r:.3e}, "
                  f"error bound vs FP64 fixed point {bound:.3e}", flush=True)
            assert abs(x.sum() - 1.0) <= 1e-5
            assert bound <= 2e-5
            if prec != "fp32":
                assert st.fp32_switch_sweep > 0  # finished in FP32
            xs[name] = x
            sweeps[name] = st.sweeps
    for prec in ("bf16", "fp16"):
        print(f"L1 {prec} vs fp32: {np.abs(xs[prec] - xs['fp32']).sum():.3e}")
    assert sweeps["fp32_extrap"] < sweeps["fp32"]
    g.free()


def test_c4_s22_ppr_batch_error_bound(ctx, s22):
    src, dst, n = s22
    print(f"case: C4 64-query BF16 batch N={n}", flush=True)
    g = ag.Graph.build(ctx, src, dst, n)
    rng = np.random.RandomState(4)
    seeds = [rng.randint(0, n, rng.randint(1, 17)).astype(np.uint32) for _ in range(64)]
    X, st, _ = ag.ppr_batch(g, [(s, 1.5e-5, 1000) for s in seeds], precision="bf16",
                            cols_per_wave=64)
    od = np.bincount(src, minlength=n).astype(np.float64)
    worst = 0.0
    for i in (0, 7, 21, 42, 63):
        r = np.abs(X[i].astype(np.float64) - _ppr_T(src, dst, n, X[i], seeds[i], od)).sum()
        worst = max(worst, r / (1 - D))
        assert st[i].residual <= 1.5e-5
    print(f"largest L1 error bound vs the FP64 fixed point: {worst:.3e}")
    assert worst <= 1e-4


def test_s22_data_driven_stops_at_epsilon_fixpoint(ctx, s22):
    src, dst, n = s22
    g = ag.Graph.build(ctx, src, dst, n)
    eps = 1e-3
    st = g.pagerank_pull(iterations=500, epsilon=eps, mode="data_driven")
    assert st.iterations < 500 and st.activeVertices == 0
    x = st.pageRanks.astype(np.float64) * n
    od = np.bincount(src, minlength=n).astype(np.float64)
    contrib = np.divide(x, od, out=np.zeros(n), where=od > 0)
    tx = 0.15 + D * np.bincount(dst, weights=contrib[src], minlength=n)
    resid = np.abs(tx - x)
    # the GPU decides |delta| >= eps on FP32 sums; hub ranks reach ~1e3-1e4
    # (un-normalised), where an FP32 ulp is ~1e-4-1e-3: allow 16 ulps of the
    # vertex's value on top of eps
    slack = 16 * 2.0 ** -24 * np.maximum(x, tx)
    w = int(np.argmax(resid - slack))
    print(f"data-driven eps {eps}: {st.iterations} sweeps, max residual {resid.max():.3e}; "
          f"worst against its slack: vertex {w}, rank {x[w]:.4g}, residual {resid[w]:.3e}")
    assert np.all(resid < eps + slack)


def test_s22_spmv_vs_fp64(ctx, s22):
    src, dst, n = s22
    g = ag.Graph.build(ctx, src, dst, n)
    x = ag.spmv_input(np.bincount(src, minlength=n))
    y, r = g.spmv_pull(x, 2)
    y64 = np.bincount(dst, weights=x[src].astype(np.float64), minlength=n)
    nz = y64 > 0
    assert np.all(y[~nz] == 0)
    assert (np.abs(y[nz] - y64[nz]) / y64[nz]).max() <= 1e-5


def test_s22_relabel_dbg_vs_oracle(ctx, s22):
    src, dst, n = s22
    lab = ag.reorder_labels(ctx, src, dst, n, "dbg_in")
    np.testing.assert_array_equal(lab, oracle.reorder_labels(src, dst, n, 5))
