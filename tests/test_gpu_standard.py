"""GPU STANDARD PageRank (x0 = 1/N, dangling mass spread uniformly; the
formulation of the CPU-runnable workload's wording and of the precision form)
against the FP64 oracle (oracle.pagerank_standard_f64, pinned by scipy and by
mass conservation in test_oracle_pagerank.py).

Bars: FP32 storage after a fixed sweep count, every vertex within 1e-5
relative of the FP64 oracle after the same count (FP32 rounding only), mass 1
within 1e-6; closed forms on a cycle (stays uniform) and an edgeless graph
(stays uniform: all mass dangling); BF16 and scaled FP16 storage with the FP32
finish, run to residual 1e-6, within d/(1-d) * 1e-6 + 5e-6 in L1 of the
converged FP64 fixed point."""
import numpy as np
import pytest

import accelgraph_express_amd as ag
import oracle

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a.astype(np.float64) - b) / np.maximum(np.abs(b), 1e-300)


def test_uniform_65536_fixed_20_sweeps(ctx):
    src, dst = ag.gen_uniform(27491095, 65536, 1048576)
    n = 65536
    print(f"case: STANDARD uniform N={n} E={src.size} 20 sweeps", flush=True)
    g = ag.Graph.build(ctx, src, dst, n)
    st = g.pagerank_pull(iterations=20, epsilon=0.0, mode="standard")
    ref, it, res = oracle.pagerank_standard_f64(oracle.build_csr(src, dst, n), 20, 0.0)
    assert st.iterations == it == 20
    assert _rel(st.pageRanks, ref).max() <= 1e-5
    assert abs(st.pageRanks.astype(np.float64).sum() - 1.0) <= 1e-6
    # the last residual is near FP32 rounding noise (N values of ~1/N, each
    # rounded to 2^-24 relative): compare within that noise
    assert abs(st.error_total - res) <= 1e-5 * res + 2e-6


def test_skewed_bins_fixed_sweeps(ctx):
    saved = ag.get_tuning()
    ag.set_tuning(low_max=24, hub_min=512, hub_chunk=256)
    try:
        src, dst = ag.gen_rmat(77, 16, 16)
        n = 1 << 16
        g = ag.Graph.build(ctx, src, dst, n)
        g.prepare()
        info = g.plan_info()
        assert min(info["n_hub"], info["n_mid"], info["n_low"], info["n_zero"]) > 0
        st = g.pagerank_pull(iterations=30, epsilon=0.0, mode="standard")
    finally:
        ag.set_tuning(**saved)
    ref, it, res = oracle.pagerank_standard_f64(oracle.build_csr(src, dst, n), 30, 0.0)
    assert _rel(st.pageRanks, ref).max() <= 1e-5
    assert abs(st.pageRanks.astype(np.float64).sum() - 1.0) <= 1e-6


def test_closed_forms(ctx):
    n = 1000
    s = np.arange(n, dtype=np.uint32)
    cyc = ag.Graph.build(ctx, s, (s + 1) % n, n)
    x = cyc.pagerank_pull(iterations=10, epsilon=0.0, mode="standard").pageRanks
    np.testing.assert_allclose(x, 1.0 / n, rtol=1e-6)
    empty = ag.Graph.build(ctx, np.zeros(0, np.uint32), np.zeros(0, np.uint32), 7)
    x = empty.pagerank_pull(iterations=5, epsilon=0.0, mode="standard").pageRanks
    np.testing.assert_allclose(x, 1.0 / 7, rtol=1e-6)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_converged_against_fp64(ctx, prec):
    src, dst = ag.gen_rmat(5, 17, 16)
    n = 1 << 17
    g = ag.Graph.build(ctx, src, dst, n)
    st = g.pagerank_pull(iterations=300, epsilon=1e-6, mode="standard", precision=prec,
                         finish_fp32=prec != "fp32")
    star, _, _ = oracle.pagerank_standard_f64(oracle.build_csr(src, dst, n), 1000, 1e-13)
    l1 = np.abs(st.pageRanks.astype(np.float64) - star).sum()
    print(f"{prec}: {st.sweeps} sweeps (FP32 from {st.fp32_switch_sweep}), L1 vs FP64 {l1:.2e}")
    assert st.error_total <= 1e-6
    assert l1 <= 0.85 / 0.15 * 1e-6 + 5e-6


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_extrapolated_converged_against_fp64(ctx, prec):
    """Extrapolated steps (Kamvar et al., as in personalized mode): the run
    still stops on a plain sweep's residual, so the contraction bound of
    test_converged_against_fp64 holds; the FP32 run takes fewer sweeps than
    plain power iteration on this graph."""
    src, dst = ag.gen_rmat(5, 17, 16)
    n = 1 << 17
    print(f"case: STANDARD extrapolated R-MAT s17 {prec} to residual 1e-6", flush=True)
    g = ag.Graph.build(ctx, src, dst, n)
    kw = dict(iterations=300, epsilon=1e-6, mode="standard", precision=prec,
              finish_fp32=prec != "fp32")
    plain = g.pagerank_pull(**kw)
    st = g.pagerank_pull(extrapolate=True, **kw)
    star, _, _ = oracle.pagerank_standard_f64(oracle.build_csr(src, dst, n), 1000, 1e-13)
    l1 = np.abs(st.pageRanks.astype(np.float64) - star).sum()
    print(f"{prec}: {st.sweeps} sweeps extrapolated vs {plain.sweeps} plain, L1 vs FP64 {l1:.2e}")
    assert st.error_total <= 1e-6
    assert l1 <= 0.85 / 0.15 * 1e-6 + 5e-6
    assert abs(st.pageRanks.astype(np.float64).sum() - 1.0) <= 1e-5
    if prec == "fp32":
        assert st.sweeps < plain.sweeps


def test_extrapolate_refused_in_reference_mode(ctx):
    src, dst = ag.gen_uniform(3, 1024, 8192)
    g = ag.Graph.build(ctx, src, dst, 1024)
    with pytest.raises(ag.AGError):
        g.pagerank_pull(iterations=10, epsilon=0.0, mode="reference", extrapolate=True)
