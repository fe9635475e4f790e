"""Personalized PageRank on the GPU against the FP64 oracle (oracle.ppr_f64).

The reference has no PPR; these results are pinned by the FP64 oracle, whose
fixed point is itself pinned by a dense linear solve (test_oracle_pagerank).
Bars:
  * FP32 storage: L1 distance to the FP64 oracle run with the same stop rule
    <= FP32_L1 (rounding only), iteration count within one sweep;
  * batch == one query alone, bit for bit, for the same columns-per-wave;
  * graph replay == eager launches, bit for bit.
"""
import numpy as np
import pytest

import accelgraph_express_amd as ag
import oracle

pytestmark = pytest.mark.gpu
FP32_L1 = 2e-5


def _graph(ctx, scale=14, seed=31):
    src, dst = ag.gen_rmat(seed, scale, 16)
    n = 1 << scale
    return ag.Graph.build(ctx, src, dst, n), oracle.build_csr(src, dst, n), n


def _queries(n, k, seed, tol=1e-6):
    rng = np.random.RandomState(seed)
    return [(rng.randint(0, n, rng.randint(1, 17)).astype(np.uint32), tol, 1000) for _ in range(k)]


def test_single_fp32_vs_fp64(ctx):
    g, csr, n = _graph(ctx)
    for seeds in ([3], [3, 17, 17, 900], list(range(0, n, n // 16))):
        print(f"case: single seeds={len(seeds)}", flush=True)
        x, st = ag.ppr_single(g, seeds, tol=1e-6)
        ref, it, res = oracle.ppr_f64(csr, np.array(seeds, np.uint32), 1e-6)
        assert abs(int(st.iterations) - it) <= 1
        assert np.abs(x - ref).sum() <= FP32_L1
        assert abs(x.sum() - 1.0) < 1e-4


@pytest.mark.parametrize("cpw", [4, 8, 16, 32, 64])
def test_batch_fp32_vs_fp64(ctx, cpw):
    g, csr, n = _graph(ctx)
    qs = _queries(n, 20, 5)
    print(f"case: batch fp32 cpw={cpw} nq=20", flush=True)
    X, stats, ms = ag.ppr_batch(g, qs, precision="fp32", cols_per_wave=cpw)
    for i, (seeds, tol, mi) in enumerate(qs):
        ref, it, res = oracle.ppr_f64(csr, seeds, tol, mi)
        assert abs(int(stats[i].iterations) - it) <= 1, i
        assert stats[i].residual <= tol
        assert np.abs(X[i] - ref).sum() <= FP32_L1, i


def test_batch_alone_equals_mixed(ctx):
    """Every query of a 40-column batch equals its one-column run bit for bit,
    including the columns the draining batch packs into lower slots (rows
    without in-edges are never rewritten by a sweep, so a moved column must
    bring both contribution buffers along)."""
    g, csr, n = _graph(ctx)
    qs = _queries(n, 40, 9)
    X, st, _ = ag.ppr_batch(g, qs, precision="fp32", cols_per_wave=16)
    for i in range(40):
        Xa, sa, _ = ag.ppr_batch(g, [qs[i]], precision="fp32", cols_per_wave=16)
        assert sa[0].iterations == st[i].iterations, i
        np.testing.assert_array_equal(Xa[0], X[i])


def test_graph_replay_matches_eager(ctx):
    g, csr, n = _graph(ctx)
    qs = _queries(n, 12, 11)
    Xr, sr, _ = ag.ppr_batch(g, qs, precision="fp32", cols_per_wave=8, replay=True)
    Xe, se, _ = ag.ppr_batch(g, qs, precision="fp32", cols_per_wave=8, replay=False)
    np.testing.assert_array_equal(Xr, Xe)
    assert [s.iterations for s in sr] == [s.iterations for s in se]


def test_more_queries_than_slots(ctx):
    g, csr, n = _graph(ctx, scale=12)
    qs = _queries(n, 150, 13, tol=1e-5)  # > 64 slots: admission as slots free up
    X, st, _ = ag.ppr_batch(g, qs, precision="fp32", cols_per_wave=32)
    for i in (0, 64, 149):
        ref, it, res = oracle.ppr_f64(csr, qs[i][0], qs[i][1])
        assert np.abs(X[i] - ref).sum() <= FP32_L1


def test_autotune_picks_a_supported_width(ctx):
    g, csr, n = _graph(ctx)
    assert ag.ppr_autotune(g, "bf16") in (4, 8, 16, 32, 64)


@pytest.mark.parametrize("cpw", [8, 64])
def test_batch_bf16_finishes_in_fp32(ctx, cpw):
    """BF16 columns read the high halves of their FP32 contributions until the
    residual stalls near the BF16 floor, then the full words: the result obeys
    the residual bound against the converged FP64 fixed point,
    ||x - x*||_1 <= d/(1-d) * residual, at tolerances far below the floor."""
    g, csr, n = _graph(ctx)
    qs = _queries(n, 16, 21, tol=1e-6)
    print(f"case: batch bf16 cpw={cpw} nq=16 tol=1e-6", flush=True)
    X, stats, ms = ag.ppr_batch(g, qs, precision="bf16", cols_per_wave=cpw)
    for i, (seeds, tol, mi) in enumerate(qs):
        assert stats[i].residual <= tol
        star, _, _ = oracle.ppr_f64(csr, seeds, 1e-13, 5000)
        err = np.abs(X[i] - star).sum()
        assert err <= 0.85 / 0.15 * tol + FP32_L1, (i, err)


def test_batch_bf16_coarse_phase_is_unbiased(ctx):
    """A BF16 column stopped by its iteration limit while still coarse (reading
    truncated high halves as interval midpoints) stays within the BF16
    rounding band of the FP32 column after the same number of sweeps: total
    mass within 1e-3 of FP32's, L1 distance within 1e-2.  Truncation without
    the midpoint would lose about 2^-8 of the mass per sweep."""
    g, csr, n = _graph(ctx)
    qs = [(np.array([3, 99, 1234], np.uint32), 1e-9, 3)]
    Xb, sb, _ = ag.ppr_batch(g, qs, precision="bf16", cols_per_wave=8)
    Xf, sf, _ = ag.ppr_batch(g, qs, precision="fp32", cols_per_wave=8)
    assert sb[0].iterations == sf[0].iterations == 3
    assert abs(Xb[0].sum() - Xf[0].sum()) < 1e-3
    assert np.abs(Xb[0] - Xf[0]).sum() < 1e-2


def test_batch_fp16_is_refused(ctx):
    g, csr, n = _graph(ctx, scale=10)
    with pytest.raises(ag.AGError):
        ag.ppr_batch(g, _queries(n, 2, 1), precision="fp16", cols_per_wave=8)


def test_bf16_batch_alone_equals_mixed(ctx):
    g, csr, n = _graph(ctx)
    qs = _queries(n, 24, 23, tol=1e-5)
    X, st, _ = ag.ppr_batch(g, qs, precision="bf16", cols_per_wave=8)
    for i in (0, 11, 23):
        Xa, sa, _ = ag.ppr_batch(g, [qs[i]], precision="bf16", cols_per_wave=8)
        assert sa[0].iterations == st[i].iterations
        np.testing.assert_array_equal(Xa[0], X[i])


def test_bf16_wide_batch_relayout_alone_equals_mixed(ctx):
    """A 64-column BF16 batch stores 64-column split blocks (one line of high
    halves per row); once it drains into at most 32 slots the rows are re-laid
    into 32-column blocks in place.  Tolerances spread over two decades so the
    columns retire unevenly and the survivors are packed and re-laid while
    running: every checked result equals the query run alone, bit for bit."""
    g, csr, n = _graph(ctx)
    rng = np.random.RandomState(41)
    qs = [(rng.randint(0, n, rng.randint(1, 17)).astype(np.uint32),
           float(10 ** rng.uniform(-6, -4)), 1000) for _ in range(64)]
    X, st, _ = ag.ppr_batch(g, qs, precision="bf16", cols_per_wave=64)
    its = np.array([s.iterations for s in st])
    assert its.max() > its.min() + 3  # uneven: the batch drains through compaction
    for i in list(np.argsort(its)[-12:]) + [0, 31, 63]:
        Xa, sa, _ = ag.ppr_batch(g, [qs[i]], precision="bf16", cols_per_wave=64)
        assert sa[0].iterations == st[i].iterations, i
        np.testing.assert_array_equal(Xa[0], X[i])


def test_single_bf16_finishes_in_fp32(ctx):
    g, csr, n = _graph(ctx)
    seeds = np.array([5, 77, 4000], np.uint32)
    x, st = ag.ppr_single(g, seeds, tol=1e-6, precision="bf16")
    star, _, _ = oracle.ppr_f64(csr, seeds, 1e-13, 5000)
    assert st.residual <= 1e-6
    assert np.abs(x - star).sum() <= 0.85 / 0.15 * 1e-6 + FP32_L1


def test_bf16_coarse_certified_stop(ctx):
    """A BF16 column may stop on a coarse sweep when its residual plus the
    coarse-read noise bound (2^-8 of the mass on rows with out-edges) meets the
    tolerance.  A query whose seeds have no out-edges is x = s from the start:
    it stops after one sweep.  Its result is s up to the midpoint reads of zero
    contributions (2^-134 each, so sums below 1e-30), and the other columns of
    the batch are unchanged (alone == mixed)."""
    g, csr, n = _graph(ctx)
    od = np.asarray(csr["vertices.out_degree"])
    sinks = np.flatnonzero(od == 0).astype(np.uint32)
    assert sinks.size >= 8
    qs = _queries(n, 10, 21)
    qs.insert(3, (sinks[:1], 1e-6, 1000))
    qs.insert(7, (sinks[1:5], 1e-6, 1000))
    X, st, _ = ag.ppr_batch(g, qs, precision="bf16", cols_per_wave=64)
    for i in (3, 7):
        seeds = qs[i][0]
        s = np.zeros(n, np.float32)
        np.add.at(s, seeds.astype(np.int64), np.float32(1.0 / len(seeds)))
        assert st[i].iterations == 1, i
        np.testing.assert_allclose(X[i], s, rtol=0, atol=1e-30)
        assert X[i][seeds].min() > 0.0 and np.abs(X[i] - s).sum() < 1e-28
    for i in (0, 5, 11):
        Xa, sa, _ = ag.ppr_batch(g, [qs[i]], precision="bf16", cols_per_wave=64)
        assert sa[0].iterations == st[i].iterations
        np.testing.assert_array_equal(Xa[0], X[i])
        ref, _, _ = oracle.ppr_f64(csr, qs[i][0], 1e-13, 5000)
        assert np.abs(X[i].astype(np.float64) - ref).sum() <= 0.85 / 0.15 * 1e-6 + FP32_L1


@pytest.mark.parametrize("damp", [0.5, 0.95])
def test_damping_factor_single_and_batch(ctx, damp):
    """ag_ppr_single_opts and ag_ppr_batch honour a damping factor other than
    0.85: within the FP32 bound of the FP64 oracle run with the same d and stop
    rule, iteration counts within one sweep."""
    g, csr, n = _graph(ctx, scale=12)
    qs = _queries(n, 6, 13)
    X, st, _ = ag.ppr_batch(g, qs, precision="fp32", cols_per_wave=8, damp=damp)
    for i, (seeds, tol, mi) in enumerate(qs):
        ref, it, _ = oracle.ppr_f64(csr, seeds, tol, mi, damp=damp)
        x, s1 = ag.ppr_single(g, seeds, tol, damp=damp)
        assert abs(int(st[i].iterations) - it) <= 1 and abs(int(s1.iterations) - it) <= 1, i
        assert np.abs(X[i] - ref).sum() <= FP32_L1, i
        assert np.abs(x - ref).sum() <= FP32_L1, i
