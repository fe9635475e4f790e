"""First sweep of an admitted batched PPR column by push (k_ppr_admit_push).

x1 = T(s) is nonzero only on the seeds and their out-neighbours, so a batch
without extrapolated steps computes a fresh column's first iterate by walking
the seeds' out-edges (forward CSR) instead of sweeping all in-edges.

Bars, against the FP64 oracle run for one sweep (oracle.ppr_f64, max_iter 1):
  * the support of x1 exactly: rows the oracle leaves at zero are exactly 0;
  * FP32 rounding on the rest (the push stores exact FP32 contributions for
    BF16 columns too): L1 <= 2e-6, residual |x1 - s|_1 within 1e-6;
  * hub seeds, isolated seeds, 64 distinct seeds (the most a push takes);
  * a query of more seeds takes the pull first sweep (FP32 storage: the same
    bars; BF16: its first sweep reads midpoints, so alone == mixed only);
  * a column alone equals its run in a mixed batch bit for bit, after one
    sweep and at convergence;
  * a graph without a forward CSR (uploaded as CSR arrays) takes the pull
    first sweep and reaches the same fixed point within the stop bound.
"""
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


def _seed_sets(csr, n):
    od = csr["vertices.out_degree"]
    ind = csr["inverse_vertices.out_degree"]
    rng = np.random.RandomState(3)
    hub = int(np.argmax(od))
    isolated = np.nonzero((od == 0) & (ind == 0))[0]
    sets = [rng.randint(0, n, rng.randint(1, 17)).astype(np.uint32) for _ in range(6)]
    sets.append(np.array([hub], np.uint32))
    sets.append(np.array([hub, int(np.argsort(od)[-2]), 5, 5, 77], np.uint32))
    sets.append(rng.choice(n, 64, replace=False).astype(np.uint32))
    sets.append(rng.choice(n, 65, replace=False).astype(np.uint32))
    sets.append(rng.randint(0, n, 3000).astype(np.uint32))
    if isolated.size:
        sets.append(isolated[:2].astype(np.uint32))
    return sets


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_first_iterate_matches_oracle(ctx, prec):
    g, csr, n = _graph(ctx)
    sets = _seed_sets(csr, n)
    qs = [(s, 1e-12, 1) for s in sets]
    print(f"case: push first sweep {prec}, {len(qs)} queries", flush=True)
    X, st, _ = ag.ppr_batch(g, qs, precision=prec, cols_per_wave=64 if prec == "bf16" else 16)
    for i, s in enumerate(sets):
        x1, it, res = oracle.ppr_f64(csr, s, 1e-12, 1)
        assert it == 1 and st[i].iterations == 1 and st[i].status == L.AG_OK
        Xa, sa, _ = ag.ppr_batch(g, [qs[i]], precision=prec,
                                 cols_per_wave=64 if prec == "bf16" else 16)
        np.testing.assert_array_equal(Xa[0], X[i])
        if prec == "bf16" and np.unique(s).size > 64:
            continue  # pull first sweep with midpoint reads
        got = X[i].astype(np.float64)
        np.testing.assert_array_equal(got == 0.0, x1 == 0.0, err_msg=f"support of query {i}")
        l1 = np.abs(got - x1).sum()
        assert l1 <= 2e-6, (i, l1)
        assert abs(st[i].residual - res) <= 1e-6, (i, st[i].residual, res)


@pytest.mark.parametrize("prec,cpw", [("fp32", 8), ("bf16", 64)])
def test_converged_alone_equals_mixed(ctx, prec, cpw):
    g, csr, n = _graph(ctx)
    sets = _seed_sets(csr, n)
    qs = [(s, 1e-6, 1000) for s in sets]
    X, st, _ = ag.ppr_batch(g, qs, precision=prec, cols_per_wave=cpw)
    for i, s in enumerate(sets):
        Xa, sa, _ = ag.ppr_batch(g, [qs[i]], precision=prec, cols_per_wave=cpw)
        assert sa[0].iterations == st[i].iterations, i
        np.testing.assert_array_equal(Xa[0], X[i])
        xs, _, _ = oracle.ppr_f64(csr, s, 1e-13, 5000)
        assert np.abs(X[i].astype(np.float64) - xs).sum() <= D / (1 - D) * 1e-6 + FP32_L1, i


def test_graph_without_forward_csr_pulls_first(ctx):
    _, csr, n = _graph(ctx)
    gp = ag.Graph.from_csr(ctx, n, csr["vertices.out_degree"], csr["inverse_vertices.out_degree"],
                           csr["inverse_vertices.edges_idx"],
                           csr["inverse_sorted_edges_array.edges_array_dest"])
    sets = _seed_sets(csr, n)
    qs = [(s, 1e-6, 1000) for s in sets]
    X, st, _ = ag.ppr_batch(gp, qs, precision="fp32", cols_per_wave=16)
    for i, s in enumerate(sets):
        assert st[i].status == L.AG_OK and st[i].residual <= 1e-6
        xs, _, _ = oracle.ppr_f64(csr, s, 1e-13, 5000)
        assert np.abs(X[i].astype(np.float64) - xs).sum() <= D / (1 - D) * 1e-6 + FP32_L1, i
