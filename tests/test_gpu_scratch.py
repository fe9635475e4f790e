"""Scratch-arena growth inside one context.

A context whose scratch slots were sized by a small graph grows them during a
larger graph's CSR and gather-layout builds, while that build's earlier
kernels are still queued: a grown slot's old buffer is retired (freed once the
stream is idle), not waited for.  The larger graph's CSR arrays and its ranks
must equal those a fresh context computes, bit for bit."""
import numpy as np
import pytest

import accelgraph_express_amd as ag

pytestmark = pytest.mark.gpu


def _run(ctx, src, dst, n):
    g = ag.Graph.build(ctx, src, dst, num_vertices=n)
    ranks = g.pagerank_pull(iterations=20, epsilon=0.0).pageRanks.copy()
    arrays = g.download()
    g.free()
    return arrays, ranks


def test_grown_scratch_matches_fresh_context():
    s_src, s_dst = ag.gen_uniform(7, 4096, 65536)
    src, dst = ag.gen_rmat(2020, 20, 16)
    n = 1 << 20
    print(f"case: small N=4096 E=65536, then R-MAT s20 N={n} E={len(src)}", flush=True)
    grown = ag.Context(0)
    try:
        small = ag.Graph.build(grown, s_src, s_dst, num_vertices=4096)
        small.pagerank_pull(iterations=5, epsilon=0.0)
        small.free()
        arrays_g, ranks_g = _run(grown, src, dst, n)
    finally:
        grown.close()
    fresh = ag.Context(0)
    try:
        arrays_f, ranks_f = _run(fresh, src, dst, n)
    finally:
        fresh.close()
    for k in arrays_f:
        np.testing.assert_array_equal(arrays_g[k], arrays_f[k], err_msg=k)
    np.testing.assert_array_equal(ranks_g, ranks_f)
