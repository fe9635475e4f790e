"""Warm start of the pull (AG_PR_FLAG_WARM_START): the sweeps begin at the
caller's rank vector, passed in the output buffer, instead of the reference's
start.  No counterpart in the reference (pageRank.c:942-1108 starts at base).

Bar: where the reference's own sweeps are bit-exact against the oracle (rows
summed in edge order) the warm-started ranks equal the oracle's warm-started
ranks bit for bit; a resumed REFERENCE run (k sweeps, then m from the
returned ranks) equals the uninterrupted k + m sweeps bit for bit at N = 2^k;
a STANDARD warm start from the uniform vector equals the cold run bit for
bit; elsewhere the PageRank parity bar of 1e-5 relative per vertex holds
(REL_TOL)."""
import ctypes as C

import numpy as np
import pytest
import torch

import accelgraph_express_amd as ag
from accelgraph_express_amd import _lib as L
import oracle

pytestmark = pytest.mark.gpu
REL_TOL = 1e-5


def _rel(a, b):
    a, b = a.astype(np.float64), b.astype(np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def _start(n, seed):
    return (np.random.RandomState(seed).rand(n).astype(np.float32) * np.float32(2.0 / n))


def test_reference_and_dd_bit_exact_on_low_degree_rows(ctx):
    n = 4096
    print(f"case: warm start, low-degree rows N={n}", flush=True)
    src, dst = ag.gen_uniform(99, n, 20000)
    g = ag.Graph.build(ctx, src, dst, n)
    g.prepare()
    info = g.plan_info()
    assert info["n_mid"] == 0 and info["n_hub"] == 0
    csr = oracle.build_csr(src, dst, n)
    x0 = _start(n, 1)
    st = g.pagerank_pull(20, 0.0, init=x0)
    ref, it, err, psum, act, errs = oracle.pagerank_pull_f32(csr, 20, 0.0, init=x0)
    np.testing.assert_array_equal(st.pageRanks, ref)
    np.testing.assert_array_equal(st.active_per_iter, act)
    assert st.iterations == it
    st = g.pagerank_pull(50, 1e-4, mode="data_driven", init=x0)
    ref, it, err, psum, act, errs, work = oracle.pagerank_dd_pull_f32(csr, 50, 1e-4, init=x0)
    np.testing.assert_array_equal(st.pageRanks, ref)
    np.testing.assert_array_equal(st.active_per_iter, act)
    assert st.iterations == it
    g.free()


def test_reference_resume_equals_uninterrupted(ctx):
    n = 1 << 16
    print(f"case: warm-start resume, R-MAT N={n}", flush=True)
    src, dst = ag.gen_rmat(77, 16, 16)
    g = ag.Graph.build(ctx, src, dst, n)
    full = g.pagerank_pull(20, 0.0)
    part = g.pagerank_pull(8, 0.0)
    resumed = g.pagerank_pull(12, 0.0, init=part.pageRanks)
    np.testing.assert_array_equal(resumed.pageRanks, full.pageRanks)
    np.testing.assert_array_equal(resumed.active_per_iter, full.active_per_iter[8:])
    # against the oracle from the same start
    csr = oracle.build_csr(src, dst, n)
    x0 = _start(n, 2)
    st = g.pagerank_pull(20, 0.0, init=x0)
    ref, *_ = oracle.pagerank_pull_f32(csr, 20, 0.0, init=x0)
    assert _rel(st.pageRanks, ref).max() <= REL_TOL
    g.free()


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_standard_uniform_start_equals_cold(ctx, prec):
    n = 1 << 16
    print(f"case: STANDARD warm start from 1/N, {prec}, N={n}", flush=True)
    src, dst = ag.gen_rmat(78, 16, 16)
    g = ag.Graph.build(ctx, src, dst, n)
    kw = dict(iterations=100, epsilon=1e-6, mode="standard", precision=prec,
              finish_fp32=prec != "fp32")
    cold = g.pagerank_pull(**kw)
    warm = g.pagerank_pull(**kw, init=np.full(n, 1.0 / n, np.float32))
    np.testing.assert_array_equal(warm.pageRanks, cold.pageRanks)
    assert warm.sweeps == cold.sweeps
    np.testing.assert_array_equal(warm.error_per_iter, cold.error_per_iter)
    g.free()


def test_standard_from_fixed_point_and_random_start(ctx):
    n = 50000
    print(f"case: STANDARD warm starts, N={n}", flush=True)
    src, dst = ag.gen_rmat(79, 16, 8)
    src, dst = src % n, dst % n
    g = ag.Graph.build(ctx, src, dst, n)
    csr = oracle.build_csr(src, dst, n)
    cold = g.pagerank_pull(200, 1e-6, mode="standard")
    # from the converged ranks the first sweep's residual is already small
    again = g.pagerank_pull(200, 1e-6, mode="standard", init=cold.pageRanks)
    print(f"cold {cold.sweeps} sweeps, warm from its ranks {again.sweeps}")
    assert again.sweeps <= 3, again.sweeps
    assert np.abs(again.pageRanks.astype(np.float64) - cold.pageRanks).sum() <= 1e-5
    # 10 sweeps from a random distribution against the FP64 oracle from it
    x0 = _start(n, 3)
    x0 /= x0.sum(dtype=np.float64)
    st = g.pagerank_pull(10, 0.0, mode="standard", init=x0)
    ref, _, _ = oracle.pagerank_standard_f64(csr, 10, init=x0.astype(np.float64))
    assert np.abs(st.pageRanks - ref).sum() <= 1e-5
    assert _rel(st.pageRanks, ref).max() <= 1e-4
    g.free()


def test_device_and_allocation_free_entries(ctx):
    n = 1 << 14
    print(f"case: warm start through the device entry points, N={n}", flush=True)
    src, dst = ag.gen_rmat(80, 14, 16)
    g = ag.Graph.build(ctx, src, dst, n)
    x0 = _start(n, 4)
    host = g.pagerank_pull(20, 0.0, init=x0)
    for noalloc in (False, True):
        d = torch.from_numpy(x0).cuda()
        if noalloc:
            g.prepare()
            g.pagerank_pull_device_noalloc(d.data_ptr(), 20, 0.0, warm_start=True)
        else:
            g.pagerank_pull_device(d.data_ptr(), 20, 0.0, warm_start=True)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d.cpu().numpy(), host.pageRanks)
    # twice in a row on the device: the second run starts where the first ended
    d = torch.from_numpy(x0).cuda()
    g.pagerank_pull_device(d.data_ptr(), 10, 0.0, warm_start=True)
    g.pagerank_pull_device(d.data_ptr(), 10, 0.0, warm_start=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), host.pageRanks)
    g.free()


def test_refusals(ctx):
    src, dst = ag.gen_uniform(5, 1000, 5000)
    g = ag.Graph.build(ctx, src, dst, 1000)
    lib = L.load()
    o = L.PrOptions()
    lib.ag_pr_options_default(C.byref(o))
    o.flags = L.AG_PR_FLAG_WARM_START
    r = L.PrResult()
    # the host entry needs the start in `ranks`
    assert lib.ag_pagerank_pull(ctx.h, g.h, C.byref(o), None, None, None, C.byref(r)) == \
        L.AG_ERR_INVALID
    assert b"warm start" in lib.ag_last_error()
    g.free()
