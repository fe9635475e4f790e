"""ag_pagerank_pull_device_noalloc: the allocation-free form of
ag_pagerank_pull_device (per-call buffers in the library's static 16 MiB
device scratch).  Called through the C ABI and checked bit for bit against
the allocating entry point, on two calls in a row, the second on a smaller
graph; sizes beyond the scratch and a graph without a gather layout are
refused."""
import ctypes as C

import numpy as np
import pytest
import torch

import accelgraph_express_amd as ag
from accelgraph_express_amd import _lib as L

pytestmark = pytest.mark.gpu

CASES = [
    dict(iterations=20, epsilon=0.0, mode="reference"),
    dict(iterations=40, epsilon=1e-7, mode="standard"),
    dict(iterations=30, epsilon=1e-6, mode="standard", precision="bf16", finish_fp32=True),
]


def _same(g, kw):
    n = g.num_vertices
    want = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
    got = torch.full((n,), -2.0, dtype=torch.float32, device="cuda")
    rw = g.pagerank_pull_device(want.data_ptr(), **kw)
    torch.cuda.synchronize()
    rg = g.pagerank_pull_device_noalloc(got.data_ptr(), **kw)
    torch.cuda.synchronize()
    a, b = want.cpu().numpy(), got.cpu().numpy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert (rw.iterations, rw.sweeps, rw.active_vertices) == (rg.iterations, rg.sweeps,
                                                             rg.active_vertices)
    assert rw.error_total == rg.error_total and rw.pr_sum == rg.pr_sum


@pytest.mark.parametrize("kw", CASES)
def test_noalloc_equals_allocating_entry(ctx, kw):
    big_src, big_dst = ag.gen_rmat(11, 17, 16)           # N = 131,072, E = 2,097,152
    small_src, small_dst = ag.gen_uniform(27491095, 4096, 65536)
    for name, (s, d, n) in (("rmat s17", (big_src, big_dst, 1 << 17)),
                            ("uniform N=4096", (small_src, small_dst, 4096))):
        print(f"case: {name} {kw}", flush=True)
        g = ag.Graph.build(ctx, s, d, num_vertices=n)
        g.prepare()
        _same(g, kw)
        g.free()


def test_noalloc_refusals(ctx):
    lib = L.load()
    src, dst = ag.gen_uniform(5, 1024, 8192)
    g = ag.Graph.build(ctx, src, dst, num_vertices=1024)
    out = torch.zeros(1024, dtype=torch.float32, device="cuda")
    o = L.PrOptions()
    lib.ag_pr_options_default(C.byref(o))
    o.epsilon = 0.0
    r = L.PrResult()
    # no gather layout yet: refused (building it would allocate)
    rc = lib.ag_pagerank_pull_device_noalloc(ctx.h, g.h, C.byref(o), C.c_void_p(out.data_ptr()),
                                             C.byref(r))
    assert rc == L.AG_ERR_INVALID and "ag_pagerank_prepare" in lib.ag_last_error().decode()
    g.prepare()
    o.flags = L.AG_PR_FLAG_TIME_SWEEPS
    rc = lib.ag_pagerank_pull_device_noalloc(ctx.h, g.h, C.byref(o), C.c_void_p(out.data_ptr()),
                                             C.byref(r))
    assert rc == L.AG_ERR_INVALID
    o.flags = 0
    assert lib.ag_pagerank_pull_device_noalloc(ctx.h, g.h, C.byref(o),
                                               C.c_void_p(out.data_ptr()), C.byref(r)) == L.AG_OK
    # more vertices than the static scratch holds: refused before any launch
    n_big = 1 << 21
    gb = ag.Graph.build(ctx, np.arange(64, dtype=np.uint32), np.arange(64, dtype=np.uint32)[::-1].copy(),
                        num_vertices=n_big)
    gb.prepare()
    outb = torch.zeros(n_big, dtype=torch.float32, device="cuda")
    rc = lib.ag_pagerank_pull_device_noalloc(ctx.h, gb.h, C.byref(o), C.c_void_p(outb.data_ptr()),
                                             C.byref(r))
    assert rc == L.AG_ERR_INVALID and "scratch" in lib.ag_last_error().decode()
    # a refused call leaves the entry point usable
    _same(g, dict(iterations=20, epsilon=0.0, mode="reference"))
