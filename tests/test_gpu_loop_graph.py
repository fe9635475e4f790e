"""The PageRank sweep loop replayed from hipGraphs equals the eager loop.

ag_pagerank_pull_device captures init + sweeps + normalise into hipGraphs
cached on the graph (one graph for a fixed sweep count, one per group of 16
sweeps plus the normalise step when epsilon > 0) and replays them on later
calls with the same arguments.  AG_PR_FLAG_TIME_SWEEPS runs the same kernels
eagerly.  Each replay is checked against the eager result bit for bit, with
eager work and synchronisations between the replays."""
import numpy as np
import pytest
import torch

import accelgraph_express_amd as ag

pytestmark = pytest.mark.gpu


def _graph(ctx):
    src, dst = ag.gen_rmat(7, 14, 16)
    g = ag.Graph.build(ctx, src, dst, num_vertices=1 << 14)
    g.prepare()
    return g


@pytest.mark.parametrize("kw", [
    dict(iterations=20, epsilon=0.0, mode="reference"),
    dict(iterations=40, epsilon=1e-7, mode="standard"),
    dict(iterations=30, epsilon=1e-4, mode="reference"),
    dict(iterations=30, epsilon=1e-6, mode="standard", precision="bf16", finish_fp32=True),
    dict(iterations=20, epsilon=1e-4, mode="data_driven"),
])
def test_replay_equals_eager(ctx, kw):
    print(f"case: R-MAT s14 {kw}", flush=True)
    g = _graph(ctx)
    n = g.num_vertices
    eager = torch.zeros(n, dtype=torch.float32, device="cuda")
    out = torch.zeros(n, dtype=torch.float32, device="cuda")
    re = g.pagerank_pull_device(eager.data_ptr(), time_sweeps=True, **kw)
    torch.cuda.synchronize()
    want = eager.cpu().numpy()
    for rep in range(4):
        out.fill_(-1.0)
        torch.cuda.synchronize()
        r = g.pagerank_pull_device(out.data_ptr(), **kw)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"replay {rep}"
        assert (r.iterations, r.sweeps, r.active_vertices) == \
            (re.iterations, re.sweeps, re.active_vertices)
        assert r.error_total == re.error_total and r.pr_sum == re.pr_sum
        assert r.loop_ms > 0.0 and r.total_ms >= r.loop_ms * 0.999
        # eager work between the replays, on the library's stream and torch's
        g.pagerank_pull_device(eager.data_ptr(), time_sweeps=True, **kw)
        torch.cuda.synchronize()
        assert np.array_equal(eager.cpu().numpy().view(np.uint32), want.view(np.uint32))
        (out * 2.0).sum().item()


def test_new_output_buffer_recaptures(ctx):
    g = _graph(ctx)
    n = g.num_vertices
    bufs = [torch.zeros(n, dtype=torch.float32, device="cuda") for _ in range(3)]
    for b in bufs:
        g.pagerank_pull_device(b.data_ptr(), iterations=20, epsilon=0.0)
    torch.cuda.synchronize()
    a = [b.cpu().numpy() for b in bufs]
    assert np.array_equal(a[0], a[1]) and np.array_equal(a[0], a[2])
    # host-ranks entry point (the drop-in's path) replays too
    st = g.pagerank_pull(iterations=20, epsilon=0.0)
    st2 = g.pagerank_pull(iterations=20, epsilon=0.0)
    assert np.array_equal(st.pageRanks, a[0]) and np.array_equal(st2.pageRanks, a[0])
