"""Serving loop: continuous batching over resident graphs, HBM-budget
admission, injected faults and cancellation.  Each query's result must be the
same alone and in the mix (bit for bit at a fixed columns-per-wave) and match
the FP64 oracle to the FP32 bound."""
import numpy as np
import pytest

import accelgraph_express_amd as ag
from accelgraph_express_amd import _lib as L
import oracle

pytestmark = pytest.mark.gpu


def _drain(srv, tickets, n_of, limit=5000):
    out = {}
    for _ in range(limit):
        active = srv.step()
        for t in list(tickets):
            code, x, st = srv.poll(t, n_of[t])
            if code != L.AG_ERR_PENDING:
                out[t] = (code, x, st)
                tickets.remove(t)
        if not tickets and active == 0:
            break
    assert not tickets, "queries did not finish"
    return out


def _setup(ctx):
    graphs = []
    for scale, seed in ((12, 1), (13, 2)):
        src, dst = ag.gen_rmat(seed, scale, 16)
        graphs.append((ag.Graph.build(ctx, src, dst, 1 << scale),
                       oracle.build_csr(src, dst, 1 << scale), 1 << scale))
    return graphs


def test_alone_equals_in_mix(ctx):
    graphs = _setup(ctx)
    rng = np.random.RandomState(3)
    srv = ag.Server(ctx, precision="fp32", cols_per_wave=16)
    gids = [srv.add_graph(g) for g, _, _ in graphs]
    tickets, meta, n_of = [], {}, {}
    for i in range(30):
        k = rng.randint(0, 2)
        n = graphs[k][2]
        seeds = rng.randint(0, n, rng.randint(1, 17)).astype(np.uint32)
        tol = float(10 ** rng.uniform(-6, -4))
        t = srv.submit(gids[k], seeds, tol)
        tickets.append(t)
        meta[t] = (k, seeds, tol)
        n_of[t] = n
        if i % 5 == 4:  # arrivals while others run
            srv.step()
    res = _drain(srv, list(tickets), n_of)
    for t in tickets[:6]:
        code, x, st = res[t]
        assert code == L.AG_OK
        k, seeds, tol = meta[t]
        Xa, sa, _ = ag.ppr_batch(graphs[k][0], [(seeds, tol, 1000)], precision="fp32",
                                 cols_per_wave=16)
        np.testing.assert_array_equal(x, Xa[0])
        ref, it, _ = oracle.ppr_f64(graphs[k][1], seeds, tol)
        assert np.abs(x - ref).sum() <= 2e-5
        assert st.latency_ms > 0
    srv.close()


def test_bf16_serving_alone_equals_in_mix(ctx):
    graphs = _setup(ctx)
    g, csr, n = graphs[1]
    rng = np.random.RandomState(8)
    srv = ag.Server(ctx, precision="bf16", cols_per_wave=8)
    gid = srv.add_graph(g)
    qs = [(rng.randint(0, n, rng.randint(1, 17)).astype(np.uint32),
           float(10 ** rng.uniform(-6, -4))) for _ in range(20)]
    tickets = []
    for i, (s, tol) in enumerate(qs):
        tickets.append(srv.submit(gid, s, tol))
        if i % 3 == 2:
            srv.step()
    res = _drain(srv, list(tickets), {t: n for t in tickets})
    for i in (0, 9, 19):
        code, x, st = res[tickets[i]]
        assert code == L.AG_OK
        Xa, sa, _ = ag.ppr_batch(g, [(qs[i][0], qs[i][1], 1000)], precision="bf16",
                                 cols_per_wave=8)
        np.testing.assert_array_equal(x, Xa[0])
        star, _, _ = oracle.ppr_f64(csr, qs[i][0], 1e-13, 5000)
        assert np.abs(x - star).sum() <= 0.85 / 0.15 * qs[i][1] + 2e-5
    srv.close()


def test_hbm_budget_admission(ctx):
    graphs = _setup(ctx)
    n = graphs[0][2]
    # room for one 8-column batch only, as the library counts it (for either
    # graph: their chunk tables differ in size)
    s2, d2 = ag.gen_rmat(7, 12, 16)
    g2 = ag.Graph.build(ctx, s2, d2, n)
    one = max(ag.ppr_batch_bytes(graphs[0][0], 8, "fp32"), ag.ppr_batch_bytes(g2, 8, "fp32"))
    assert one > 8 * n * 12
    budget = one + 100
    assert budget < 2 * min(ag.ppr_batch_bytes(graphs[0][0], 8, "fp32"),
                            ag.ppr_batch_bytes(g2, 8, "fp32"))
    srv = ag.Server(ctx, hbm_budget_bytes=budget, max_columns=8, precision="fp32",
                    cols_per_wave=8)
    gid = srv.add_graph(graphs[0][0])
    plan_bytes0 = graphs[0][0].plan_info()["plan_bytes"]
    rng = np.random.RandomState(4)
    tickets = [srv.submit(gid, rng.randint(0, n, 3).astype(np.uint32), 1e-5) for _ in range(20)]
    srv.step()
    assert 0 < srv.reserved_bytes <= budget
    res = _drain(srv, list(tickets), {t: n for t in tickets})
    assert all(c == L.AG_OK for c, _, _ in res.values())
    held = srv.reserved_bytes
    assert 0 < held <= budget  # the drained batch stays resident for the next query
    # a query on another graph of the same size evicts the idle batch instead
    # of waiting forever (the budget holds one batch)
    gid2 = srv.add_graph(g2)
    t2 = srv.submit(gid2, [1, 2, 3], 1e-5)
    res2 = _drain(srv, [t2], {t2: n})
    assert res2[t2][0] == L.AG_OK
    assert 0 < srv.reserved_bytes <= budget
    # the evicted graph's column buffers were freed, not parked on the graph:
    # its plan holds only its gather layout again
    assert graphs[0][0].plan_info()["plan_bytes"] == plan_bytes0
    ref, _, _ = oracle.ppr_f64(oracle.build_csr(s2, d2, n), np.array([1, 2, 3], np.uint32), 1e-5)
    assert np.abs(res2[t2][1] - ref).sum() <= 2e-5
    # a graph whose column state never fits stays queued, and nothing breaks
    tiny = ag.Server(ctx, hbm_budget_bytes=1024, precision="fp32", cols_per_wave=8)
    g2 = tiny.add_graph(graphs[0][0])
    t = tiny.submit(g2, [1, 2], 1e-4)
    for _ in range(3):
        tiny.step()
    assert tiny.poll(t, n)[0] == L.AG_ERR_PENDING
    tiny.close()
    srv.close()


def test_injected_faults_and_cancel(ctx):
    graphs = _setup(ctx)
    g, csr, n = graphs[0]
    srv = ag.Server(ctx, precision="fp32", cols_per_wave=8, inject_fail_alloc=1,
                    inject_fail_query=3)
    gid = srv.add_graph(g)
    rng = np.random.RandomState(5)
    seeds = [rng.randint(0, n, 4).astype(np.uint32) for _ in range(6)]
    tickets = [srv.submit(gid, s, 1e-6) for s in seeds]
    srv.step()  # first allocation attempt fails by injection: nothing admitted yet
    srv.step()
    srv.cancel(tickets[4])  # cancel one mid-flight
    res = _drain(srv, [t for t in tickets if t != tickets[4]], {t: n for t in tickets})
    assert srv.poll(tickets[4], n)[0] == L.AG_ERR_CANCELLED
    assert res[tickets[2]][0] == L.AG_ERR_INJECTED  # ticket 3
    for i in (0, 1, 3, 5):
        code, x, st = res[tickets[i]]
        assert code == L.AG_OK
        ref, it, _ = oracle.ppr_f64(csr, seeds[i], 1e-6)
        assert np.abs(x - ref).sum() <= 2e-5
    srv.close()


def test_bad_requests(ctx):
    graphs = _setup(ctx)
    srv = ag.Server(ctx, precision="fp32", cols_per_wave=8)
    gid = srv.add_graph(graphs[0][0])
    with pytest.raises(ag.AGError):
        srv.submit(gid, [graphs[0][2] + 5], 1e-4)
    with pytest.raises(ag.AGError):
        srv.submit(gid + 7, [1], 1e-4)
    assert srv.poll(12345, 10)[0] == L.AG_ERR_NOT_FOUND
    srv.close()


def test_peek_is_zero_copy_and_matches_poll(ctx):
    """peek returns a view of the pinned buffer the GPU wrote (same values as a
    poll copy of the same query), valid until release; release frees it."""
    graphs = _setup(ctx)
    g, csr, n = graphs[0]
    srv = ag.Server(ctx, precision="fp32", cols_per_wave=8)
    gid = srv.add_graph(g)
    seeds = np.array([4, 40, 400], np.uint32)
    ta, tb = srv.submit(gid, seeds, 1e-6), srv.submit(gid, seeds, 1e-6)
    assert srv.peek(ta)[0] == L.AG_ERR_PENDING
    for _ in range(5000):
        srv.step()
        if srv.peek(ta)[0] != L.AG_ERR_PENDING and srv.peek(tb)[0] != L.AG_ERR_PENDING:
            break
    code, xa, sa = srv.peek(ta)
    assert code == L.AG_OK and xa.shape == (n,)
    code_b, xb, sb = srv.poll(tb, n)
    assert code_b == L.AG_OK
    np.testing.assert_array_equal(xa, xb)
    ref, _, _ = oracle.ppr_f64(csr, seeds, 1e-6)
    assert np.abs(xa - ref).sum() <= 2e-5
    srv.release(ta)
    with pytest.raises(ag.AGError):
        srv.release(ta)
    assert srv.peek(ta)[0] == L.AG_ERR_NOT_FOUND
    srv.close()


def test_bf16_layout_switches_while_running(ctx):
    """A 64-slot BF16 batch that starts narrow (32-column split blocks), widens
    past 48 running columns (re-laid to 64-column blocks while running) and
    drains again (re-laid back): every result equals the query alone."""
    graphs = _setup(ctx)
    g, csr, n = graphs[1]
    rng = np.random.RandomState(12)
    qs = [(rng.randint(0, n, rng.randint(1, 17)).astype(np.uint32),
           float(10 ** rng.uniform(-6, -5))) for _ in range(64)]
    srv = ag.Server(ctx, precision="bf16", cols_per_wave=64, max_columns=64)
    gid = srv.add_graph(g)
    tickets = [srv.submit(gid, s, tol) for s, tol in qs[:10]]
    srv.step()
    srv.step()
    tickets += [srv.submit(gid, s, tol) for s, tol in qs[10:]]
    res = _drain(srv, list(tickets), {t: n for t in tickets})
    for i in (0, 5, 9, 10, 40, 63):
        code, x, st = res[tickets[i]]
        assert code == L.AG_OK
        Xa, sa, _ = ag.ppr_batch(g, [(qs[i][0], qs[i][1], 1000)], precision="bf16",
                                 cols_per_wave=64)
        np.testing.assert_array_equal(x, Xa[0])
    srv.close()
