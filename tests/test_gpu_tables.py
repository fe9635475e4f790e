"""The partitioned-table gather layout (AG_LAYOUT_TABLES, the default).

Source labels are cut into LDS-sized tables; each row's in-edges from one
table form a sub-row summed in edge order (one lane, or a lane group for long
sub-rows, or wave chunks for hub rows), and the combine kernel adds a row's
sub-row sums in table order.  Parity bar as for the row layout: after 20
sweeps every vertex within 1e-5 relative of the FP64 oracle, and within 1e-5
of the FP32 restatement except where the restatement's sequential sum is the
farther of the two from FP64.  Rows whose in-edges all come from one table
and fit one lane sum exactly as the reference does, so such graphs are
bit-exact.  The cases force every path: several LDS tables plus the cold
table (max_tables), lane / 16 / 32 / 64-lane sub-rows (lane_max), hub chunks
(hub_min, hub_chunk), and the precision, data-driven, STANDARD and SPMV
modes."""
import numpy as np
import pytest

import accelgraph_express_amd as ag
import oracle
from conftest import tuned

pytestmark = pytest.mark.gpu
REL_TOL = 1e-5


def _rel(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def _check(gpu, ref32, ref64):
    r32 = _rel(gpu, ref32)
    bad = r32 > REL_TOL
    if bad.any():
        assert np.all(_rel(gpu[bad], ref64[bad]) <= _rel(ref32[bad], ref64[bad]))
    assert _rel(gpu, ref64).max() <= REL_TOL


CASES = [
    ("s16 default", dict()),
    ("s16 one table + cold", dict(max_tables=1)),
    ("s16 lane_max 4 (long sub-rows)", dict(lane_max=4)),
    ("s16 lane_max 64", dict(lane_max=64)),
    ("s16 hubs chunked", dict(low_max=24, hub_min=512, hub_chunk=256)),
]


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_reference_mode_parity(ctx, name, kw):
    src, dst = ag.gen_rmat(77, 16, 16)
    n = 1 << 16
    print(f"case: {name} {kw}", flush=True)
    with tuned(layout="tables", **kw):
        g = ag.Graph.build(ctx, src, dst, n)
        g.prepare()
        info = g.plan_info()
        st = g.pagerank_pull(iterations=20, epsilon=0.0)
        st_eps = g.pagerank_pull(iterations=100, epsilon=1e-4)
    assert info["layout"] == ag._lib.AG_LAYOUT_TABLES
    print(f"plan: tables {info['tables']} subrows {info['subrows']} cold {info['cold_edges']} "
          f"wg {info['table_workgroups']} hub chunks {info['hub_chunks']}", flush=True)
    if "max_tables" in kw:
        assert info["tables"] == 1 and info["cold_edges"] > 0
    if "hub_min" in kw:
        assert info["n_hub"] > 0 and info["hub_chunks"] > info["n_hub"]
    csr = oracle.build_csr(src, dst, n)
    ref, it, err, psum, act, _ = oracle.pagerank_pull_f32(csr, 20, 0.0)
    f64, _, _ = oracle.pagerank_pull_f64(csr, 20, 0.0)
    _check(st.pageRanks, ref, f64)
    assert st.iterations == it
    np.testing.assert_array_equal(st.active_per_iter, act)
    assert abs(st.pr_sum - psum) <= 1e-6 * psum
    # epsilon 1e-4: both stop; the stopping sweep depends on when the last
    # vertex's |delta| (unnormalised ranks up to ~1e2, FP32 ulp ~1e-5) crosses
    # epsilon, which another summation order shifts by a few sweeps on skewed
    # graphs; the early sweeps' active counts agree
    ref_e, it_e, _, _, act_e, _ = oracle.pagerank_pull_f32(csr, 100, 1e-4)
    assert st_eps.iterations < 100 and st_eps.activeVertices == 0
    assert abs(st_eps.iterations - it_e) <= 10
    k = min(10, it_e)
    assert np.abs(st_eps.active_per_iter[:k].astype(np.int64) -
                  act_e[:k].astype(np.int64)).max() <= 8


@pytest.mark.parametrize("max_tables", [20, 40])
def test_many_tables_grouped_records(ctx, max_tables):
    """More tables than one combine group holds (17): the records come in
    groups loaded one after another and the partials are still added in table
    order.  R-MAT scale 21 has enough sources for 40 tables of 40,440."""
    src, dst = ag.gen_rmat(2021, 21, 16)
    n = 1 << 21
    print(f"case: s21 max_tables={max_tables}", flush=True)
    with tuned(layout="tables", max_tables=max_tables):
        g = ag.Graph.build(ctx, src, dst, n)
        g.prepare()
        info = g.plan_info()
        st = g.pagerank_pull(iterations=20, epsilon=0.0)
    print(f"plan: tables {info['tables']} subrows {info['subrows']} cold {info['cold_edges']}")
    assert info["tables"] > 17
    csr = oracle.build_csr(src, dst, n)
    ref, it, err, psum, act, _ = oracle.pagerank_pull_f32(csr, 20, 0.0)
    f64, _, _ = oracle.pagerank_pull_f64(csr, 20, 0.0)
    _check(st.pageRanks, ref, f64)
    np.testing.assert_array_equal(st.active_per_iter, act)
    # the same sums as with the default table count, up to summation order
    g.free()
    g2 = ag.Graph.build(ctx, src, dst, n)
    base = g2.pagerank_pull(iterations=20, epsilon=0.0).pageRanks
    assert _rel(st.pageRanks, base.astype(np.float64)).max() <= 2 * REL_TOL


def test_single_table_low_rows_bit_exact(ctx):
    """All sources in table 0 and every row short: each row is one sub-row
    summed by one lane in the reference's edge order."""
    src, dst = ag.gen_uniform(99, 4096, 20000)
    with tuned(layout="tables"):
        g = ag.Graph.build(ctx, src, dst, 4096)
        g.prepare()
        assert g.plan_info()["tables"] == 1 and g.plan_info()["cold_edges"] == 0
        st = g.pagerank_pull(20, 0.0)
    ref, *_ = oracle.pagerank_pull_f32(oracle.build_csr(src, dst, 4096), 20, 0.0)
    np.testing.assert_array_equal(st.pageRanks, ref)


def test_tables_vs_rows_layout(ctx):
    """Both layouts on the same graph: within 1e-5 of each other, and the
    tables' ranks as close to FP64 as the rows'."""
    src, dst = ag.gen_rmat(5, 18, 16)
    n = 1 << 18
    out = {}
    for lay in ("rows", "tables"):
        with tuned(layout=lay):
            g = ag.Graph.build(ctx, src, dst, n)
            out[lay] = g.pagerank_pull(20, 0.0).pageRanks
            g.free()
    f64, _, _ = oracle.pagerank_pull_f64(oracle.build_csr(src, dst, n), 20, 0.0)
    assert _rel(out["tables"], out["rows"]).max() <= 2 * REL_TOL
    assert _rel(out["tables"], f64).max() <= REL_TOL


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_low_precision_finish(ctx, prec):
    src, dst = ag.gen_rmat(31, 16, 16)
    n = 1 << 16
    with tuned(layout="tables", max_tables=1):
        g = ag.Graph.build(ctx, src, dst, n)
        st = g.pagerank_pull(iterations=200, epsilon=1e-7, mode="standard", precision=prec,
                             finish_fp32=True)
    x64, _, _ = oracle.pagerank_standard_f64(oracle.build_csr(src, dst, n), 2000, 1e-14)
    assert st.fp32_switch_sweep > 0
    assert np.abs(st.pageRanks.astype(np.float64) - x64).sum() <= 5e-6
    assert abs(st.pageRanks.astype(np.float64).sum() - 1.0) <= 1e-5


def test_data_driven_and_spmv(ctx):
    src, dst = ag.gen_rmat(13, 16, 16)
    n = 1 << 16
    csr = oracle.build_csr(src, dst, n)
    with tuned(layout="tables", max_tables=1):
        g = ag.Graph.build(ctx, src, dst, n)
        a = g.pagerank_pull(20, 0.0)
        b = g.pagerank_pull(20, 0.0, mode="data_driven")
        x = ag.spmv_input(csr["vertices.out_degree"])
        y, _ = g.spmv_pull(x, 3)
    # epsilon 0: the data-driven run equals the pull run bit for bit
    np.testing.assert_array_equal(a.pageRanks, b.pageRanks)
    y64 = np.bincount(dst, weights=x[src].astype(np.float64), minlength=n)
    assert np.all(np.abs(y - y64) <= 1e-5 * np.maximum(y64, 1e-30))
    dd, it, _, _, _, _, _ = oracle.pagerank_dd_pull_f32(csr, 30, 1e-4)
    c = g.pagerank_pull(30, 1e-4, mode="data_driven")
    assert c.iterations > 0 and np.abs(c.pageRanks - dd).sum() <= 2 * 1e-4 / 0.15


def test_personalized_single(ctx):
    src, dst = ag.gen_rmat(17, 16, 16)
    n = 1 << 16
    csr = oracle.build_csr(src, dst, n)
    seeds = np.array([3, 9000, 41000, 65000], np.uint32)
    with tuned(layout="tables", max_tables=1):
        g = ag.Graph.build(ctx, src, dst, n)
        x, st = ag.ppr_single(g, seeds, 1e-7, precision="fp32")
    ref, _, _ = oracle.ppr_f64(csr, seeds, 1e-12, 5000)
    assert np.abs(x - ref).sum() <= 0.85 / 0.15 * 1e-7 + 2e-5


def test_c1_tables_parity_and_epsilon(ctx):
    """C1 (N=65,536 > one table): rows are split over two tables, so sums are
    not in the reference's order; 20 sweeps hold the 1e-5 / FP64 bar, SPMV
    likewise, and at epsilon 1e-4 the pull and data-driven runs stop within a
    few sweeps of the oracle on an epsilon-fixpoint."""
    src, dst = ag.gen_uniform(27491095, 65536, 1048576)
    n = 65536
    csr = oracle.build_csr(src, dst, n)
    with tuned(layout="tables"):
        g = ag.Graph.build(ctx, src, dst, n)
        st = g.pagerank_pull(20, 0.0)
        se = g.pagerank_pull(100, 1e-4)
        sd = g.pagerank_pull(100, 1e-4, mode="data_driven")
        x = ag.spmv_input(csr["vertices.out_degree"])
        y, _ = g.spmv_pull(x, 1)
    ref, it, _, _, act, _ = oracle.pagerank_pull_f32(csr, 20, 0.0)
    f64, _, _ = oracle.pagerank_pull_f64(csr, 20, 0.0)
    _check(st.pageRanks, ref, f64)
    np.testing.assert_array_equal(st.active_per_iter, act)
    _, it_e, _, _, _, _ = oracle.pagerank_pull_f32(csr, 100, 1e-4)
    _, it_d, _, _, _, _, _ = oracle.pagerank_dd_pull_f32(csr, 100, 1e-4)
    for s_, it_o in ((se, it_e), (sd, it_d)):
        assert s_.activeVertices == 0 and abs(s_.iterations - it_o) <= 5
    y32 = oracle.spmv_pull_f32(csr, x)
    y64 = np.bincount(dst, weights=x[src].astype(np.float64), minlength=n)
    _check(y, y32, y64)
