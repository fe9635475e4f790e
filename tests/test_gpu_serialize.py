"""CSR binary serialization on the device: the library writes exactly the
documented AGCSRv1 bytes (compared with the host-side writer of the tests),
loads them back bit for bit, and PageRank on a loaded graph equals PageRank
on the graph it was saved from."""
import numpy as np
import pytest

import accelgraph_express_amd as ag
import oracle
from test_serialize_host import MASK_FULL, write_agcsr

pytestmark = pytest.mark.gpu


def _graph(ctx, n=65536, e=1048576, seed=27491095):
    src, dst = ag.gen_uniform(seed, n, e)
    return ag.Graph.build(ctx, src, dst, num_vertices=n), src, dst


def test_round_trip_bit_exact_and_pagerank_equal(ctx, tmp_path):
    g, src, dst = _graph(ctx)
    p = str(tmp_path / "c1.agcsr")
    print("case: save/load C1", flush=True)
    g.save(p)
    info = ag.csr_file_info(p)
    assert info == {"num_vertices": 65536, "num_edges": 1048576, "sections": MASK_FULL}
    h = ag.Graph.load(ctx, p)
    a, b = g.download(), h.download()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    sa = g.pagerank_pull(iterations=20, epsilon=1e-4)
    sb = h.pagerank_pull(iterations=20, epsilon=1e-4)
    np.testing.assert_array_equal(sa.pageRanks, sb.pageRanks)
    assert sa.iterations == sb.iterations


def test_library_bytes_equal_documented_layout(ctx, tmp_path):
    rng = np.random.RandomState(8)
    n, e = 3000, 40000
    src, dst = rng.randint(0, n, e).astype(np.uint32), rng.randint(0, n, e).astype(np.uint32)
    g = ag.Graph.build(ctx, src, dst, num_vertices=n)
    p_lib, p_doc = str(tmp_path / "lib.agcsr"), str(tmp_path / "doc.agcsr")
    g.save(p_lib)
    write_agcsr(p_doc, oracle.build_csr(src, dst, n), n, e)
    assert open(p_lib, "rb").read() == open(p_doc, "rb").read()
    h = ag.Graph.load(ctx, p_doc)
    want = oracle.build_csr(src, dst, n)
    got = h.download()
    for k in oracle.CSR_KEYS:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)


def test_pagerank_only_graph_round_trip(ctx, tmp_path):
    rng = np.random.RandomState(9)
    n, e = 2000, 30000
    src, dst = rng.randint(0, n, e), rng.randint(0, n, e)
    csr = oracle.build_csr(src, dst, n)
    g = ag.Graph.from_csr(ctx, n, csr["vertices.out_degree"], csr["inverse_vertices.out_degree"],
                          csr["inverse_vertices.edges_idx"],
                          csr["inverse_sorted_edges_array.edges_array_dest"])
    p = str(tmp_path / "pr.agcsr")
    g.save(p)
    assert ag.csr_file_info(p)["sections"] == 0x1C4
    h = ag.Graph.load(ctx, p)
    np.testing.assert_array_equal(g.pagerank_pull(iterations=10, epsilon=0.0).pageRanks,
                                  h.pagerank_pull(iterations=10, epsilon=0.0).pageRanks)


def test_out_of_range_vertex_refused(ctx, tmp_path):
    rng = np.random.RandomState(10)
    n, e = 100, 500
    src, dst = rng.randint(0, n, e), rng.randint(0, n, e)
    csr = oracle.build_csr(src, dst, n)
    csr["sorted_edges_array.edges_array_dest"] = csr["sorted_edges_array.edges_array_dest"].copy()
    csr["sorted_edges_array.edges_array_dest"][7] = n + 3  # consistent checksum, bad id
    p = str(tmp_path / "bad.agcsr")
    write_agcsr(p, csr, n, e)
    ag.csr_file_info(p)  # the container itself is well formed
    with pytest.raises(ag.AGError):
        ag.Graph.load(ctx, p)


@pytest.mark.parametrize("defect", ["idx_past_E", "offsets_decrease", "degree_vs_offsets",
                                    "inverse_idx_past_E"])
def test_inconsistent_ranges_refused_before_upload(ctx, tmp_path, defect):
    """A file with a correct checksum whose per-vertex ranges leave the edge
    array is refused on the host (the gather and push kernels would read and
    write outside it)."""
    rng = np.random.RandomState(11)
    n, e = 200, 3000
    src, dst = rng.randint(0, n, e), rng.randint(0, n, e)
    csr = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in oracle.build_csr(src, dst, n).items()}
    deg = csr["vertices.out_degree"]
    v = int(np.argmax(deg > 0))
    if defect == "idx_past_E":
        csr["vertices.edges_idx"][v] = e - 1
    elif defect == "offsets_decrease":
        off = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint32)
        off[v + 1], off[v + 2] = off[v + 2] + 1, off[v + 1]
        csr["offsets"] = off
    elif defect == "degree_vs_offsets":
        # degrees still sum to E, but no longer match the offsets
        csr["offsets"] = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint32)
        w = int(np.nonzero(deg > 0)[0][1])
        d2 = deg.copy()
        d2[v] += 1
        d2[w] -= 1
        csr["vertices.out_degree"] = d2
    else:
        csr["inverse_vertices.edges_idx"][int(np.argmax(csr["inverse_vertices.out_degree"]))] = e
    p = str(tmp_path / f"{defect}.agcsr")
    write_agcsr(p, csr, n, e)
    ag.csr_file_info(p)  # checksum and size are consistent
    with pytest.raises(ag.AGError) as ei:
        ag.Graph.load(ctx, p)
    assert "CSR" in str(ei.value)
