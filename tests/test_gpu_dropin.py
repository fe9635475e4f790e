"""The drop-in C API end to end: a C program written against the
reference-named headers only (tests/c/dropin_pagerank.c) builds GraphCSRs two
ways and runs pageRankGraphCSR; CSR arrays must equal the oracle's bit for bit
and ranks must be within REL_TOL of the oracle's FP32 restatement."""
import os
import subprocess

import numpy as np
import pytest

import accelgraph_express_amd as ag
import oracle

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "tests", "c", "_bin", "dropin_pagerank")
REL_TOL = 1e-5


def _run(path, fmt, iterations, epsilon, tmp_path, pushpull=0):
    out = str(tmp_path / "out.bin")
    print(f"case: dropin {os.path.basename(os.path.dirname(path))}/{os.path.basename(path)} "
          f"fmt={fmt} pushpull={pushpull}", flush=True)
    env = dict(os.environ, AG_QUIET="1")
    r = subprocess.run([BIN, path, str(fmt), str(iterations), str(epsilon), out, str(pushpull)],
                       capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(out, "rb").read()
    V, E, it, same = np.frombuffer(raw[:16], np.uint32)
    ea, eb = np.frombuffer(raw[16:32], np.float64)
    itb = np.frombuffer(raw[32:36], np.uint32)[0]
    off = 36
    ra = np.frombuffer(raw[off:off + 4 * V], np.float32); off += 4 * V
    rb = np.frombuffer(raw[off:off + 4 * V], np.float32); off += 4 * V
    fdst = np.frombuffer(raw[off:off + 4 * E], np.uint32); off += 4 * E
    idst = np.frombuffer(raw[off:off + 4 * E], np.uint32); off += 4 * E
    iidx = np.frombuffer(raw[off:off + 4 * V], np.uint32)
    return dict(V=V, E=E, it=it, itb=itb, same=same, ea=ea, eb=eb, ra=ra, rb=rb, fdst=fdst,
                idst=idst, iidx=iidx)


def _read_list(path, fmt):
    if fmt == 0:
        a = np.loadtxt(path, dtype=np.int64).reshape(-1, 2)
    else:
        a = np.fromfile(path, dtype="<i4").reshape(-1, 3 if fmt == 2 else 2)[:, :2]
    return a[:, 0].astype(np.uint32), a[:, 1].astype(np.uint32)


@pytest.mark.parametrize("rel,fmt", [("v51_e1021/graph", 0), ("v51_e1021/graph.bin", 1),
                                     ("v300_e2730/graph.wbin", 2), ("graphbrew/graph.bin", 1)])
def test_reference_fixture_graphs(rel, fmt, tmp_path):
    path = os.path.join(GOLDEN, "reference_graphs", rel)
    r = _run(path, fmt, 20, 0.0, tmp_path)
    src, dst = _read_list(path, fmt)
    csr = oracle.build_csr(src, dst)
    assert r["same"] == 1
    np.testing.assert_array_equal(r["fdst"], csr["sorted_edges_array.edges_array_dest"])
    np.testing.assert_array_equal(r["idst"], csr["inverse_sorted_edges_array.edges_array_dest"])
    np.testing.assert_array_equal(r["iidx"], csr["inverse_vertices.edges_idx"])
    ref, it, err, *_ = oracle.pagerank_pull_f32(csr, 20, 0.0)
    assert r["it"] == it == r["itb"]
    for ranks in (r["ra"], r["rb"]):
        rel_err = np.abs(ranks.astype(np.float64) - ref) / ref
        assert rel_err.max() <= REL_TOL


def test_c1_size_binary_file(tmp_path):
    src, dst = ag.gen_uniform(27491095, 65536, 1048576)
    path = str(tmp_path / "c1.bin")
    np.stack([src, dst], 1).astype("<i4").tofile(path)
    r = _run(path, 1, 100, 1e-4, tmp_path)  # the reference CLI's defaults: epsilon 1e-4
    csr = oracle.build_csr(src, dst)
    assert r["same"] == 1
    np.testing.assert_array_equal(r["fdst"], csr["sorted_edges_array.edges_array_dest"])
    ref, it, err, *_ = oracle.pagerank_pull_f32(csr, 100, 1e-4)
    assert r["it"] == it
    assert np.abs(r["ra"].astype(np.float64) - ref).max() / ref.min() <= REL_TOL * 10


@pytest.mark.parametrize("rel,fmt", [("v51_e1021/graph", 0), ("v300_e2730/graph.wbin", 2),
                                     ("graphbrew/graph.bin", 1)])
def test_data_driven_pull_dispatch(rel, fmt, tmp_path):
    """pushpull 6 routes pageRankGraphCSR to pageRankDataDrivenPullGraphCSR
    (pageRank.c:900-902): iteration count and ranks follow the data-driven
    oracle, which differs from plain pull once epsilon stops sub-epsilon
    updates."""
    path = os.path.join(GOLDEN, "reference_graphs", rel)
    src, dst = _read_list(path, fmt)
    csr = oracle.build_csr(src, dst)
    for eps in (0.0, 1e-4):
        r = _run(path, fmt, 50, eps, tmp_path, pushpull=6)
        ref, it, *_ = oracle.pagerank_dd_pull_f32(csr, 50, eps)
        assert r["it"] == it == r["itb"]
        for ranks in (r["ra"], r["rb"]):
            rel_err = np.abs(ranks.astype(np.float64) - ref) / ref
            assert rel_err.max() <= REL_TOL
