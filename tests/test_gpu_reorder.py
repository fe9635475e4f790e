"""GPU locality relabelling (the reference's -l/-L/-O options, README.md:362-411)
against the oracle.

Bar: labels are integer work, so every mode's labels equal the oracle's bit
for bit (the oracle restates the published algorithms; parity with the absent
OpenGraph module's exact labels is unpinned, see test_oracle_reorder.py).
Relabelled edge lists equal labels[src], labels[dst]; labels that are not a
permutation are refused.  Through the drop-in: graphCSRPreProcessingStep with
one to three layers (and mode 11, labels from the reference's Rabbit-order
file) gives the composed label_array / inverse_label_array, CSR arrays equal
to the oracle's build of the relabelled list, and PageRank equal to the
unrelabelled graph's read through the labels (within 1e-5)."""
import os
import subprocess

import numpy as np
import pytest

import accelgraph_express_amd as ag
import oracle

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "tests", "c", "_bin", "dropin_reorder")


def _skewed(seed, n, m):
    rng = np.random.RandomState(seed)
    src = (rng.zipf(1.6, m) % n).astype(np.uint32)
    dst = ((rng.zipf(1.6, m) * 7919) % n).astype(np.uint32)
    return src, dst


@pytest.mark.parametrize("mode", list(range(11)))
def test_labels_equal_oracle(ctx, mode):
    src, dst = _skewed(1, 5000, 80000)
    print(f"case: relabel mode {mode} N=5000 E=80000", flush=True)
    got = ag.reorder_labels(ctx, src, dst, 5000, mode, seed=77)
    np.testing.assert_array_equal(got, oracle.reorder_labels(src, dst, 5000, mode, seed=77))


@pytest.mark.parametrize("mode", [1, 5, 6, 9, 10])
def test_rmat_labels_and_edges(ctx, mode):
    src, dst = ag.gen_rmat(12, 16, 16)
    n = 1 << 16
    lab = ag.reorder_labels(ctx, src, dst, n, mode, seed=3)
    np.testing.assert_array_equal(lab, oracle.reorder_labels(src, dst, n, mode, seed=3))
    s2, d2 = ag.relabel_edges(ctx, lab, src, dst)
    np.testing.assert_array_equal(s2, lab[src])
    np.testing.assert_array_equal(d2, lab[dst])


def test_refusals(ctx):
    src = np.array([0, 1, 2], np.uint32)
    dst = np.array([1, 2, 3], np.uint32)
    with pytest.raises(ag.AGError):
        ag.relabel_edges(ctx, np.array([0, 1, 1, 3], np.uint32), src, dst)  # repeats 1
    with pytest.raises(ag.AGError):
        ag.relabel_edges(ctx, np.array([0, 1, 2, 4], np.uint32), src, dst)  # 4 >= N
    with pytest.raises(ag.AGError):
        ag.reorder_labels(ctx, src, dst, 4, 11)
    with pytest.raises(ag.AGError):
        ag.reorder_labels(ctx, src, np.array([1, 2, 9], np.uint32), 4, 1)  # id >= N
    assert ag.reorder_labels(ctx, np.zeros(0, np.uint32), np.zeros(0, np.uint32), 5, 1).tolist() \
        == [0, 1, 2, 3, 4]


def _run(path, fmt, layers, label_file, tmp_path):
    out = str(tmp_path / "reorder.bin")
    print(f"case: dropin_reorder {os.path.basename(path)} layers {layers}", flush=True)
    r = subprocess.run([BIN, path, str(fmt)] + [str(m) for m in layers] +
                       [label_file or "-", out], capture_output=True, text=True,
                       env=dict(os.environ, AG_QUIET="1"), timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, np.uint32)
    V, E = int(raw[0]), int(raw[1])
    o = 2
    parts = {}
    for k, z in (("lab", V), ("inv", V), ("ilab", V), ("iinv", V), ("fdst", E), ("idst", E)):
        parts[k] = raw[o:o + z]
        o += z
    parts["ranks"] = raw[o:o + V].view(np.float32)
    return V, E, parts


def _read(path, fmt):
    if fmt == 0:
        a = np.loadtxt(path, dtype=np.int64).reshape(-1, 2)
    else:
        a = np.fromfile(path, dtype="<i4").reshape(-1, 3 if fmt == 2 else 2)[:, :2]
    return a[:, 0].astype(np.uint32), a[:, 1].astype(np.uint32)


@pytest.mark.parametrize("rel,fmt,layers", [
    ("v300_e2730/graph.wbin", 2, (5, 0, 0)),
    ("v300_e2730/graph.wbin", 2, (6, 2, 0)),
    ("v51_e1021/graph", 0, (9, 4, 10)),
    ("graphbrew/graph.org", 0, (11, 0, 0)),
    ("graphbrew/graph.org", 0, (11, 1, 0)),
])
def test_dropin_layers(rel, fmt, layers, tmp_path):
    path = os.path.join(GOLDEN, "reference_graphs", rel)
    label_file = os.path.join(GOLDEN, "reference_graphs", "graphbrew", "graph_Rabbit.org.labels")
    V, E, p = _run(path, fmt, layers, label_file if 11 in layers else None, tmp_path)
    src, dst = _read(path, fmt)
    n = int(max(src.max(), dst.max())) + 1
    assert V == n and E == src.size
    total = np.arange(n, dtype=np.uint32)
    s, d = src, dst
    for m in layers:
        if m == 0:
            continue
        if m == 11:
            lab = np.loadtxt(label_file, dtype=np.uint32)
        else:
            lab = oracle.reorder_labels(s, d, n, m, seed=27491095)
        s, d = lab[s], lab[d]
        total = lab[total]
    np.testing.assert_array_equal(p["lab"], total)
    np.testing.assert_array_equal(p["ilab"], total)
    np.testing.assert_array_equal(p["inv"][total], np.arange(n))
    np.testing.assert_array_equal(p["iinv"], p["inv"])
    csr = oracle.build_csr(s, d, n)
    np.testing.assert_array_equal(p["fdst"], csr["sorted_edges_array.edges_array_dest"])
    np.testing.assert_array_equal(p["idst"], csr["inverse_sorted_edges_array.edges_array_dest"])
    ref, *_ = oracle.pagerank_pull_f32(oracle.build_csr(src, dst, n), 20, 0.0)
    rel_err = np.abs(p["ranks"][total].astype(np.float64) - ref) / ref
    assert rel_err.max() <= 1e-5


def test_degenerate_graphs(ctx):
    # one vertex, self-loops only, and a vertex id range with isolated vertices
    for src, dst, n in ((np.array([0], np.uint32), np.array([0], np.uint32), 1),
                        (np.array([3, 3, 3], np.uint32), np.array([3, 3, 3], np.uint32), 5),
                        (np.array([0, 9], np.uint32), np.array([9, 0], np.uint32), 10)):
        for mode in range(11):
            got = ag.reorder_labels(ctx, src, dst, n, mode, seed=5)
            np.testing.assert_array_equal(got, oracle.reorder_labels(src, dst, n, mode, seed=5))
