"""GPU CSR builder parity: every array bit-exact against the oracle.

Covers the configurations' shapes (C1 uniform N=65,536 E=1,048,576; R-MAT
scale 22 E=67,108,864 at full size), every sort mode, the inverse built from
input order, and the edge cases the builder meets: empty lists, one edge,
self-loops, duplicates, isolated and trailing vertices, hub vertices."""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

import accelgraph_express_amd as ag
import oracle

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _cmp(ctx, src, dst, n=0, sort=0, from_input=False):
    print(f"case: N={n} E={len(src)} sort={sort} from_input={from_input}", flush=True)
    g = ag.Graph.build(ctx, src, dst, num_vertices=n, sort=sort, inverse_from_input=from_input)
    got = g.download()
    nn = g.num_vertices
    want = oracle.build_csr(src, dst, nn, sort, from_input)
    for k in oracle.CSR_KEYS:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    g.free()
    return nn


def test_c1_uniform_bit_exact_and_golden(ctx):
    golden = json.load(open(os.path.join(GOLDEN, "c1_uniform.json")))
    src, dst = ag.gen_uniform(27491095, 65536, 1048576)
    assert hashlib.sha256(src.tobytes()).hexdigest() == golden["input_sha256"]["src"]
    g = ag.Graph.build(ctx, src, dst, num_vertices=65536)
    got = g.download()
    for k, h in golden["csr_sha256"].items():
        assert hashlib.sha256(np.ascontiguousarray(got[k]).tobytes()).hexdigest() == h, k
    _cmp(ctx, src, dst, 65536)


@pytest.mark.parametrize("sort", [0, 1, 2, 3])
@pytest.mark.parametrize("from_input", [False, True])
def test_sort_modes(ctx, sort, from_input):
    rng = np.random.RandomState(10 + sort)
    src = rng.randint(0, 3000, 50000).astype(np.uint32)
    dst = rng.randint(0, 3000, 50000).astype(np.uint32)
    _cmp(ctx, src, dst, 3000, sort, from_input)


def test_edge_cases(ctx):
    assert _cmp(ctx, np.zeros(0, np.uint32), np.zeros(0, np.uint32), 5) == 5
    assert _cmp(ctx, np.zeros(0, np.uint32), np.zeros(0, np.uint32), 0) == 0
    assert _cmp(ctx, np.array([3], np.uint32), np.array([3], np.uint32)) == 4  # self loop
    # duplicates + isolated vertices + a trailing isolated vertex id
    _cmp(ctx, np.array([2, 2, 2, 0, 2], np.uint32), np.array([1, 1, 1, 1, 5], np.uint32), 9)
    # one vertex: 0-bit keys (the sort is a copy), ids checked after it
    for sort in range(4):
        assert _cmp(ctx, np.zeros(5000, np.uint32), np.zeros(5000, np.uint32), 1, sort) == 1
    with pytest.raises(ag.AGError) as e:
        ag.Graph.build(ctx, np.zeros(3, np.uint32), np.array([0, 1, 0], np.uint32), 1)
    assert e.value.code == ag._lib.AG_ERR_INVALID
    # one hub with every edge (one segment spanning many tiles)
    m = 200000
    _cmp(ctx, np.full(m, 7, np.uint32), np.arange(m, dtype=np.uint32) % 1000, 1000)
    # sizes that straddle the 4096-key tile
    rng = np.random.RandomState(4)
    for m in (4095, 4096, 4097, 8191, 12289):
        _cmp(ctx, rng.randint(0, 70000, m).astype(np.uint32),
             rng.randint(0, 70000, m).astype(np.uint32), 70000)


def test_random_small_graphs_bit_exact(ctx):
    """300 random edge lists on at most 6 vertices with at most 10 edges, every
    sort mode, against the oracle (which the CPU tests pin to the closed forms)."""
    rng = np.random.RandomState(91)
    print("case: 300 small random graphs", flush=True)
    for i in range(300):
        n = rng.randint(1, 7)
        m = rng.randint(0, 11)
        src = rng.randint(0, n, m).astype(np.uint32)
        dst = rng.randint(0, n, m).astype(np.uint32)
        sort = i % 4
        g = ag.Graph.build(ctx, src, dst, num_vertices=n, sort=sort)
        got = g.download()
        want = oracle.build_csr(src, dst, n, sort, False)
        for k in oracle.CSR_KEYS:
            np.testing.assert_array_equal(got[k], want[k], err_msg=f"{k} graph {i}")
        g.free()


def test_vertex_out_of_range_rejected(ctx):
    with pytest.raises(ag.AGError) as e:
        ag.Graph.build(ctx, np.array([1, 9], np.uint32), np.array([0, 0], np.uint32), 5)
    assert e.value.code == ag._lib.AG_ERR_INVALID
    # With num_vertices given, the ids are checked after the sort (the first
    # radix pass of each direction takes the largest key): one id past N among
    # 1 M edges, as a source or a destination, in every sort mode and with a
    # forward-only build, is refused with the id named, and a good build
    # follows.
    n, m = 100000, 1 << 20
    rng = np.random.RandomState(3)
    for side in ("src", "dst"):
        for sort in range(4):
            src = rng.randint(0, n, m).astype(np.uint32)
            dst = rng.randint(0, n, m).astype(np.uint32)
            (src if side == "src" else dst)[m // 2 + sort] = n + 7
            with pytest.raises(ag.AGError) as e:
                ag.Graph.build(ctx, src, dst, n, sort=sort)
            assert e.value.code == ag._lib.AG_ERR_INVALID
            assert f"vertex id {n + 7} >= num_vertices {n}" in str(e.value)
        lib = ag._lib.load()
        h = C.c_void_p()
        rc = lib.ag_csr_build(ctx.h, src.ctypes.data_as(C.c_void_p), dst.ctypes.data_as(C.c_void_p),
                              m, n, ag._lib.AG_BUILD_NO_INVERSE, C.byref(h))
        assert rc == ag._lib.AG_ERR_INVALID
    src = rng.randint(0, n, m).astype(np.uint32)
    dst = rng.randint(0, n, m).astype(np.uint32)
    _cmp(ctx, src, dst, n)


def test_rmat_s22_full_size_bit_exact(ctx):
    src, dst = ag.gen_rmat(2022, 22, 16)
    n = 1 << 22
    _cmp(ctx, src, dst, n)


def test_deterministic_repeat(ctx):
    src, dst = ag.gen_rmat(5, 18, 16)
    a = ag.Graph.build(ctx, src, dst, 1 << 18).download()
    b = ag.Graph.build(ctx, src, dst, 1 << 18).download()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
