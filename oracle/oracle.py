"""ctypes wrapper of oracle/liboracle.so (TEST INFRASTRUCTURE ONLY).

Each function mirrors one C function of oracle.h; the reference lines it
restates are cited there.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    U, D, I = C.c_uint32, C.c_double, C.c_int
    sig = {
        "oracle_num_vertices": (U, [U, u32p, u32p]),
        "oracle_count_sort": (None, [U, U, u32p, u32p, u32p, u32p]),
        "oracle_vertex_map": (None, [U, U, u32p, u32p, u32p]),
        "oracle_build_csr": (None, [U, U, u32p, u32p, U] + [u32p] * 10),
        "oracle_pagerank_pull_f32": (U, [U, u32p, u32p, u32p, u32p, U, D, f32p, f64p, f64p,
                                         u32p, f64p]),
        "oracle_pagerank_pull_f32_from": (U, [U, u32p, u32p, u32p, u32p, U, D, f32p, f32p, f64p,
                                              f64p, u32p, f64p]),
        "oracle_pagerank_pull_f64": (U, [U, u32p, u32p, u32p, u32p, U, D, f64p, f64p]),
        "oracle_spmv_input": (None, [U, u32p, f32p]),
        "oracle_reorder_labels": (I, [U, U, u32p, u32p, U, U, u32p]),
        "oracle_spmv_pull_f32": (None, [U, u32p, u32p, u32p, f32p, f32p]),
        "oracle_spmv_sum": (D, [U, f32p]),
        "oracle_pagerank_dd_pull_f32": (U, [U] + [u32p] * 6 + [U, D, f32p, f64p, f64p, u32p,
                                                              f64p, u32p]),
        "oracle_pagerank_dd_pull_f32_from": (U, [U] + [u32p] * 6 + [U, D, f32p, f32p, f64p, f64p,
                                                                   u32p, f64p, u32p]),
        "oracle_pagerank_standard_f64": (U, [U, u32p, u32p, u32p, u32p, D, U, D, f64p, f64p]),
        "oracle_pagerank_standard_f64_from": (U, [U, u32p, u32p, u32p, u32p, D, U, D, f64p, f64p,
                                                  f64p]),
        "oracle_ppr_f64": (U, [U, u32p, u32p, u32p, u32p, D, u32p, U, D, U, f64p, f64p]),
        "oracle_csr_build_omp": (D, [U, U, u32p, u32p] + [u32p] * 8 + [I]),
        "oracle_pagerank_pull_omp": (D, [U, u32p, u32p, u32p, u32p, U, D, f32p, f64p, u32p, I]),
        "oracle_omp_max_threads": (I, []),
    }
    for k, (r, a) in sig.items():
        fn = getattr(lib, k)
        fn.restype = r
        fn.argtypes = a
    _lib = lib
    return lib


def _u(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _p(a, t=C.c_uint32):
    return a.ctypes.data_as(C.POINTER(t))


def num_vertices(src, dst) -> int:
    s, d = _u(src), _u(dst)
    return int(load().oracle_num_vertices(s.size, _p(s), _p(d)))


def count_sort(n: int, key, val):
    k, v = _u(key), _u(val)
    ko, vo = np.empty_like(k), np.empty_like(v)
    load().oracle_count_sort(n, k.size, _p(k), _p(v), _p(ko), _p(vo))
    return ko, vo


def vertex_map(n: int, sorted_key):
    k = _u(sorted_key)
    deg, idx = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    load().oracle_vertex_map(n, k.size, _p(k), _p(deg), _p(idx))
    return deg, idx


CSR_KEYS = ["sorted_edges_array.edges_array_src", "sorted_edges_array.edges_array_dest",
            "vertices.out_degree", "vertices.in_degree", "vertices.edges_idx",
            "inverse_sorted_edges_array.edges_array_src",
            "inverse_sorted_edges_array.edges_array_dest", "inverse_vertices.out_degree",
            "inverse_vertices.in_degree", "inverse_vertices.edges_idx"]


def build_csr(src, dst, n: int | None = None, sort_mode: int = 0,
              inverse_from_input: bool = False) -> dict:
    s, d = _u(src), _u(dst)
    if n is None:
        n = num_vertices(s, d)
    m = s.size
    sizes = [m, m, n, n, n, m, m, n, n, n]
    out = {k: np.zeros(z, np.uint32) for k, z in zip(CSR_KEYS, sizes)}
    mode = sort_mode | (0x100 if inverse_from_input else 0)
    load().oracle_build_csr(n, m, _p(s), _p(d), mode, *[_p(out[k]) for k in CSR_KEYS])
    out["num_vertices"] = n
    out["num_edges"] = m
    return out


def _inv(csr):
    return (_u(csr["vertices.out_degree"]), _u(csr["inverse_vertices.out_degree"]),
            _u(csr["inverse_vertices.edges_idx"]),
            _u(csr["inverse_sorted_edges_array.edges_array_dest"]))


def _init32(init, n):
    if init is None:
        return None
    x = np.ascontiguousarray(init, np.float32)
    assert x.size == n
    return x


def pagerank_pull_f32(csr: dict, iterations: int = 20, epsilon: float = 1e-4, init=None):
    """Returns (ranks f32, iterations, error_total, pr_sum, active[], err[]).
    init (N floats, the returned ranks' scale) replaces the reference's start."""
    n = csr["num_vertices"]
    od, idg, iidx, idst = _inv(csr)
    x0 = _init32(init, n)
    pr = np.zeros(n, np.float32)
    et, ps = C.c_double(), C.c_double()
    act = np.zeros(max(iterations, 1), np.uint32)
    err = np.zeros(max(iterations, 1), np.float64)
    it = load().oracle_pagerank_pull_f32_from(n, _p(od), _p(idg), _p(iidx), _p(idst), iterations,
                                              epsilon, None if x0 is None else _p(x0, C.c_float),
                                              _p(pr, C.c_float), C.byref(et), C.byref(ps),
                                              _p(act), _p(err, C.c_double))
    sweeps = min(iterations, it + 1) if iterations else 0
    return pr, int(it), et.value, ps.value, act[:sweeps].copy(), err[:sweeps].copy()


def pagerank_dd_pull_f32(csr: dict, iterations: int = 20, epsilon: float = 1e-4, init=None):
    """Data-driven pull.  Returns (ranks f32, iterations, error_total, pr_sum,
    active[], err[], work[]) with per-sweep updated counts, error totals and
    worklist sizes.  init as in pagerank_pull_f32."""
    n = csr["num_vertices"]
    x0 = _init32(init, n)
    od, idg, iidx, idst = _inv(csr)
    fidx = _u(csr["vertices.edges_idx"])
    fdst = _u(csr["sorted_edges_array.edges_array_dest"])
    pr = np.zeros(n, np.float32)
    et, ps = C.c_double(), C.c_double()
    k = max(iterations, 1)
    act, err, work = np.zeros(k, np.uint32), np.zeros(k, np.float64), np.zeros(k, np.uint32)
    it = load().oracle_pagerank_dd_pull_f32_from(n, _p(od), _p(fidx), _p(fdst), _p(idg), _p(iidx),
                                            _p(idst), iterations, epsilon,
                                            None if x0 is None else _p(x0, C.c_float),
                                            _p(pr, C.c_float),
                                            C.byref(et), C.byref(ps), _p(act),
                                            _p(err, C.c_double), _p(work))
    sweeps = min(iterations, it + 1) if iterations else 0
    return (pr, int(it), et.value, ps.value, act[:sweeps].copy(), err[:sweeps].copy(),
            work[:sweeps].copy())


def reorder_labels(src, dst, n: int, mode: int, seed: int = 0) -> np.ndarray:
    """labels[v] = new label of v under relabelling `mode` (0-10)."""
    s, d = _u(src), _u(dst)
    lab = np.zeros(max(n, 1), np.uint32)
    rc = load().oracle_reorder_labels(n, s.size, _p(s), _p(d), mode, seed, _p(lab))
    if rc != 0:
        raise ValueError(f"unknown relabelling mode {mode}")
    return lab[:n]


def spmv_input(csr: dict) -> np.ndarray:
    """The reference's benchmark input vector: 1/out_degree, 0.001 without out-edges."""
    n = csr["num_vertices"]
    x = np.zeros(n, np.float32)
    load().oracle_spmv_input(n, _p(_u(csr["vertices.out_degree"])), _p(x, C.c_float))
    return x


def spmv_pull_f32(csr: dict, x) -> np.ndarray:
    """y[v] = sum of x over v's in-edges (inverse CSR order), FP32."""
    n = csr["num_vertices"]
    _, idg, iidx, idst = _inv(csr)
    xs = np.ascontiguousarray(x, np.float32)
    assert xs.size == n
    y = np.zeros(n, np.float32)
    load().oracle_spmv_pull_f32(n, _p(idg), _p(iidx), _p(idst), _p(xs, C.c_float),
                                _p(y, C.c_float))
    return y


def spmv_sum(y) -> float:
    ys = np.ascontiguousarray(y, np.float32)
    return float(load().oracle_spmv_sum(ys.size, _p(ys, C.c_float)))


def pagerank_pull_f64(csr: dict, iterations: int = 20, epsilon: float = 1e-4):
    n = csr["num_vertices"]
    od, idg, iidx, idst = _inv(csr)
    pr = np.zeros(n, np.float64)
    et = C.c_double()
    it = load().oracle_pagerank_pull_f64(n, _p(od), _p(idg), _p(iidx), _p(idst), iterations,
                                         epsilon, _p(pr, C.c_double), C.byref(et))
    return pr, int(it), et.value


def pagerank_standard_f64(csr: dict, iterations: int, tol: float = 0.0, damp: float = 0.85,
                          init=None):
    """STANDARD PageRank in FP64 from x0 = 1/N, or from init (N values)."""
    n = csr["num_vertices"]
    od, idg, iidx, idst = _inv(csr)
    x0 = None if init is None else np.ascontiguousarray(init, np.float64)
    assert x0 is None or x0.size == n
    x = np.zeros(n, np.float64)
    res = C.c_double()
    it = load().oracle_pagerank_standard_f64_from(n, _p(od), _p(idg), _p(iidx), _p(idst), damp,
                                                  iterations, tol,
                                                  None if x0 is None else _p(x0, C.c_double),
                                                  _p(x, C.c_double), C.byref(res))
    return x, int(it), res.value


def ppr_f64(csr: dict, seeds, tol: float, max_iter: int = 1000, damp: float = 0.85):
    n = csr["num_vertices"]
    od, idg, iidx, idst = _inv(csr)
    sd = _u(seeds)
    x = np.zeros(n, np.float64)
    res = C.c_double()
    it = load().oracle_ppr_f64(n, _p(od), _p(idg), _p(iidx), _p(idst), damp, _p(sd), sd.size,
                               tol, max_iter, _p(x, C.c_double), C.byref(res))
    return x, int(it), res.value


def csr_build_omp(src, dst, n: int, threads: int = 0):
    """OpenMP CSR build (CPU baseline).  Returns (seconds, csr dict)."""
    s, d = _u(src), _u(dst)
    m = s.size
    keys = ["sorted_edges_array.edges_array_src", "sorted_edges_array.edges_array_dest",
            "vertices.out_degree", "vertices.edges_idx",
            "inverse_sorted_edges_array.edges_array_src",
            "inverse_sorted_edges_array.edges_array_dest", "inverse_vertices.out_degree",
            "inverse_vertices.edges_idx"]
    sizes = [m, m, n, n, m, m, n, n]
    out = {k: np.zeros(z, np.uint32) for k, z in zip(keys, sizes)}
    secs = load().oracle_csr_build_omp(n, m, _p(s), _p(d), *[_p(out[k]) for k in keys], threads)
    out["vertices.in_degree"] = out["inverse_vertices.out_degree"]
    out["inverse_vertices.in_degree"] = out["vertices.out_degree"]
    out["num_vertices"] = n
    out["num_edges"] = m
    return secs, out


def pagerank_pull_omp(csr: dict, iterations: int = 20, epsilon: float = 0.0, threads: int = 0):
    """OpenMP pull PageRank (CPU baseline).  Returns (loop seconds, ranks, iterations, err)."""
    n = csr["num_vertices"]
    od, idg, iidx, idst = _inv(csr)
    pr = np.zeros(n, np.float32)
    et = C.c_double()
    it = C.c_uint32()
    secs = load().oracle_pagerank_pull_omp(n, _p(od), _p(idg), _p(iidx), _p(idst), iterations,
                                           epsilon, _p(pr, C.c_float), C.byref(et), C.byref(it),
                                           threads)
    return secs, pr, it.value, et.value


def omp_max_threads() -> int:
    return int(load().oracle_omp_max_threads())
