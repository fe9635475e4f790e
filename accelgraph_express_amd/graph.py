"""Python host-side mirror of the CSR builder / pull-PageRank path.

Thin wrappers over the C ABI (``_lib``) with numpy arrays at the boundary.  The
names follow the reference's C API where one exists:

* ``Graph.build``          GraphCSR builder: edge list -> stable sort by src ->
                           vertex mapping, plus the inverse CSR
                           (SSSP.c:485-488,518-520,548-563 call sequence)
* ``Graph.download``       the GraphCSR arrays (``vertices->out_degree`` ...,
                           ``sorted_edges_array->edges_array_src`` ...)
* ``Graph.pagerank_pull``  pageRankPullGraphCSR (pageRank.c:942-1108) and the
                           PageRankStats fields it fills (pageRank.h:19-30);
                           mode="data_driven" is pageRankDataDrivenPullGraphCSR
                           (pageRank.c:2810-2968)
* ``Graph.spmv_pull``      SPMVPullGraphCSR (SPMV.c:622-750), unweighted
* ``reorder_labels`` / ``relabel_edges``  the -l/-L/-O locality relabelling
                           (README.md:362-411)
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def _ptr(a: np.ndarray, ct=C.c_uint32):
    return a.ctypes.data_as(C.POINTER(ct))


class Context:
    """One HIP device + one stream (``ag_ctx``)."""

    def __init__(self, device: int = -1):
        lib = L.load()
        h = C.c_void_p()
        L.check(lib.ag_ctx_create(device, C.byref(h)))
        self.h = h

    def synchronize(self) -> None:
        L.check(L.load().ag_ctx_synchronize(self.h))

    @property
    def stream(self) -> int:
        return L.load().ag_ctx_stream(self.h) or 0

    def close(self) -> None:
        if getattr(self, "h", None):
            L.load().ag_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class PageRankStats:
    """Mirror of struct PageRankStats (pageRank.h:19-30) plus device timings."""
    damp: float
    base_pr: float
    iterations: int
    num_vertices: int
    activeVertices: int
    pageRanks: np.ndarray
    error_total: float
    pr_sum: float
    sweeps: int
    loop_ms: float
    total_ms: float
    plan_ms: float
    fp32_switch_sweep: int = 0
    active_per_iter: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    error_per_iter: np.ndarray = field(default_factory=lambda: np.zeros(0))

    @property
    def realRanks(self) -> np.ndarray:
        return np.arange(self.num_vertices, dtype=np.uint32)


_MODES = {"reference": L.AG_PR_REFERENCE, "standard": L.AG_PR_STANDARD,
          "data_driven": L.AG_PR_DATA_DRIVEN}
_PRECS = {"fp32": L.AG_PREC_FP32, "bf16": L.AG_PREC_BF16, "fp16": L.AG_PREC_FP16}
_SORTS = {"radix_src": 0, "radix_src_dst": 1, "count_src": 2, "count_src_dst": 3}


class Graph:
    """A device-resident graph (``ag_graph``)."""

    def __init__(self, ctx: Context, handle: C.c_void_p):
        self.ctx = ctx
        self.h = handle

    # ---- construction ----
    @classmethod
    def build(cls, ctx: Context, src, dst, num_vertices: int = 0, sort=0,
              inverse_from_input: bool = False, device_ptrs: tuple | None = None,
              num_edges: int | None = None) -> "Graph":
        """Builds forward + inverse CSR on the GPU.

        ``sort`` is the reference's sort mode (0 radix-src, 1 radix-src-dst,
        2 count-src, 3 count-src-dst) or its name.  ``device_ptrs`` =
        (src_ptr, dst_ptr) builds from device memory instead of host arrays.
        """
        lib = L.load()
        mode = _SORTS[sort] if isinstance(sort, str) else int(sort)
        flags = mode | (L.AG_BUILD_INVERSE_FROM_INPUT if inverse_from_input else 0)
        h = C.c_void_p()
        if device_ptrs is not None:
            flags |= L.AG_BUILD_DEVICE_INPUT
            L.check(lib.ag_csr_build(ctx.h, C.c_void_p(device_ptrs[0]), C.c_void_p(device_ptrs[1]),
                                     int(num_edges), num_vertices, flags, C.byref(h)))
        else:
            s, d = _u32(src), _u32(dst)
            if s.shape != d.shape:
                raise ValueError("src and dst differ in length")
            L.check(lib.ag_csr_build(ctx.h, s.ctypes.data_as(C.c_void_p),
                                     d.ctypes.data_as(C.c_void_p), s.size, num_vertices, flags,
                                     C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_csr(cls, ctx: Context, num_vertices: int, fwd_out_degree, inv_out_degree,
                 inv_edges_idx, inv_dest) -> "Graph":
        """Uploads an existing (host) CSR: what pull PageRank reads."""
        lib = L.load()
        a, b, c, d = (_u32(x) for x in (fwd_out_degree, inv_out_degree, inv_edges_idx, inv_dest))
        h = C.c_void_p()
        L.check(lib.ag_graph_from_csr(ctx.h, num_vertices, d.size, _ptr(a), _ptr(b), _ptr(c),
                                      _ptr(d), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def load(cls, ctx: Context, path: str) -> "Graph":
        """Loads an AGCSRv1 file (see ag_graph_save) onto the device."""
        h = C.c_void_p()
        L.check(L.load().ag_graph_load(ctx.h, os.fsencode(path), C.byref(h)))
        return cls(ctx, h)

    def save(self, path: str) -> None:
        """Writes the graph's CSR arrays to an AGCSRv1 file."""
        L.check(L.load().ag_graph_save(self.ctx.h, self.h, os.fsencode(path)))

    def free(self) -> None:
        if getattr(self, "h", None):
            L.load().ag_graph_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def num_vertices(self) -> int:
        return int(L.load().ag_graph_num_vertices(self.h))

    @property
    def num_edges(self) -> int:
        return int(L.load().ag_graph_num_edges(self.h))

    @property
    def device_bytes(self) -> int:
        return int(L.load().ag_graph_device_bytes(self.h))

    def view(self) -> L.CsrView:
        v = L.CsrView()
        L.check(L.load().ag_graph_view(self.h, C.byref(v)))
        return v

    def download(self) -> dict:
        """GraphCSR arrays on the host, keyed by the reference's field paths."""
        n, m = self.num_vertices, self.num_edges
        names = ["vertices.out_degree", "vertices.in_degree", "vertices.edges_idx",
                 "sorted_edges_array.edges_array_src", "sorted_edges_array.edges_array_dest",
                 "inverse_vertices.out_degree", "inverse_vertices.in_degree",
                 "inverse_vertices.edges_idx", "inverse_sorted_edges_array.edges_array_src",
                 "inverse_sorted_edges_array.edges_array_dest"]
        sizes = [n, n, n, m, m, n, n, n, m, m]
        out = {k: np.zeros(s, np.uint32) for k, s in zip(names, sizes)}
        L.check(L.load().ag_graph_download(self.ctx.h, self.h, *[_ptr(out[k]) for k in names]))
        return out

    # ---- PageRank ----
    def prepare(self) -> float:
        ms = C.c_double(0.0)
        L.check(L.load().ag_pagerank_prepare(self.ctx.h, self.h, C.byref(ms)))
        return ms.value

    def autotune(self, sweeps: int = 20) -> tuple:
        """Times candidate gather layouts of this graph on the device and keeps
        the fastest (ag_pagerank_autotune).  Returns (tuning dict, wall ms)."""
        t = L.Tuning()
        ms = C.c_double(0.0)
        L.check(L.load().ag_pagerank_autotune(self.ctx.h, self.h, sweeps, C.byref(t), C.byref(ms)))
        return {f: getattr(t, f) for f, _ in L.Tuning._fields_}, ms.value

    def plan_info(self) -> dict:
        info = L.PlanInfo()
        L.check(L.load().ag_pagerank_plan_info(self.h, C.byref(info)))
        return {k: getattr(info, k) for k, _ in L.PlanInfo._fields_}

    def pagerank_pull(self, iterations: int = 20, epsilon: float = 1e-4, mode="reference",
                      precision="fp32", finish_fp32: bool = False,
                      damp: float = 0.85, extrapolate: bool = False,
                      init=None) -> PageRankStats:
        """Pull PageRank on the resident graph.  `init` (N floats, the scale of
        the returned ranks, e.g. a previous call's pageRanks) replaces the
        uniform start (AG_PR_FLAG_WARM_START; not in the reference)."""
        lib = L.load()
        o = L.PrOptions()
        lib.ag_pr_options_default(C.byref(o))
        o.iterations = iterations
        o.epsilon = epsilon
        o.mode = _MODES[mode] if isinstance(mode, str) else int(mode)
        o.precision = _PRECS[precision] if isinstance(precision, str) else int(precision)
        o.flags = (L.AG_PR_FLAG_FINISH_FP32 if finish_fp32 else 0) | \
                  (L.AG_PR_FLAG_EXTRAPOLATE if extrapolate else 0)
        o.damp = damp
        n = self.num_vertices
        ranks = np.zeros(n, np.float32)
        if init is not None:
            x0 = np.ascontiguousarray(init, np.float32)
            if x0.size != n:
                raise ValueError(f"init has {x0.size} entries, graph has {n} vertices")
            ranks[:] = x0
            o.flags |= L.AG_PR_FLAG_WARM_START
        act = np.zeros(max(iterations, 1), np.uint32)
        err = np.zeros(max(iterations, 1), np.float64)
        r = L.PrResult()
        L.check(lib.ag_pagerank_pull(self.ctx.h, self.h, C.byref(o), _ptr(ranks, C.c_float),
                                     _ptr(act), _ptr(err, C.c_double), C.byref(r)))
        damp32 = np.float32(damp)
        return PageRankStats(
            damp=float(damp32), base_pr=float(np.float32(1.0) - damp32), iterations=r.iterations,
            num_vertices=n, activeVertices=r.active_vertices, pageRanks=ranks,
            error_total=r.error_total, pr_sum=r.pr_sum, sweeps=r.sweeps, loop_ms=r.loop_ms,
            total_ms=r.total_ms, plan_ms=r.plan_ms, fp32_switch_sweep=r.fp32_switch_sweep,
            active_per_iter=act[:r.sweeps].copy(), error_per_iter=err[:r.sweeps].copy())

    def spmv_pull(self, x, iterations: int = 1):
        """y[v] = sum of x[u] over the in-edges u -> v (FP32), run `iterations`
        times on the same input.  Returns (y, SpmvResult)."""
        n = self.num_vertices
        xs = np.ascontiguousarray(x, np.float32)
        if xs.size != n:
            raise ValueError(f"x has {xs.size} entries, graph has {n} vertices")
        y = np.zeros(n, np.float32)
        r = L.SpmvResult()
        L.check(L.load().ag_spmv_pull(self.ctx.h, self.h, _ptr(xs, C.c_float), _ptr(y, C.c_float),
                                      iterations, C.byref(r)))
        return y, r

    def spmv_pull_device(self, d_x_ptr: int, d_y_ptr: int, iterations: int = 1) -> L.SpmvResult:
        """Same on device buffers of N floats (not aliased)."""
        r = L.SpmvResult()
        L.check(L.load().ag_spmv_pull_device(self.ctx.h, self.h, C.c_void_p(d_x_ptr),
                                             C.c_void_p(d_y_ptr), iterations, C.byref(r)))
        return r

    def pagerank_pull_device(self, d_ranks_ptr: int, iterations: int = 20, epsilon: float = 0.0,
                             mode="reference", precision="fp32", finish_fp32=False,
                             time_sweeps: bool = False, extrapolate: bool = False,
                             warm_start: bool = False) -> L.PrResult:
        """Ranks to a device buffer of N floats; time_sweeps fills
        sweep_kernel_ms (HIP events around each sweep kernel); extrapolate
        (STANDARD mode) takes Aitken-type extrapolated steps; warm_start reads
        the start from the same buffer first."""
        lib = L.load()
        o = L.PrOptions()
        lib.ag_pr_options_default(C.byref(o))
        o.iterations, o.epsilon = iterations, epsilon
        o.mode = _MODES[mode] if isinstance(mode, str) else int(mode)
        o.precision = _PRECS[precision] if isinstance(precision, str) else int(precision)
        o.flags = (L.AG_PR_FLAG_FINISH_FP32 if finish_fp32 else 0) | \
                  (L.AG_PR_FLAG_TIME_SWEEPS if time_sweeps else 0) | \
                  (L.AG_PR_FLAG_EXTRAPOLATE if extrapolate else 0) | \
                  (L.AG_PR_FLAG_WARM_START if warm_start else 0)
        r = L.PrResult()
        L.check(lib.ag_pagerank_pull_device(self.ctx.h, self.h, C.byref(o),
                                            C.c_void_p(d_ranks_ptr), C.byref(r)))
        return r

    def pagerank_pull_device_noalloc(self, d_ranks_ptr: int, iterations: int = 20,
                                     epsilon: float = 0.0, mode="reference", precision="fp32",
                                     finish_fp32=False, warm_start: bool = False) -> L.PrResult:
        """ag_pagerank_pull_device_noalloc: the same loop with no device
        allocation (per-call buffers in the library's static 16 MiB scratch;
        the gather layout must exist, see prepare())."""
        lib = L.load()
        o = L.PrOptions()
        lib.ag_pr_options_default(C.byref(o))
        o.iterations, o.epsilon = iterations, epsilon
        o.mode = _MODES[mode] if isinstance(mode, str) else int(mode)
        o.precision = _PRECS[precision] if isinstance(precision, str) else int(precision)
        o.flags = (L.AG_PR_FLAG_FINISH_FP32 if finish_fp32 else 0) | \
                  (L.AG_PR_FLAG_WARM_START if warm_start else 0)
        r = L.PrResult()
        L.check(lib.ag_pagerank_pull_device_noalloc(self.ctx.h, self.h, C.byref(o),
                                                    C.c_void_p(d_ranks_ptr), C.byref(r)))
        return r


REORDER_MODES = {"none": 0, "out_degree": 1, "in_degree": 2, "inout_degree": 3, "dbg_out": 4,
                 "dbg_in": 5, "hubsort_out": 6, "hubsort_in": 7, "hubcluster_out": 8,
                 "hubcluster_in": 9, "random": 10}


def reorder_labels(ctx: Context, src, dst, num_vertices: int, mode, seed: int = 0) -> np.ndarray:
    """labels[v] = new label of vertex v under a locality relabelling mode
    (the reference's -l/-L/-O options 0-10), computed on the GPU."""
    s, d = _u32(src), _u32(dst)
    m = REORDER_MODES[mode] if isinstance(mode, str) else int(mode)
    lab = np.zeros(max(num_vertices, 1), np.uint32)
    L.check(L.load().ag_reorder_labels(ctx.h, _ptr(s), _ptr(d), s.size, num_vertices, m, seed,
                                       _ptr(lab)))
    return lab[:num_vertices]


def relabel_edges(ctx: Context, labels, src, dst):
    """(labels[src], labels[dst]) on the GPU; labels must be a permutation."""
    lab = _u32(labels)
    s, d = _u32(src).copy(), _u32(dst).copy()
    L.check(L.load().ag_relabel_edges(ctx.h, _ptr(lab), lab.size, _ptr(s), _ptr(d), s.size))
    return s, d


def spmv_input(out_degree) -> np.ndarray:
    """The reference's SPMV benchmark input (SPMV.c:699-706): 1/out_degree in
    FP32, 0.001 for a vertex without out-edges."""
    od = np.asarray(out_degree, np.uint32)
    x = np.full(od.size, np.float32(0.001), np.float32)
    nz = od > 0
    x[nz] = np.float32(1.0) / od[nz].astype(np.float32)
    return x


def csr_file_info(path: str) -> dict:
    """Header of an AGCSRv1 file after its size and checksum are verified
    (host only)."""
    n, e, m = C.c_uint32(), C.c_uint32(), C.c_uint32()
    L.check(L.load().ag_csr_file_info(os.fsencode(path), C.byref(n), C.byref(e), C.byref(m)))
    return {"num_vertices": n.value, "num_edges": e.value, "sections": m.value}


LAYOUTS = {"tables": L.AG_LAYOUT_TABLES, "rows": L.AG_LAYOUT_ROWS, "auto": L.AG_LAYOUT_AUTO}


def set_tuning(low_max: int = 0, hub_min: int = 0, hub_chunk: int = 0,
               hot_bytes: int = 0, layout=0, max_tables: int = 0, lane_max: int = 0) -> None:
    """0 keeps a knob; hot_bytes=1 disables the LDS table of hot contributions;
    hub_chunk=0xFFFFFFFF (AG_TUNE_AUTO, the default) picks it per graph;
    layout "auto" (default: tables when the sources outnumber one LDS table),
    "tables" or "rows".  Applies to gather layouts built later."""
    lay = LAYOUTS[layout] if isinstance(layout, str) else int(layout)
    t = L.Tuning(low_max, hub_min, hub_chunk, hot_bytes, lay, max_tables, lane_max)
    L.check(L.load().ag_set_tuning(C.byref(t)))


def get_tuning() -> dict:
    t = L.Tuning()
    L.check(L.load().ag_get_tuning(C.byref(t)))
    lay = {v: k for k, v in LAYOUTS.items()}.get(t.layout, t.layout)
    return {"low_max": t.low_max, "hub_min": t.hub_min, "hub_chunk": t.hub_chunk,
            "hot_bytes": t.hot_bytes, "layout": lay, "max_tables": t.max_tables,
            "lane_max": t.lane_max}


# ---- synthetic inputs ----
def gen_uniform(seed: int, num_vertices: int, num_edges: int):
    src = np.empty(num_edges, np.uint32)
    dst = np.empty(num_edges, np.uint32)
    L.check(L.load().ag_gen_uniform(seed, num_vertices, num_edges, _ptr(src), _ptr(dst)))
    return src, dst


def gen_rmat(seed: int, scale: int, edge_factor: int = 16, a=0.57, b=0.19, c=0.19,
             permute: bool = True, threads: int = 0):
    m = (1 << scale) * edge_factor
    src = np.empty(m, np.uint32)
    dst = np.empty(m, np.uint32)
    L.check(L.load().ag_gen_rmat(seed, scale, edge_factor, a, b, c, 1 if permute else 0,
                                 _ptr(src), _ptr(dst), threads))
    return src, dst


# ---- personalized PageRank ----
def _query(seeds, tol, max_iter):
    s = _u32(seeds)
    q = L.PprQuery(_ptr(s), s.size, max_iter, tol)
    return q, s


def ppr_single(graph: Graph, seeds, tol: float = 1e-4, max_iter: int = 1000,
               precision="fp32", return_x: bool = True, extrapolate: bool = False,
               damp: float = 0.85):
    """One query with the single-vector sweep kernel.  Returns (x or None, stat).
    extrapolate: Aitken-type extrapolated steps once the residual falls by a
    steady ratio (ag_ppr_options.extrapolate)."""
    lib = L.load()
    q, keep = _query(seeds, tol, max_iter)
    x = np.zeros(graph.num_vertices, np.float32) if return_x else None
    st = L.PprStat()
    o = L.PprOptions()
    lib.ag_ppr_options_default(C.byref(o))
    o.precision = _PRECS[precision] if isinstance(precision, str) else int(precision)
    o.extrapolate = 1 if extrapolate else 0
    o.damp = damp
    L.check(lib.ag_ppr_single_opts(graph.ctx.h, graph.h, C.byref(q), C.byref(o),
                                   _ptr(x, C.c_float) if return_x else None, C.byref(st)))
    del keep
    return x, st


def ppr_batch(graph: Graph, queries, precision="bf16", cols_per_wave: int = 0,
              replay: bool = True, damp: float = 0.85, return_x: bool = True,
              extrapolate: bool = False):
    """queries: list of (seeds, tol, max_iter).  Returns (X [nq, N] or None, stats,
    elapsed_ms); with return_x=False the results stay on the device.
    extrapolate: per-column Aitken-type extrapolated steps."""
    lib = L.load()
    nq = len(queries)
    arr = (L.PprQuery * nq)()
    keep = []
    for i, (seeds, tol, mi) in enumerate(queries):
        q, s = _query(seeds, tol, mi)
        arr[i] = q
        keep.append(s)
    o = L.PprOptions()
    lib.ag_ppr_options_default(C.byref(o))
    o.precision = _PRECS[precision] if isinstance(precision, str) else int(precision)
    o.cols_per_wave = cols_per_wave
    o.use_graph_replay = 1 if replay else 0
    o.extrapolate = 1 if extrapolate else 0
    o.damp = damp
    X = np.zeros((nq, graph.num_vertices), np.float32) if return_x else None
    stats = (L.PprStat * nq)()
    ms = C.c_double()
    L.check(lib.ag_ppr_batch(graph.ctx.h, graph.h, arr, nq, C.byref(o),
                             _ptr(X, C.c_float) if return_x else None, stats, C.byref(ms)))
    return X, list(stats), ms.value


def ppr_batch_bytes(graph: Graph, capacity: int, precision="bf16") -> int:
    """Device bytes a batch of `capacity` columns holds on the graph (the
    serving loop's per-batch reservation)."""
    b = C.c_uint64()
    prec = _PRECS[precision] if isinstance(precision, str) else int(precision)
    L.check(L.load().ag_ppr_batch_bytes(graph.ctx.h, graph.h, capacity, prec, C.byref(b)))
    return b.value


def ppr_autotune(graph: Graph, precision="bf16") -> int:
    c = C.c_uint32()
    prec = _PRECS[precision] if isinstance(precision, str) else int(precision)
    L.check(L.load().ag_ppr_autotune(graph.ctx.h, graph.h, prec, C.byref(c)))
    return c.value


class Server:
    """Continuous-batching PPR server over resident graphs (``ag_server``)."""

    def __init__(self, ctx: Context, hbm_budget_bytes: int = 0, max_columns: int = 0,
                 precision="bf16", cols_per_wave: int = 0, inject_fail_alloc: int = 0,
                 inject_fail_query: int = 0, extrapolate: bool = False):
        lib = L.load()
        cfg = L.ServeConfig()
        lib.ag_serve_config_default(C.byref(cfg))
        if hbm_budget_bytes:
            cfg.hbm_budget_bytes = hbm_budget_bytes
        if max_columns:
            cfg.max_columns = max_columns
        cfg.precision = _PRECS[precision] if isinstance(precision, str) else int(precision)
        cfg.cols_per_wave = cols_per_wave
        cfg.inject_fail_alloc = inject_fail_alloc
        cfg.inject_fail_query = inject_fail_query
        cfg.extrapolate = 1 if extrapolate else 0
        h = C.c_void_p()
        L.check(lib.ag_serve_create(ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h
        self.ctx = ctx
        self.graphs = []

    def add_graph(self, g: Graph) -> int:
        gid = C.c_uint32()
        L.check(L.load().ag_serve_add_graph(self.h, g.h, C.byref(gid)))
        self.graphs.append(g)
        return gid.value

    def submit(self, gid: int, seeds, tol: float = 1e-4, max_iter: int = 1000) -> int:
        s = _u32(seeds)
        t = C.c_uint64()
        L.check(L.load().ag_serve_submit(self.h, gid, _ptr(s), s.size, tol, max_iter,
                                         C.byref(t)))
        return t.value

    def cancel(self, ticket: int) -> None:
        L.check(L.load().ag_serve_cancel(self.h, ticket))

    def step(self) -> int:
        n = C.c_uint32()
        L.check(L.load().ag_serve_step(self.h, C.byref(n)))
        return n.value

    def poll(self, ticket: int, n: int):
        """Returns (code, x or None, stat).  code AG_ERR_PENDING while running
        (x untouched: np.empty maps no pages until the result is copied in)."""
        x = np.empty(n, np.float32)
        st = L.PprStat()
        code = L.load().ag_serve_poll(self.h, ticket, _ptr(x, C.c_float), n, C.byref(st))
        return code, (x if code == L.AG_OK else None), st

    def peek(self, ticket: int):
        """Zero-copy result: (code, x, stat) where x is a numpy view of the
        pinned host buffer the GPU wrote (valid until release(ticket))."""
        p = L.f32p()
        n = C.c_uint32()
        st = L.PprStat()
        code = L.load().ag_serve_peek(self.h, ticket, C.byref(p), C.byref(n), C.byref(st))
        x = np.ctypeslib.as_array(p, shape=(n.value,)) if code == L.AG_OK and p else None
        return code, x, st

    def release(self, ticket: int) -> None:
        L.check(L.load().ag_serve_release(self.h, ticket))

    @property
    def reserved_bytes(self) -> int:
        return int(L.load().ag_serve_reserved_bytes(self.h))

    def close(self) -> None:
        if getattr(self, "h", None):
            L.load().ag_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
