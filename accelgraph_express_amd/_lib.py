"""ctypes binding of libaccelgraph_amd.so (the C ABI in include/accelgraph_amd.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (make in
``accelgraph_express_amd/csrc``).  There is no fallback: importing this module
without the built library raises, and every compute call on a machine without
a HIP device returns AG_ERR_NO_DEVICE, which is raised as ``AGError``.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaccelgraph_amd.so")

AG_OK = 0
AG_ERR_INVALID = 1
AG_ERR_HIP = 2
AG_ERR_OOM = 3
AG_ERR_ADMISSION = 4
AG_ERR_NOT_FOUND = 5
AG_ERR_CANCELLED = 6
AG_ERR_IO = 7
AG_ERR_INJECTED = 8
AG_ERR_PENDING = 9
AG_ERR_NO_DEVICE = 10

AG_BUILD_INVERSE_FROM_INPUT = 0x100
AG_BUILD_NO_INVERSE = 0x200
AG_BUILD_DEVICE_INPUT = 0x400

AG_PR_REFERENCE = 0
AG_PR_STANDARD = 1
AG_PR_DATA_DRIVEN = 2
AG_PREC_FP32 = 0
AG_PREC_BF16 = 1
AG_PREC_FP16 = 2
AG_PR_FLAG_FINISH_FP32 = 0x1
AG_PR_FLAG_TIME_SWEEPS = 0x2
AG_PR_FLAG_EXTRAPOLATE = 0x4
AG_PR_FLAG_WARM_START = 0x8
AG_LAYOUT_TABLES = 1
AG_LAYOUT_ROWS = 2
AG_LAYOUT_AUTO = 3

u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
vp = C.c_void_p


class AGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[AG error {code}] {msg}")
        self.code = code


class PrOptions(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("mode", C.c_uint32), ("precision", C.c_uint32),
                ("flags", C.c_uint32), ("epsilon", C.c_double), ("damp", C.c_float),
                ("reserved", C.c_float)]


class SpmvResult(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("reserved", C.c_uint32), ("y_sum", C.c_double),
                ("loop_ms", C.c_double), ("total_ms", C.c_double), ("plan_ms", C.c_double)]


class PrResult(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("sweeps", C.c_uint32),
                ("active_vertices", C.c_uint32), ("fp32_switch_sweep", C.c_uint32),
                ("error_total", C.c_double), ("pr_sum", C.c_double), ("loop_ms", C.c_double),
                ("total_ms", C.c_double), ("plan_ms", C.c_double),
                ("sweep_kernel_ms", C.c_double)]


class PlanInfo(C.Structure):
    _fields_ = [("n_hub", C.c_uint32), ("n_mid", C.c_uint32), ("n_low", C.c_uint32),
                ("n_zero", C.c_uint32), ("hub_chunks", C.c_uint32), ("low_max", C.c_uint32),
                ("hub_min", C.c_uint32), ("sell_slots", C.c_uint64), ("plan_bytes", C.c_uint64),
                ("hub_chunk", C.c_uint32), ("layout", C.c_uint32), ("tables", C.c_uint32),
                ("table_workgroups", C.c_uint32), ("subrows", C.c_uint64),
                ("cold_edges", C.c_uint64), ("lane_max", C.c_uint32), ("max_tables", C.c_uint32),
                ("autotuned", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("low_max", C.c_uint32), ("hub_min", C.c_uint32), ("hub_chunk", C.c_uint32),
                ("hot_bytes", C.c_uint32), ("layout", C.c_uint32), ("max_tables", C.c_uint32),
                ("lane_max", C.c_uint32)]


class PprQuery(C.Structure):
    _fields_ = [("seeds", u32p), ("nseeds", C.c_uint32), ("max_iter", C.c_uint32),
                ("tol", C.c_double)]


class PprStat(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("status", C.c_uint32), ("residual", C.c_double),
                ("latency_ms", C.c_double), ("fine_sweep", C.c_uint32), ("reserved", C.c_uint32)]


class PprOptions(C.Structure):
    _fields_ = [("precision", C.c_uint32), ("cols_per_wave", C.c_uint32), ("damp", C.c_float),
                ("use_graph_replay", C.c_uint32), ("extrapolate", C.c_uint32)]


class ServeConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_columns", C.c_uint32),
                ("precision", C.c_uint32), ("cols_per_wave", C.c_uint32),
                ("inject_fail_alloc", C.c_uint32), ("inject_fail_query", C.c_uint32),
                ("damp", C.c_float), ("extrapolate", C.c_uint32)]


class CsrView(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_edges", C.c_uint32)] + [
        (n, vp) for n in ("out_degree", "in_degree", "edges_idx", "sorted_src", "sorted_dest",
                          "inv_out_degree", "inv_in_degree", "inv_edges_idx", "inv_sorted_src",
                          "inv_sorted_dest", "offsets", "inv_offsets")]


# name -> (restype, argtypes)
_SIGS = {
    "ag_last_error": (C.c_char_p, []),
    "ag_version": (C.c_char_p, []),
    "ag_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "ag_ctx_destroy": (C.c_int, [vp]),
    "ag_ctx_synchronize": (C.c_int, [vp]),
    "ag_ctx_stream": (vp, [vp]),
    "ag_default_ctx": (C.c_int, [C.POINTER(vp)]),
    "ag_csr_build": (C.c_int, [vp, vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(vp)]),
    "ag_graph_from_csr": (C.c_int, [vp, C.c_uint32, C.c_uint32, u32p, u32p, u32p, u32p,
                                    C.POINTER(vp)]),
    "ag_graph_free": (C.c_int, [vp]),
    "ag_graph_num_vertices": (C.c_uint32, [vp]),
    "ag_graph_num_edges": (C.c_uint32, [vp]),
    "ag_graph_device_bytes": (C.c_uint64, [vp]),
    "ag_graph_view": (C.c_int, [vp, C.POINTER(CsrView)]),
    "ag_graph_download": (C.c_int, [vp, vp] + [u32p] * 10),
    "ag_graph_save": (C.c_int, [vp, vp, C.c_char_p]),
    "ag_graph_load": (C.c_int, [vp, C.c_char_p, C.POINTER(vp)]),
    "ag_csr_file_info": (C.c_int, [C.c_char_p, u32p, u32p, u32p]),
    "ag_sort_pairs_host": (C.c_int, [vp, u32p, u32p, C.c_uint32, C.c_uint32, u32p, u32p]),
    "ag_vertex_map_host": (C.c_int, [vp, u32p, C.c_uint32, C.c_uint32, u32p, u32p]),
    "ag_pr_options_default": (None, [C.POINTER(PrOptions)]),
    "ag_pagerank_pull": (C.c_int, [vp, vp, C.POINTER(PrOptions), f32p, u32p, f64p,
                                   C.POINTER(PrResult)]),
    "ag_pagerank_pull_device": (C.c_int, [vp, vp, C.POINTER(PrOptions), vp,
                                          C.POINTER(PrResult)]),
    "ag_pagerank_pull_device_noalloc": (C.c_int, [vp, vp, C.POINTER(PrOptions), vp,
                                                  C.POINTER(PrResult)]),
    "ag_pagerank_prepare": (C.c_int, [vp, vp, f64p]),
    "ag_spmv_pull": (C.c_int, [vp, vp, f32p, f32p, C.c_uint32, C.POINTER(SpmvResult)]),
    "ag_reorder_labels": (C.c_int, [vp, u32p, u32p, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_uint32, u32p]),
    "ag_reorder_labels_device": (C.c_int, [vp, vp, vp, C.c_uint32, C.c_uint32, C.c_uint32,
                                           C.c_uint32, vp]),
    "ag_relabel_edges": (C.c_int, [vp, u32p, C.c_uint32, u32p, u32p, C.c_uint32]),
    "ag_relabel_edges_device": (C.c_int, [vp, vp, C.c_uint32, vp, vp, C.c_uint32]),
    "ag_spmv_pull_device": (C.c_int, [vp, vp, vp, vp, C.c_uint32, C.POINTER(SpmvResult)]),
    "ag_pagerank_plan_info": (C.c_int, [vp, C.POINTER(PlanInfo)]),
    "ag_pagerank_autotune": (C.c_int, [vp, vp, C.c_uint32, C.POINTER(Tuning), f64p]),
    "ag_set_tuning": (C.c_int, [C.POINTER(Tuning)]),
    "ag_get_tuning": (C.c_int, [C.POINTER(Tuning)]),
    "ag_ppr_options_default": (None, [C.POINTER(PprOptions)]),
    "ag_ppr_batch": (C.c_int, [vp, vp, C.POINTER(PprQuery), C.c_uint32, C.POINTER(PprOptions),
                               f32p, C.POINTER(PprStat), f64p]),
    "ag_ppr_batch_bytes": (C.c_int, [vp, vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    "ag_ppr_autotune": (C.c_int, [vp, vp, C.c_uint32, u32p]),
    "ag_ppr_single": (C.c_int, [vp, vp, C.POINTER(PprQuery), C.c_uint32, f32p,
                                C.POINTER(PprStat)]),
    "ag_ppr_single_opts": (C.c_int, [vp, vp, C.POINTER(PprQuery), C.POINTER(PprOptions), f32p,
                                     C.POINTER(PprStat)]),
    "ag_serve_config_default": (None, [C.POINTER(ServeConfig)]),
    "ag_serve_create": (C.c_int, [vp, C.POINTER(ServeConfig), C.POINTER(vp)]),
    "ag_serve_destroy": (C.c_int, [vp]),
    "ag_serve_add_graph": (C.c_int, [vp, vp, u32p]),
    "ag_serve_submit": (C.c_int, [vp, C.c_uint32, u32p, C.c_uint32, C.c_double, C.c_uint32,
                                  C.POINTER(C.c_uint64)]),
    "ag_serve_cancel": (C.c_int, [vp, C.c_uint64]),
    "ag_serve_step": (C.c_int, [vp, u32p]),
    "ag_serve_poll": (C.c_int, [vp, C.c_uint64, f32p, C.c_uint32, C.POINTER(PprStat)]),
    "ag_serve_peek": (C.c_int, [vp, C.c_uint64, C.POINTER(f32p), u32p, C.POINTER(PprStat)]),
    "ag_serve_release": (C.c_int, [vp, C.c_uint64]),
    "ag_serve_reserved_bytes": (C.c_uint64, [vp]),
    "ag_gen_uniform": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, u32p, u32p]),
    "ag_gen_rmat": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                              C.c_double, C.c_int, u32p, u32p, C.c_int]),
}

_lib = None


def load() -> C.CDLL:
    """Loads the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
            "g.build()'` (there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != AG_OK:
        raise AGError(code, load().ag_last_error().decode(errors="replace"))
