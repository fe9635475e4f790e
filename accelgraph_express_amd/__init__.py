"""MI355X-native CSR builder + pull PageRank (the hot path of AccelGraph_Express).

The compute lives in ``libaccelgraph_amd.so`` (hand-written HIP kernels for
gfx950 behind the C ABI in ``include/accelgraph_amd.h``).  This package is the
Python host-side mirror used by tests and ``bench.py``.
"""
from . import _lib
from ._lib import AGError, load
from .graph import (Context, Graph, PageRankStats, Server, csr_file_info, gen_rmat, gen_uniform,
                    get_tuning, ppr_autotune, ppr_batch, ppr_batch_bytes, ppr_single,
                    relabel_edges,
                    reorder_labels, set_tuning, spmv_input)

__all__ = ["AGError", "Context", "Graph", "PageRankStats", "Server", "csr_file_info", "gen_rmat",
           "gen_uniform",
           "get_tuning", "ppr_autotune", "ppr_batch", "ppr_batch_bytes", "ppr_single", "relabel_edges",
           "reorder_labels", "set_tuning", "spmv_input",
           "load", "_lib"]
