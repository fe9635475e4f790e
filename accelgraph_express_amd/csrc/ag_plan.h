// Gather layout ("plan") shared by the PageRank and personalized-PageRank
// kernels.  Internal; not part of the C ABI.
#pragma once

#include <mutex>
#include <vector>

#include "ag_internal.h"

namespace ag {

// BF16 storage word (round-to-nearest-even from FP32, widened by a shift).
struct bf16_t {
    uint16_t x;
};

struct Work {
    uint32_t max_iter = 0;
    uint32_t grid = 0;
    DevBuf pr, c0, c1, blk_err, blk_act, blk_dang, state, err_iter, act_iter, fmt, sum_part, out;
};

// A captured run of the PageRank sweep loop (init, sweeps, finish), replayed
// while its launch arguments (`key`) stay the same.
struct LoopGraph {
    std::vector<unsigned char> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

// Buffers of a personalized-PageRank batch that a finished batch leaves to
// the next one on the same graph (see PprBatch).
constexpr int kKeepBufs = 11;

struct Plan {
    uint32_t N = 0, E = 0;
    uint32_t n_hub = 0, n_w = 0, n_h = 0, n_nonlow = 0, n_nonzero = 0, n_dangling = 0;
    uint32_t low_slices = 0, hub_chunks = 0;
    uint32_t low_max = 0, hub_min = 0, hub_chunk = 0, hot_bytes = 0;
    uint64_t csr_edges = 0, sell_slots = 0;
    // order[i] = vertex with label i; rank[v] = label of vertex v
    DevBuf order, rank, roff, cols, slice_ptr, outdeg, chunk_row, chunk_first, hub_partial;
    uint32_t nb_hub = 0, nb_w = 0, nb_h = 0, nb_q = 0, nb_low = 0, nb_zero = 0, nvblocks = 0;
    Work work;
    // single-query personalized PageRank scratch (reused across calls)
    DevBuf ppr_bits, ppr_ids, ppr_labels, ppr_w, ppr_out;
    // batched personalized PageRank: fixed-size edge chunks of the rows
    // [0, n_nonlow), built on first use
    DevBuf ppr_ch_row, ppr_ch_first, ppr_multi;
    uint32_t ppr_nch = 0, ppr_nmulti = 0;
    bool ppr_ready = false;
    std::vector<uint32_t> rank_host; // host copy of rank (seed ids -> labels), on first use
    // buffers of the last finished PprBatch (columns, state, partials,
    // staging), lent to the next batch on this graph
    DevBuf ppr_keep[kKeepBufs];
    // captured sweep loops, most recent last (at most kLoopGraphs)
    static constexpr size_t kLoopGraphs = 8;
    std::vector<LoopGraph> loop_graphs;
    std::mutex mu;
    Plan() = default;
    Plan(const Plan &) = delete;
    Plan &operator=(const Plan &) = delete;
    ~Plan()
    {
        for (LoopGraph &lg : loop_graphs) {
            if (lg.exec) (void)hipGraphExecDestroy(lg.exec);
            if (lg.graph) (void)hipGraphDestroy(lg.graph);
        }
    }
    uint64_t bytes() const
    {
        return order.bytes + rank.bytes + roff.bytes + cols.bytes + slice_ptr.bytes + outdeg.bytes +
               chunk_row.bytes + chunk_first.bytes + hub_partial.bytes + ppr_ch_row.bytes +
               ppr_ch_first.bytes + ppr_multi.bytes + keep_bytes();
    }
    uint64_t keep_bytes() const
    {
        uint64_t b = 0;
        for (const DevBuf &k : ppr_keep) b += k.bytes;
        return b;
    }
};


// Internal PageRank mode: single-query personalized PageRank (teleport and
// dangling mass to the seed set), used by ag_ppr_single.  Outside the public
// AG_PR_* range.
constexpr uint32_t kModePPR = 0x100;
// Internal sweep mode for SPMV pull: the row's gathered sum is the output
// (no damping, no contribution written); the input sits in label space in c0.
constexpr uint32_t kModeSpmv = 0x101;

// Seed table of one personalized query, in label space (device arrays).
struct SeedTable {
    const uint32_t *bits = nullptr; // N-bit membership bitmap
    const uint32_t *v = nullptr;    // labels, unique
    const float *w = nullptr;       // teleport weights
    uint32_t n = 0;
};

Plan *ensure_plan(ag_ctx *ctx, ag_graph *g, double *ms);
void run_pagerank_seeded(ag_ctx *ctx, ag_graph *g, const ag_pr_options &o, const SeedTable &st,
                         float *d_out, ag_pr_result *res);

} // namespace ag
