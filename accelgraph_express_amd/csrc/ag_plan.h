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

// Partitioned-table gather layout, the pull sweep's default (ag_pagerank.hip).
// Source labels are cut into tables of T labels; the sweep's workgroups each
// hold one table's contributions in LDS and sum, for every row, the row's
// in-edges from that table (a "sub-row"), so that nearly every gather is an
// LDS read.  Sources past the last LDS table form the cold table, gathered
// from global memory.  A combine kernel then sums each row's sub-row partials
// in table order and updates the vertex; hub rows' sub-rows are cut into
// chunks that the finalize kernel sums.
//
// Sub-rows are numbered table by table, rows ascending inside a table (the
// hub rows' sub-rows first).  A sub-row's source indices are contiguous and
// padded with a sentinel to a multiple of 4: u16 table-local indices for LDS
// tables (sentinel T reads the zero after the table), u32 labels for the cold
// table (sentinel N reads the zero kept after the contribution vector).
struct TabPlan {
    uint32_t T = 0;          // sources per LDS table
    uint32_t nt = 0;         // LDS tables; table nt is the cold table
    uint32_t NT = 0;         // nt + 1
    uint32_t S = 0;          // sub-rows
    uint32_t cold_base4 = 0; // sub_start value at which the cold table's indices begin
    uint32_t nunits = 0;     // work units of the table kernel
    uint32_t cold_unit0 = 0; // units [cold_unit0, nunits) belong to the cold table
    uint32_t cold_lane_u0 = 0, cold_lane_s0 = 0, cold_lane_s1 = 0; // cold lane units: first unit,
                                                            // its first sub-row, the cold table's end
    uint32_t grid = 0;       // workgroups of the table kernel (one per CU)
    uint32_t hub_chunk = 0;  // edges per hub-row chunk (multiple of 4)
    uint32_t hub_chunks = 0;
    uint32_t nwaves = 0;     // 64-row waves of the combine kernel over the non-hub rows
    uint32_t NTpad = 0;      // tables per combine record (NT rounded up to 4, 8, 12, 17, 24 or 34)
    uint32_t grp = 0;        // tables per group of a record (the combine kernel loads one at a time)
    uint32_t row_blocks = 0; // combine blocks over those waves
    uint32_t hub_blocks = 0; // combine blocks over the hub rows (one wave per row)
    uint64_t idx16_len = 0, idx32_len = 0;
    std::vector<uint32_t> SB;         // first sub-row of each table, SB[NT] = S (host copy)
    std::vector<uint64_t> tab_edges;  // edges per table (host copy)
    // idx16 / idx32: source indices; sub_start[S + 1]: start of each sub-row's
    // indices in units of 4 entries; part[S]: sub-row partial sums; units,
    // wg_begin: the table kernel's work units and each workgroup's range;
    // items: sub-rows sorted by (table, class), the long ones' lists;
    // hub_items: (sub-row, chunk, output slot) per hub chunk; chunk_first,
    // hub_partial: hub rows' chunk slots, in (table, chunk) order per row;
    // wtab: per 64-row wave of the non-hub rows, 3 * NTpad words in groups of
    // grp tables: for each table of the group the lanes (rows) with a
    // sub-row in it (two words: lanes 0-31, 32-63), then for each table of the
    // group the first sub-row of the table at or after the wave's first row.
    // Tables NT .. NTpad - 1 have no lanes.
    DevBuf idx16, idx32, sub_start, part, units, wg_begin, items, hub_items, chunk_first,
        hub_partial, wtab;
    uint64_t bytes() const
    {
        return idx16.bytes + idx32.bytes + sub_start.bytes + part.bytes + units.bytes +
               wg_begin.bytes + items.bytes + hub_items.bytes + chunk_first.bytes +
               hub_partial.bytes + wtab.bytes;
    }
};

// Buffers of a personalized-PageRank batch that a finished batch leaves to
// the next one on the same graph (see PprBatch).
constexpr int kKeepBufs = 11;

struct Plan {
    uint32_t N = 0, E = 0;
    uint32_t n_hub = 0, n_w = 0, n_h = 0, n_nonlow = 0, n_nonzero = 0, n_dangling = 0;
    uint32_t low_slices = 0, hub_chunks = 0;
    uint32_t low_max = 0, hub_min = 0, hub_chunk = 0, hot_bytes = 0, lane_max = 16;
    uint32_t max_tables = 16;  // partitioned layout: LDS tables at most (tuning)
    bool autotuned = false;    // chosen by ag_pagerank_autotune
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
    // partitioned-table layout (null: the row layout above serves the sweep)
    std::unique_ptr<TabPlan> tab;
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
               ppr_ch_first.bytes + ppr_multi.bytes + keep_bytes() + (tab ? tab->bytes() : 0);
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
