// Batched personalized PageRank: a device-resident set of column slots over
// one graph, advanced sweep by sweep (internal; not part of the C ABI).
#pragma once

#include <algorithm>
#include <deque>
#include <memory>
#include <vector>

#include "ag_plan.h"

namespace ag {

// Per-column device state.
struct PprCol {
    float t;           // teleport factor for the next sweep: (1 - d) + d * dangling
    uint32_t done;     // residual <= tol (in FP32) or iters == max_iter
    uint32_t iters;
    uint32_t max_iter;
    double tol;
    double residual;
    double dangling;
    uint32_t coarse;   // BF16 storage: 1 while sweeps read the high halves only
    uint32_t fine_at;  // sweep count when the column switched to full FP32 reads
    float w;           // > 0: the next sweep takes an extrapolated step of weight w
    uint32_t since;    // plain sweeps since the start or the last extrapolated step
    double rho;        // last residual ratio R_k / R_(k-1)
};

struct PprBatchState {
    uint32_t K;      // one past the highest used column slot
    uint32_t nzrows; // seed rows without in-edges
    uint32_t sweeps;
    uint32_t nseeds; // entries in the seed table
};

// A query as the batch sees it: unique labels and weights.
struct PprSeedSet {
    std::vector<uint32_t> labels;
    std::vector<float> w;
    double tol = 1e-4;
    uint32_t max_iter = 1000;
};

// Column slots of one graph.  Storage AG_PREC_FP32 keeps FP32 contributions;
// AG_PREC_BF16 keeps each FP32 contribution split into two 16-bit planes (the
// high halves, read alone as BF16 while a column is coarse, and the low
// halves, read as well once the column's residual stalls), so a column
// finishes in exact FP32 in place.
class PprBatch {
  public:
    // capacity: column slots (multiple of 8, <= 64); cpw: widest column group
    // one row access of a wave covers, {4, 8, 16, 32, 64} (BF16: at least 8).
    // extrapolate: columns take Aitken-type extrapolated steps once their
    // residual falls geometrically (fewer sweeps; same stop rule and bound).
    // exact_keep: buffers left on the graph by an earlier batch are lent only
    // when their size is exactly this batch's (others are freed), so the
    // batch holds exactly bytes_for(capacity) (the serving loop's budget).
    // `stream`: the stream all of the batch's work is queued on (null: the
    // context's); the serving loop gives each graph's batch its own.
    PprBatch(ag_ctx *ctx, ag_graph *g, uint32_t capacity, uint32_t cpw, uint32_t fmt,
             float damp, bool extrapolate = false, bool exact_keep = false,
             hipStream_t stream = nullptr);
    ~PprBatch();
    PprBatch(const PprBatch &) = delete;
    PprBatch &operator=(const PprBatch &) = delete;

    // Device bytes of a batch of `capacity` columns on g (builds g's plan):
    // the buffers the constructor allocates plus 1 MiB for the seed tables.
    static uint64_t bytes_for(ag_ctx *ctx, ag_graph *g, uint32_t capacity, uint32_t fmt);
    uint32_t capacity() const { return cap_; }
    uint32_t active() const { return used_count_; }
    hipStream_t stream() const { return s_; }
    uint32_t free_slots() const { return cap_ - used_count_; }
    uint64_t bytes() const;
    uint32_t cpw() const { return cpw_; }
    void set_cpw(uint32_t cpw);

    // Maps vertex ids to labels and merges duplicates (weights 1/n each).
    PprSeedSet make_seeds(const uint32_t *seeds, uint32_t n, double tol, uint32_t max_iter);
    // Takes the lowest free slot for a query; its column is initialised on the
    // next sweep (all admissions of a step in one launch).  Returns the slot.
    uint32_t admit(const PprSeedSet &q, uint64_t tag);
    // Runs `n` sweeps over all active columns (graph replay when enabled).
    // With `timed` it waits for them and records the mean sweep time.
    void sweep(uint32_t n, bool replay, bool timed = true);
    struct Done {
        uint32_t slot;
        uint64_t tag;
        uint32_t iters;
        double residual;
        uint32_t fine_at;
    };
    // Synchronises and returns the slots whose columns finished.
    std::vector<Done> poll();
    // Sweeps until the first running column can meet its tolerance, from the
    // residuals seen by the last poll (each sweep assumed to cut the residual
    // by at least 0.6), clamped to [1, 8]; 4 before any residual is known.  A
    // finished column is frozen on the device, so sweeping past it does not
    // change its result; polling less often saves the host round trips.
    uint32_t sweeps_hint() const { return hint_; }
    // Sweeps until the slowest running column can meet its tolerance, from the
    // same estimate (64 when unknown); compact() skips short tails.
    uint32_t tail_sweeps() const { return tail_; }
    // Copies a finished column (caller's vertex order) to host memory.
    void extract(uint32_t slot, float *host_out);
    // Writes a finished column in vertex order to `dst`, memory the device can
    // write (device memory or mapped pinned host memory); asynchronous, in
    // stream order before any later reuse of the slot.
    void extract_to(uint32_t slot, float *dst);
    // Frees a slot (a running column is stopped first).
    void release(uint32_t slot);
    // Moves running columns from the top slots into free lower ones when that
    // narrows the column groups the sweeps walk (no admissions pending).
    void compact();
    bool cancel(uint64_t tag);
    bool used(uint32_t slot) const { return slot < cap_ && used_[slot]; }
    uint64_t tag(uint32_t slot) const { return tags_[slot]; }
    double last_sweep_ms() const { return last_ms_; }

    struct Finished {
        uint64_t tag;
        uint32_t iters;
        double residual;
        uint32_t fine_at;
    };
    // Retires finished columns; `out(tag)` names the buffer for a column's
    // result (nullptr: none): host memory filled synchronously, or with
    // `device_out` device-writable memory filled asynchronously.  With
    // `may_compact` the remaining columns are packed afterwards.
    template <typename F>
    std::vector<Finished> retire(F out, bool may_compact, bool device_out = false);

  private:
    void flush();
    void choose_layout();
    // split block the layout rule asks for with `hi` slots in use
    uint32_t layout_for(uint32_t hi) const;
    void upload_seed_table();
    void launch_sweep(bool from0);
    uint32_t lanes_per_row() const;
    ag_ctx *ctx_;
    hipStream_t s_;
    Plan *p_;
    uint32_t N_, cap_, cpw_ = 0, fmt_;
    uint32_t used_count_ = 0, hi_ = 0; // hi_: one past the highest used slot
    float damp_;
    bool extrap_;
    // forward CSR of the graph (vertex ids): admissions run their first sweep
    // as a push over the seeds' out-edges; null without one or with
    // extrapolated steps (pull sweep)
    const uint32_t *fwd_off_ = nullptr, *fwd_dst_ = nullptr;
    int cur_ = 0;          // contribution buffer read by the next sweep
    uint32_t graph_lpr_ = 0; // lanes per row the captured graph was built for
    uint32_t blk_ = 64;      // BF16 split block (columns per highs/lows block)
    uint32_t graph_blk_ = 0; // split block the captured graph was built for
    bool dirty_ = true;
    int blocks_[5] = {0, 0, 0, 0, 0}; // sweep workgroups per CU by log2(lanes per row)
    DevBuf X_, C_[2], cols_, bstate_, blk_res_, blk_dang_, part_;
    DevBuf seed_bits_, seed_v_, seed_c_, seed_w_, zrows_, out_, desc_;
    std::vector<PprSeedSet> q_;    // per slot
    std::vector<uint64_t> tags_;   // per slot
    std::vector<uint8_t> used_;    // per slot
    std::vector<uint32_t> pending_; // admitted, columns not initialised yet
    std::vector<uint32_t> stop_;    // released while running: mark done on device
    hipGraphExec_t gexec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    double last_ms_ = 0.0;
    uint32_t hint_ = 4;
    uint32_t tail_ = 64;
};

template <typename F>
std::vector<PprBatch::Finished> PprBatch::retire(F out, bool may_compact, bool device_out)
{
    std::vector<Finished> fin;
    for (auto &d : poll()) {
        float *dst = out(d.tag);
        if (device_out) {
            if (dst) extract_to(d.slot, dst);
        } else {
            extract(d.slot, dst);
        }
        fin.push_back({d.tag, d.iters, d.residual, d.fine_at});
        release(d.slot);
    }
    if (may_compact) compact();
    return fin;
}

// Columns-per-wave for g, timed on a batch of `capacity` columns (default
// 64; the result is cached on the graph only for 64).
uint32_t ppr_autotune(ag_ctx *ctx, ag_graph *g, uint32_t fmt, uint32_t capacity = 64,
                      bool exact_keep = false);
// Frees the column buffers a finished batch left on the graph's plan.
void ppr_release_kept(ag_graph *g);
void ensure_ppr_chunks(ag_ctx *ctx, Plan *p);
// Host copy of the plan's rank array (vertex id -> label), made on first use.
const std::vector<uint32_t> &rank_host(ag_ctx *ctx, Plan *p);

} // namespace ag
