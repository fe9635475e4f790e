// Batched personalized PageRank: a device-resident set of column slots over
// one graph, advanced sweep by sweep (internal; not part of the C ABI).
#pragma once

#include <deque>
#include <vector>

#include "ag_plan.h"

namespace ag {

// Per-column device state.
struct PprCol {
    float t;           // teleport factor for the next sweep: (1 - d) + d * dangling
    uint32_t done;     // residual <= tol or iters == max_iter
    uint32_t iters;
    uint32_t max_iter;
    double tol;
    double residual;
    double dangling;
};

struct PprBatchState {
    uint32_t K;      // active (compacted) column slots
    uint32_t pad0;
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

class PprBatch {
  public:
    // capacity: column slots (multiple of 32, <= 64); cpw: columns per wave
    // group in {4, 8, 16, 32}; fmt: AG_PREC_* storage of contributions.
    PprBatch(ag_ctx *ctx, ag_graph *g, uint32_t capacity, uint32_t cpw, uint32_t fmt,
             float damp);
    ~PprBatch();
    PprBatch(const PprBatch &) = delete;
    PprBatch &operator=(const PprBatch &) = delete;

    static uint64_t bytes_for(uint32_t N, uint32_t capacity, uint32_t fmt);
    uint32_t capacity() const { return cap_; }
    uint32_t active() const { return K_; }
    uint32_t free_slots() const { return cap_ - K_; }
    uint64_t bytes() const;
    uint32_t cpw() const { return cpw_; }
    void set_cpw(uint32_t cpw);

    // Maps vertex ids to labels and merges duplicates (weights 1/n each).
    PprSeedSet make_seeds(const uint32_t *seeds, uint32_t n, double tol, uint32_t max_iter);
    // Admits a query into the next slot; returns the slot.  Enqueued on the
    // context stream (no host synchronisation).
    uint32_t admit(const PprSeedSet &q, uint64_t tag);
    // Runs `n` sweeps over all active columns (graph replay when enabled).
    void sweep(uint32_t n, bool replay);
    // Synchronises and returns the slots whose columns finished.
    struct Done {
        uint32_t slot;
        uint64_t tag;
        uint32_t iters;
        double residual;
    };
    std::vector<Done> poll();
    // Copies a finished column (caller's vertex order) to host memory.
    void extract(uint32_t slot, float *host_out);
    // Drops a slot; the last active slot moves into it.  Returns the tag that
    // moved (or ~0 when none moved) so callers can update their maps.
    uint64_t remove(uint32_t slot);
    uint64_t tag(uint32_t slot) const { return tags_[slot]; }
    double last_sweep_ms() const { return last_ms_; }

  private:
    void upload_seed_table();
    void launch_sweep(bool even);
    ag_ctx *ctx_;
    ag_graph *g_;
    Plan *p_;
    uint32_t N_, cap_, cpw_, fmt_, K_ = 0;
    float damp_;
    int cur_ = 0; // contribution buffer read by the next sweep
    uint32_t grid_;
    DevBuf X_, C_[2], cols_, bstate_, blk_res_, blk_dang_, hub_part_, hub_res_, hub_dang_;
    DevBuf seed_bits_, seed_v_, seed_c_, seed_w_, out_;
    std::vector<PprSeedSet> q_;    // per slot
    std::vector<uint64_t> tags_;   // per slot
    hipGraphExec_t gexec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    double last_ms_ = 0.0;
};

uint32_t ppr_autotune(ag_ctx *ctx, ag_graph *g, uint32_t fmt);

} // namespace ag
