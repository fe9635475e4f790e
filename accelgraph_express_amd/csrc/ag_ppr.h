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
    uint32_t done;     // residual <= tol or iters == max_iter
    uint32_t iters;
    uint32_t max_iter;
    double tol;
    double residual;
    double dangling;
    uint32_t stalled;  // low-precision column frozen for its FP32 finish
    uint32_t finish;   // 1: freeze at the stall / tolerance instead of finishing
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
    // capacity: column slots (multiple of 4, <= 64); cpw: columns per lane
    // group in {4, 8, 16, 32, 64}; fmt: AG_PREC_* storage of contributions.
    PprBatch(ag_ctx *ctx, ag_graph *g, uint32_t capacity, uint32_t cpw, uint32_t fmt,
             float damp);
    ~PprBatch();
    PprBatch(const PprBatch &) = delete;
    PprBatch &operator=(const PprBatch &) = delete;

    static uint64_t bytes_for(uint32_t N, uint32_t capacity, uint32_t fmt);
    uint32_t capacity() const { return cap_; }
    uint32_t active() const { return used_count_; }
    uint32_t free_slots() const { return cap_ - used_count_; }
    uint64_t bytes() const;
    uint32_t cpw() const { return cpw_; }
    void set_cpw(uint32_t cpw);

    // Maps vertex ids to labels and merges duplicates (weights 1/n each).
    PprSeedSet make_seeds(const uint32_t *seeds, uint32_t n, double tol, uint32_t max_iter);
    // Takes a free slot for a query; its columns are initialised on the next
    // sweep (all admissions of a step in one launch).  Returns the slot.
    uint32_t admit(const PprSeedSet &q, uint64_t tag);
    // Runs `n` sweeps over all active columns (graph replay when enabled).
    void sweep(uint32_t n, bool replay);
    // Synchronises and returns the slots whose columns finished or stalled.
    struct Done {
        uint32_t slot;
        uint64_t tag;
        uint32_t iters;
        double residual;
        bool stalled; // low precision: ready to move to an FP32 batch
    };
    std::vector<Done> poll();
    // Moves stalled columns into `dst` (an FP32 batch) in one launch, iterate
    // and state included; the source slots are released.
    void migrate(const std::vector<uint32_t> &slots, PprBatch &dst);
    // Copies a finished column (caller's vertex order) to host memory.
    void extract(uint32_t slot, float *host_out);
    // Frees a slot (a running column is stopped first).
    void release(uint32_t slot);
    bool used(uint32_t slot) const { return slot < cap_ && used_[slot]; }
    uint64_t tag(uint32_t slot) const { return tags_[slot]; }
    double last_sweep_ms() const { return last_ms_; }

  private:
    void flush();
    void upload_seed_table();
    void launch_sweep(bool even);
    ag_ctx *ctx_;
    Plan *p_;
    uint32_t N_, cap_, cpw_, fmt_;
    uint32_t used_count_ = 0, hi_ = 0; // hi_: one past the highest used slot
    float damp_;
    int cur_ = 0; // contribution buffer read by the next sweep
    uint32_t grid_;
    bool dirty_ = true;
    DevBuf X_, C_[2], cols_, bstate_, blk_res_, blk_dang_, hub_part_, hub_res_, hub_dang_;
    DevBuf seed_bits_, seed_v_, seed_c_, seed_w_, out_, desc_;
    std::vector<PprSeedSet> q_;    // per slot
    std::vector<uint64_t> tags_;   // per slot
    std::vector<uint8_t> used_;    // per slot
    std::vector<uint32_t> pending_; // admitted, columns not initialised yet
    std::vector<uint32_t> stop_;    // released while running: mark done on device
    hipGraphExec_t gexec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    double last_ms_ = 0.0;
};

// A graph's column slots in two stages: columns run in the requested storage
// precision; with BF16/FP16 storage a column that stalls (or meets its
// tolerance) is frozen and finished in an FP32 batch.  FP32 runs use one stage.
class PprPipeline {
  public:
    PprPipeline(ag_ctx *ctx, ag_graph *g, uint32_t cap, uint32_t cpw, uint32_t fmt, float damp);
    static uint64_t bytes_for(uint32_t N, uint32_t cap, uint32_t fmt);
    uint32_t free_slots() const { return first_->free_slots(); }
    uint32_t active() const { return first_->active() + (fp32_ ? fp32_->active() : 0u); }
    PprSeedSet make_seeds(const uint32_t *seeds, uint32_t n, double tol, uint32_t max_iter)
    {
        return first_->make_seeds(seeds, n, tol, max_iter);
    }
    void admit(const PprSeedSet &q, uint64_t tag) { first_->admit(q, tag); }
    void sweep(uint32_t n, bool replay);
    struct Finished {
        uint64_t tag;
        uint32_t iters;
        double residual;
    };
    // Migrates stalled columns, retires finished ones; `out(tag)` names the
    // host buffer for a finished column's result (nullptr: drop it).
    template <typename F> std::vector<Finished> retire(F out);
    bool cancel(uint64_t tag);

  private:
    std::unique_ptr<PprBatch> first_, fp32_;
};

template <typename F>
std::vector<PprPipeline::Finished> PprPipeline::retire(F out)
{
    std::vector<Finished> fin;
    std::vector<uint32_t> moving;
    for (auto &d : first_->poll()) {
        if (d.stalled) {
            if (fp32_ && moving.size() < fp32_->free_slots()) moving.push_back(d.slot);
            continue; // without room it stays frozen until a later step
        }
        first_->extract(d.slot, out(d.tag));
        fin.push_back({d.tag, d.iters, d.residual});
        first_->release(d.slot);
    }
    if (!moving.empty()) first_->migrate(moving, *fp32_);
    if (fp32_) {
        for (auto &d : fp32_->poll()) {
            fp32_->extract(d.slot, out(d.tag));
            fin.push_back({d.tag, d.iters, d.residual});
            fp32_->release(d.slot);
        }
    }
    return fin;
}

uint32_t ppr_autotune(ag_ctx *ctx, ag_graph *g, uint32_t fmt);

} // namespace ag
