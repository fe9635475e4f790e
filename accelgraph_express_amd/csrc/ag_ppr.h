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
        bool stalled; // low precision: ready to move to an FP32 batch
    };
    std::vector<Done> poll();
    // Moves a stalled column into `dst` (an FP32 batch), iterate and state
    // included; returns the new slot in dst.  The slot here is not removed.
    uint32_t migrate(uint32_t slot, PprBatch &dst);
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
    auto by_slot = [](const PprBatch::Done &a, const PprBatch::Done &b) { return a.slot > b.slot; };
    auto done = first_->poll();
    std::sort(done.begin(), done.end(), by_slot);
    for (auto &d : done) {
        if (d.stalled) {
            if (!fp32_ || fp32_->free_slots() == 0) continue; // stays frozen until room
            first_->migrate(d.slot, *fp32_);
        } else {
            first_->extract(d.slot, out(d.tag));
            fin.push_back({d.tag, d.iters, d.residual});
        }
        first_->remove(d.slot);
    }
    if (fp32_) {
        auto d2 = fp32_->poll();
        std::sort(d2.begin(), d2.end(), by_slot);
        for (auto &d : d2) {
            fp32_->extract(d.slot, out(d.tag));
            fin.push_back({d.tag, d.iters, d.residual});
            fp32_->remove(d.slot);
        }
    }
    return fin;
}

uint32_t ppr_autotune(ag_ctx *ctx, ag_graph *g, uint32_t fmt);

} // namespace ag
