// Admission policy of the serving loop: which idle batches to evict and what
// column capacity a new batch gets under the HBM budget.  Pure host code with
// no HIP dependency, so that it is unit-tested on the CPU with a mock byte
// counter (tests/c/test_admission.cpp); ag_serve.hip applies its decisions.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace ag {

// A batch resident for another graph.
struct AdmitHolder {
    uint64_t reserved; // bytes it holds against the budget
    bool idle;         // no running column: may be evicted
};

struct AdmitPlan {
    uint32_t cap = 0;           // columns of the new batch; 0: no room, stay queued
    std::vector<size_t> evict;  // indices into the holders, in eviction order
};

// want: the configured capacity (a multiple of 8, at most 64).  bytes(cap):
// the new batch's reservation for `cap` columns.  Idle holders are evicted,
// largest first, while the full capacity does not fit; then the capacity is
// lowered in steps of 8 until it fits.  Busy holders are never evicted, and
// nothing is evicted when even the full eviction would not make room for 8
// columns.
template <typename Bytes>
AdmitPlan plan_admission(uint64_t budget, uint64_t reserved, uint32_t want,
                         const std::vector<AdmitHolder> &holders, Bytes bytes)
{
    AdmitPlan p;
    auto fits = [&](uint64_t held, uint32_t cap) { return held + bytes(cap) <= budget; };
    std::vector<size_t> idle;
    for (size_t i = 0; i < holders.size(); i++)
        if (holders[i].idle) idle.push_back(i);
    std::stable_sort(idle.begin(), idle.end(), [&](size_t a, size_t b) {
        return holders[a].reserved > holders[b].reserved;
    });
    uint64_t held = reserved;
    size_t k = 0;
    while (!fits(held, want) && k < idle.size()) {
        held -= std::min(held, holders[idle[k]].reserved);
        p.evict.push_back(idle[k]);
        k++;
    }
    uint32_t cap = want;
    while (cap >= 8 && !fits(held, cap)) cap -= 8;
    if (cap < 8) {
        p.evict.clear(); // evicting would not make room: keep every batch
        return p;
    }
    p.cap = cap;
    return p;
}

} // namespace ag
