// Internal declarations shared by the library's translation units.
// Not part of the C ABI (include/accelgraph_amd.h is).
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "accelgraph_amd.h"

namespace ag {

// ---- errors -----------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(int code, const char *msg);
[[noreturn]] void fail(int code, const char *fmt, ...);

#define AG_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            ::ag::fail(AG_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                       __FILE__, __LINE__);                                            \
    } while (0)

// Names one library phase for a profiler (rocprofv3 --marker-trace shows it as
// a range on the calling thread, around the host work and the kernel launches
// it issues).  The phase ends when the scope does, also by an exception.
// Without a profiler attached a push/pop pair costs a few tens of ns.
struct PhaseRange {
    explicit PhaseRange(const char *name) { roctxRangePushA(name); }
    ~PhaseRange() { roctxRangePop(); }
    PhaseRange(const PhaseRange &) = delete;
    PhaseRange &operator=(const PhaseRange &) = delete;
};

// Every extern "C" entry point runs its body through guard(): exceptions become
// status codes plus ag_last_error() text.
template <typename F>
int guard(F &&f)
{
    try {
        f();
        set_last_error(AG_OK, "");
        return AG_OK;
    } catch (const Error &e) {
        set_last_error(e.code, e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error(AG_ERR_OOM, "host allocation failed");
        return AG_ERR_OOM;
    } catch (const std::exception &e) {
        set_last_error(AG_ERR_INVALID, e.what());
        return AG_ERR_INVALID;
    }
}

// ---- device memory ----------------------------------------------------------
// Owning device buffer; frees on destruction.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n)
    {
        release();
        if (n == 0) n = 4;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            fail(AG_ERR_OOM, "hipMalloc(%zu bytes): %s", n, hipGetErrorString(e));
        }
        bytes = n;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// ---- context ----------------------------------------------------------------
} // namespace ag

struct ag_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    std::string arch;
    // Scratch arena of the builders (radix passes, scans, uploads): buffers
    // grow on demand and are kept until the context is destroyed, so a build
    // does not pay hipMalloc / hipFree (which waits for the device) per step.
    // The mutex serialises the host threads that use it.
    static constexpr int kScratchSlots = 26;
    std::recursive_mutex scratch_mu;
    ag::DevBuf scratch[kScratchSlots];
    // buffers of grown slots that queued work may still read: freed once the
    // stream is seen idle (or with the context, after its stream is drained)
    std::vector<ag::DevBuf> retired;
};

namespace ag {

// Scratch slots (see ag_ctx::scratch).
enum ScratchSlot : int {
    kScrHist = 0, kScrCsum, kScrSortK, kScrSortV, kScrScan0, // kScrScan0 .. +3: scan levels
    kScrMax = 8, kScrPairA, kScrPairB, kScrInSrc, kScrInDst,
    kScrLay0 = 13, // kScrLay0 .. +7: gather-layout build temporaries
    kScrBlk = 21,  // first row of each 256-slot block of an edge-slot fill
    kScrSort2K = 22, kScrSort2V = 23, // second ping-pong pair of radix_sort_pairs_late
    kScrTileMax = 24,                 // per-tile key maxima of a sort's first pass
    kScrLay8 = 25,                    // gather-layout build: rows of the table-sorted edges
};
// Device buffer of at least `bytes` in `slot`.  Growing a slot does not wait
// for the stream: the old buffer, which queued work may still read, is
// retired and freed at a later call that finds the stream idle, so the
// allocation overlaps the queued kernels.  Sizes are rounded up (1 MiB, plus
// 1 MiB of slack) so that an array of n + 1 words after one of n words does
// not reallocate.
inline void *scratch(ag_ctx *ctx, int slot, size_t bytes)
{
    if (!ctx->retired.empty() && hipStreamQuery(ctx->stream) == hipSuccess)
        ctx->retired.clear();
    DevBuf &b = ctx->scratch[slot];
    if (b.bytes < bytes) {
        if (b.p) ctx->retired.push_back(std::move(b));
        constexpr size_t kMiB = (size_t)1 << 20;
        const size_t n = (bytes + kMiB - 1) / kMiB * kMiB + kMiB;
        try {
            b.alloc(n);
        } catch (const Error &) {
            if (ctx->retired.empty()) throw;
            // short of memory: wait for the queued work, free what it held
            AG_HIP(hipStreamSynchronize(ctx->stream));
            ctx->retired.clear();
            b.alloc(n);
        }
    }
    return b.p;
}
template <typename T> inline T *scratch_as(ag_ctx *ctx, int slot, size_t count)
{
    return static_cast<T *>(scratch(ctx, slot, count * sizeof(T)));
}

// Gather layout of a graph for the pull kernels (built lazily, cached).
struct Plan;
struct PprState;
void plan_destroy(Plan *p);

// ---- tuning ----------------------------------------------------------------
struct Tuning {
    uint32_t low_max = 64;
    uint32_t hub_min = 6144;   // s22: rows of 4,096-6,143 in-edges do better as 64-lane rows
    // AG_TUNE_AUTO: chosen per graph from its hub rows (build_plan)
    uint32_t hub_chunk = AG_TUNE_AUTO;
    uint32_t hot_bytes = 160 * 1024 - 2048; // LDS table of the hottest contributions
    uint32_t layout = AG_LAYOUT_AUTO;       // gather layout of the pull sweep
    uint32_t max_tables = 16;               // LDS tables of the partitioned layout
    uint32_t lane_max = 16;                 // sub-rows up to this many edges: one lane each
};

} // namespace ag

struct ag_graph {
    uint32_t N = 0, E = 0;
    bool has_full_fwd = false; // sorted src/dst arrays + forward idx present
    bool has_inverse_full = false;
    ag::DevBuf fwd_src, fwd_dst, fwd_outdeg, fwd_idx, fwd_off;
    ag::DevBuf inv_src, inv_dst, inv_outdeg, inv_idx, inv_off;
    ag::Plan *plan = nullptr;
    // the graph's own gather-layout tuning (set by ag_pagerank_autotune);
    // without it a layout is built with the process-wide tuning
    bool has_tune = false;
    ag::Tuning tune;
    uint32_t ppr_cpw[3] = {0, 0, 0}; // autotuned columns-per-wave per precision
    ~ag_graph() { ag::plan_destroy(plan); }
    uint64_t bytes() const;
};

namespace ag {

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }
inline uint32_t bits_for(uint64_t limit) // bits needed for keys in [0, limit)
{
    uint32_t b = 0;
    while (b < 32 && (1ull << b) < limit) b++;
    return b;
}

ag_ctx *default_ctx();
void check_ctx(ag_ctx *ctx);
// Device address of the allocation-free entry point's static scratch on
// `device` (the current device), looked up once per device (ag_pagerank.hip).
void noalloc_scratch_init(int device);

// ---- sort / scan primitives (ag_sort.hip) ------------------------------------
// Exclusive scan of n u32 values into out[0..n]; out[n] = total.  in == out allowed.
void scan_exclusive_u32(ag_ctx *ctx, const uint32_t *in, uint32_t *out, uint32_t n);
// Stable LSD radix sort of (key, val) pairs by the low `key_bits` bits of key.
// Inputs are not modified; outputs may not alias inputs.
void radix_sort_pairs(ag_ctx *ctx, const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                      uint32_t *vout, uint32_t n, uint32_t key_bits);
// The same with a second value array moving with the first.
void radix_sort_triples(ag_ctx *ctx, const uint32_t *kin, const uint32_t *vin, const uint32_t *vin2,
                        uint32_t *kout, uint32_t *vout, uint32_t *vout2, uint32_t n,
                        uint32_t key_bits);
// The same as radix_sort_pairs, with the output arrays supplied by `out` (which may allocate
// them) after the first pass is queued, so that a host-side allocation runs
// while the device sorts.  The first pass writes to scratch.
void radix_sort_pairs_late(ag_ctx *ctx, const uint32_t *kin, const uint32_t *vin, uint32_t n,
                           uint32_t key_bits,
                           const std::function<void(uint32_t **kout, uint32_t **vout)> &out,
                           uint32_t *key_max = nullptr /* device word: atomicMax of the keys */);
// Degree / edges_idx / offsets of a key-sorted array (keys >= N are skipped).
void vertex_map(ag_ctx *ctx, const uint32_t *sorted_keys, uint32_t n, uint32_t N,
                uint32_t *degree, uint32_t *edges_idx, uint32_t *offsets /* N+1 or null */);
// atomicMax(*out, max over a and b), queued on the context's stream.
void device_max_async(ag_ctx *ctx, const uint32_t *a, const uint32_t *b, uint32_t n,
                      uint32_t *out);
// max over two arrays (returns host value); 0 for n == 0.
uint32_t device_max2(ag_ctx *ctx, const uint32_t *a, const uint32_t *b, uint32_t n);
uint32_t device_max1(ag_ctx *ctx, const uint32_t *a, uint32_t n);

Tuning &tuning();

} // namespace ag
