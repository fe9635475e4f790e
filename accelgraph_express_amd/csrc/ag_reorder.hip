// Locality relabelling of an edge list: the reference's -l / -L / -O options
// (README.md:362-411; Arguments.lmode / lmode_l2 / lmode_l3,
// graphConfig.h:28-30).  The reference's implementation lives in the OpenGraph
// reorder module, which is not available; the modes restate the published
// algorithms with these conventions:
//
//   degrees count the edge list: out = occurrences as src, in = as dst,
//   in+out = both; avg = E / N in uint32 (the structs' avg_degree); a hub has
//   degree > avg.
//     1-3   degree sort: descending degree, ties by ascending id
//     4-5   DBG (Faldu et al., IISWC 2019): groups d >= 32avg, >= 16avg,
//           >= 8avg, >= 4avg, >= 2avg, >= avg, 2d >= avg, rest, in that order;
//           ascending id inside a group
//     6-7   HUBSort (Zhang et al., 2017): hubs by descending degree (ties by
//           id), then the rest in id order
//     8-9   HUBCluster (Balaji and Lucia, IISWC 2018): hubs in id order, then
//           the rest in id order
//     10    random: ascending key lowbias32(v * 0x9E3779B9 ^ seed), ties by id
//   (11, labels from a file, is host code in the drop-in layer.)
//
// On the device: one pass of atomics counts degrees, one kernel turns them
// into sort keys (ascending order = the mode's order), the builder's stable
// LSD radix sort orders (key, id) pairs with as many 8-bit passes as the key
// range needs, and a scatter writes labels[order[i]] = i.  Relabelling an edge
// list is one gather per endpoint, after a check that the labels are a
// permutation of [0, N).
#include <algorithm>

#include "ag_internal.h"

namespace ag {

namespace {

constexpr int kBlock = 256;

__global__ void k_add_u32(uint32_t *__restrict__ a, const uint32_t *__restrict__ b, uint32_t n)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] += b[i];
}

__device__ __forceinline__ uint32_t lowbias32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ void k_reorder_keys(const uint32_t *__restrict__ deg, uint32_t N, uint32_t mode,
                               uint64_t avg, uint32_t maxdeg, uint32_t seed,
                               uint32_t *__restrict__ key, uint32_t *__restrict__ id)
{
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= N) return;
    const uint64_t d = deg ? deg[v] : 0;
    uint32_t k = 0;
    switch (mode) {
    case 1: case 2: case 3: k = maxdeg - (uint32_t)d; break;
    case 4: case 5:
        k = d >= 32 * avg ? 0u : d >= 16 * avg ? 1u : d >= 8 * avg ? 2u : d >= 4 * avg ? 3u
          : d >= 2 * avg ? 4u : d >= avg ? 5u : 2 * d >= avg ? 6u : 7u;
        break;
    case 6: case 7: k = d > avg ? maxdeg - (uint32_t)d : maxdeg; break;
    case 8: case 9: k = d > avg ? 0u : 1u; break;
    case 10: k = lowbias32(v * 0x9E3779B9u ^ seed); break;
    default: k = 0; break;
    }
    key[v] = k;
    id[v] = v;
}

__global__ void k_scatter_labels(const uint32_t *__restrict__ order, uint32_t N,
                                 uint32_t *__restrict__ labels)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N) labels[order[i]] = i;
}

// seen[labels[v]] += 1 for every v (labels already checked < N)
__global__ void k_mark_labels(const uint32_t *__restrict__ labels, uint32_t N,
                              uint32_t *__restrict__ seen)
{
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v < N) atomicAdd(&seen[labels[v]], 1u);
}

// bad[0] += number of labels not hit exactly once
__global__ void k_count_unseen(const uint32_t *__restrict__ seen, uint32_t N, uint32_t *bad)
{
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    const bool b = v < N && seen[v] != 1u;
    const uint64_t m = __ballot(b);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(bad, (uint32_t)__popcll(m));
}

__global__ void k_apply_labels(const uint32_t *__restrict__ labels, uint32_t *__restrict__ a,
                               uint32_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kBlock)
        a[i] = labels[a[i]];
}

uint32_t stride_grid(uint64_t n)
{
    return std::max<uint32_t>(1, std::min<uint32_t>(ceil_div(n, kBlock), 65536));
}

void check_ids(ag_ctx *ctx, const uint32_t *src, const uint32_t *dst, uint32_t E, uint32_t N)
{
    if (E == 0) return;
    if (N == 0) fail(AG_ERR_INVALID, "edges given for a graph of 0 vertices");
    const uint32_t m = device_max2(ctx, src, dst, E); // one pass, one wait
    if (m >= N) fail(AG_ERR_INVALID, "vertex id %u >= num_vertices %u", m, N);
}

// deg[v] = number of ids equal to v (ids < N), from a stable radix sort of
// the ids (the values ride along unused) and the lengths of the sorted runs.
// One device atomic per id took 3.1 ms on R-MAT scale 22's 67 M in-edges;
// the sort and the run bounds take about half of that.  Scratch slots:
// the sorted pair and the bounds' start array.
void count_ids(ag_ctx *ctx, const uint32_t *ids, uint32_t E, uint32_t N, uint32_t *deg)
{
    uint32_t *sk = static_cast<uint32_t *>(scratch(ctx, kScrPairA, (size_t)E * 4));
    uint32_t *sv = static_cast<uint32_t *>(scratch(ctx, kScrPairB, (size_t)E * 4));
    radix_sort_pairs(ctx, ids, ids, sk, sv, E, bits_for(N));
    uint32_t *first = static_cast<uint32_t *>(scratch(ctx, kScrInSrc, (size_t)N * 4));
    vertex_map(ctx, sk, E, N, deg, first, nullptr);
}

} // namespace

void reorder_labels_dev(ag_ctx *ctx, const uint32_t *src, const uint32_t *dst, uint32_t E,
                        uint32_t N, uint32_t mode, uint32_t seed, uint32_t *labels)
{
    if (mode > 10) fail(AG_ERR_INVALID, "unknown relabelling mode %u (0-10)", mode);
    if (N == 0) return;
    hipStream_t s = ctx->stream;
    std::lock_guard<std::recursive_mutex> lk(ctx->scratch_mu); // count_ids' scratch slots
    check_ids(ctx, src, dst, E, N);
    const bool out = mode == 1 || mode == 3 || mode == 4 || mode == 6 || mode == 8;
    const bool in = mode == 2 || mode == 3 || mode == 5 || mode == 7 || mode == 9;
    DevBuf deg, key((size_t)N * 4), id((size_t)N * 4), skey((size_t)N * 4), order((size_t)N * 4);
    uint32_t maxdeg = 0;
    if (out || in) {
        deg.alloc((size_t)N * 4);
        AG_HIP(hipMemsetAsync(deg.p, 0, (size_t)N * 4, s));
        if (E && out && in) {
            DevBuf d2((size_t)N * 4);
            count_ids(ctx, src, E, N, deg.as<uint32_t>());
            count_ids(ctx, dst, E, N, d2.as<uint32_t>());
            hipLaunchKernelGGL(k_add_u32, dim3(ceil_div(N, kBlock)), dim3(kBlock), 0, s,
                               deg.as<uint32_t>(), d2.as<uint32_t>(), N);
            AG_HIP(hipGetLastError());
            AG_HIP(hipStreamSynchronize(s)); // d2 is freed on return
        } else if (E) {
            count_ids(ctx, out ? src : dst, E, N, deg.as<uint32_t>());
        }
        maxdeg = device_max1(ctx, deg.as<uint32_t>(), N);
    }
    const uint64_t avg = E / N;
    hipLaunchKernelGGL(k_reorder_keys, dim3(ceil_div(N, kBlock)), dim3(kBlock), 0, s,
                       deg.p ? deg.as<uint32_t>() : nullptr, N, mode, avg, maxdeg, seed,
                       key.as<uint32_t>(), id.as<uint32_t>());
    AG_HIP(hipGetLastError());
    uint32_t bits = 0; // mode 0: one key, the sort keeps id order
    switch (mode) {
    case 1: case 2: case 3: case 6: case 7: bits = bits_for((uint64_t)maxdeg + 1); break;
    case 4: case 5: bits = 3; break;
    case 8: case 9: bits = 1; break;
    case 10: bits = 32; break;
    default: break;
    }
    radix_sort_pairs(ctx, key.as<uint32_t>(), id.as<uint32_t>(), skey.as<uint32_t>(),
                     order.as<uint32_t>(), N, bits);
    hipLaunchKernelGGL(k_scatter_labels, dim3(ceil_div(N, kBlock)), dim3(kBlock), 0, s,
                       order.as<uint32_t>(), N, labels);
    AG_HIP(hipGetLastError());
    AG_HIP(hipStreamSynchronize(s)); // the temporaries are freed on return
}

void relabel_edges_dev(ag_ctx *ctx, const uint32_t *labels, uint32_t N, uint32_t *src, uint32_t *dst,
                       uint32_t E)
{
    hipStream_t s = ctx->stream;
    check_ids(ctx, src, dst, E, N);
    if (N) {
        const uint32_t m = device_max1(ctx, labels, N);
        if (m >= N) fail(AG_ERR_INVALID, "label %u >= num_vertices %u", m, N);
        DevBuf seen((size_t)N * 4), bad(4);
        AG_HIP(hipMemsetAsync(seen.p, 0, (size_t)N * 4, s));
        AG_HIP(hipMemsetAsync(bad.p, 0, 4, s));
        hipLaunchKernelGGL(k_mark_labels, dim3(ceil_div(N, kBlock)), dim3(kBlock), 0, s, labels, N,
                           seen.as<uint32_t>());
        hipLaunchKernelGGL(k_count_unseen, dim3(ceil_div(N, kBlock)), dim3(kBlock), 0, s,
                           seen.as<uint32_t>(), N, bad.as<uint32_t>());
        AG_HIP(hipGetLastError());
        uint32_t nb = 0;
        AG_HIP(hipMemcpyAsync(&nb, bad.p, 4, hipMemcpyDeviceToHost, s));
        AG_HIP(hipStreamSynchronize(s));
        if (nb) fail(AG_ERR_INVALID, "labels are not a permutation of [0, %u): %u values repeat or miss", N, nb);
    }
    if (E) {
        hipLaunchKernelGGL(k_apply_labels, dim3(stride_grid(E)), dim3(kBlock), 0, s, labels, src, E);
        hipLaunchKernelGGL(k_apply_labels, dim3(stride_grid(E)), dim3(kBlock), 0, s, labels, dst, E);
        AG_HIP(hipGetLastError());
    }
    AG_HIP(hipStreamSynchronize(s));
}

} // namespace ag

using namespace ag;

extern "C" {

int ag_reorder_labels_device(ag_ctx *ctx, const uint32_t *d_src, const uint32_t *d_dst,
                             uint32_t num_edges, uint32_t num_vertices, uint32_t mode,
                             uint32_t seed, uint32_t *d_labels)
{
    return guard([&] {
        PhaseRange phase_("ag_reorder_labels_device");
        check_ctx(ctx);
        if (num_edges && (!d_src || !d_dst)) fail(AG_ERR_INVALID, "null edge array");
        if (num_vertices && !d_labels) fail(AG_ERR_INVALID, "null labels");
        reorder_labels_dev(ctx, d_src, d_dst, num_edges, num_vertices, mode, seed, d_labels);
    });
}

int ag_reorder_labels(ag_ctx *ctx, const uint32_t *src, const uint32_t *dst, uint32_t num_edges,
                      uint32_t num_vertices, uint32_t mode, uint32_t seed, uint32_t *labels)
{
    return guard([&] {
        PhaseRange phase_("ag_reorder_labels");
        check_ctx(ctx);
        if (num_edges && (!src || !dst)) fail(AG_ERR_INVALID, "null edge array");
        if (num_vertices && !labels) fail(AG_ERR_INVALID, "null labels");
        hipStream_t s = ctx->stream;
        DevBuf ds((size_t)std::max<uint32_t>(num_edges, 1) * 4),
            dd((size_t)std::max<uint32_t>(num_edges, 1) * 4),
            dl((size_t)std::max<uint32_t>(num_vertices, 1) * 4);
        if (num_edges) {
            AG_HIP(hipMemcpyAsync(ds.p, src, (size_t)num_edges * 4, hipMemcpyHostToDevice, s));
            AG_HIP(hipMemcpyAsync(dd.p, dst, (size_t)num_edges * 4, hipMemcpyHostToDevice, s));
        }
        reorder_labels_dev(ctx, ds.as<uint32_t>(), dd.as<uint32_t>(), num_edges, num_vertices, mode,
                           seed, dl.as<uint32_t>());
        if (num_vertices)
            AG_HIP(hipMemcpyAsync(labels, dl.p, (size_t)num_vertices * 4, hipMemcpyDeviceToHost, s));
        AG_HIP(hipStreamSynchronize(s));
    });
}

int ag_relabel_edges_device(ag_ctx *ctx, const uint32_t *d_labels, uint32_t num_vertices,
                            uint32_t *d_src, uint32_t *d_dst, uint32_t num_edges)
{
    return guard([&] {
        PhaseRange phase_("ag_relabel_edges_device");
        check_ctx(ctx);
        if (num_edges && (!d_src || !d_dst)) fail(AG_ERR_INVALID, "null edge array");
        if (num_vertices && !d_labels) fail(AG_ERR_INVALID, "null labels");
        relabel_edges_dev(ctx, d_labels, num_vertices, d_src, d_dst, num_edges);
    });
}

int ag_relabel_edges(ag_ctx *ctx, const uint32_t *labels, uint32_t num_vertices, uint32_t *src,
                     uint32_t *dst, uint32_t num_edges)
{
    return guard([&] {
        PhaseRange phase_("ag_relabel_edges");
        check_ctx(ctx);
        if (num_edges && (!src || !dst)) fail(AG_ERR_INVALID, "null edge array");
        if (num_vertices && !labels) fail(AG_ERR_INVALID, "null labels");
        hipStream_t s = ctx->stream;
        DevBuf ds((size_t)std::max<uint32_t>(num_edges, 1) * 4),
            dd((size_t)std::max<uint32_t>(num_edges, 1) * 4),
            dl((size_t)std::max<uint32_t>(num_vertices, 1) * 4);
        if (num_edges) {
            AG_HIP(hipMemcpyAsync(ds.p, src, (size_t)num_edges * 4, hipMemcpyHostToDevice, s));
            AG_HIP(hipMemcpyAsync(dd.p, dst, (size_t)num_edges * 4, hipMemcpyHostToDevice, s));
        }
        if (num_vertices)
            AG_HIP(hipMemcpyAsync(dl.p, labels, (size_t)num_vertices * 4, hipMemcpyHostToDevice, s));
        relabel_edges_dev(ctx, dl.as<uint32_t>(), num_vertices, ds.as<uint32_t>(), dd.as<uint32_t>(),
                          num_edges);
        if (num_edges) {
            AG_HIP(hipMemcpyAsync(src, ds.p, (size_t)num_edges * 4, hipMemcpyDeviceToHost, s));
            AG_HIP(hipMemcpyAsync(dst, dd.p, (size_t)num_edges * 4, hipMemcpyDeviceToHost, s));
        }
        AG_HIP(hipStreamSynchronize(s));
    });
}

} // extern "C"
