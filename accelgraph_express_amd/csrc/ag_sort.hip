// Device sort / scan / vertex-mapping primitives for the GraphCSR builder.
//
// The reference builds its CSR with OpenGraph's stable counting / LSD radix
// sort by source (sortRunAlgorithms, modes at README.md:434-439, called at
// SSSP.c:548-553) followed by a vertex mapping that yields out_degree and
// edges_idx (fields read at capienv.c:409-426).  Here the same permutation is
// produced by stable LSD radix passes of digits of at most 8 bits (the widths
// balanced over the key: 22 bits are 8 + 7 + 7), each pass being:
//
//   k_radix_hist     per-tile digit histogram in LDS (tile = 4096 keys)
//   k_hist_*         digit-major exclusive scan of the tile histograms
//                    (chunk sums -> one-block chunk scan -> chunk apply)
//   k_radix_scatter  per-wave stable ranking with ballot-based digit
//                    matching, cross-wave prefix in LDS, scatter
//
// Every step is deterministic and atomic-free in global memory, so the sorted
// arrays are bit-identical run to run and identical to a serial stable
// counting sort.
#include <algorithm>

#include "ag_internal.h"

namespace ag {

namespace {

constexpr int kBlock = 256;       // threads per workgroup (4 wave64)
constexpr int kItems = 16;        // keys per thread per tile
constexpr int kTile = kBlock * kItems;
constexpr int kRadix = 256;       // 8-bit digits
constexpr int kWaves = kBlock / 64;
static_assert(kWaves == 4, "k_radix_hist combines four wave maxima");

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, int lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

// ---- generic exclusive scan ---------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t *__restrict__ in,
                                                        uint32_t n,
                                                        uint32_t *__restrict__ block_sums)
{
    __shared__ uint32_t ws[kWaves];
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) ws[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < kWaves; i++) t += ws[i];
        block_sums[blockIdx.x] = t;
    }
}

// Block-exclusive scan of one tile (blocked: thread t owns [t*16, t*16+16)),
// plus a carry-in per block.  The last block writes out[n] = total.  With
// `vec` (in and out 16-byte aligned) a full tile moves as 16-byte words.
__global__ __launch_bounds__(kBlock) void k_scan_down(const uint32_t *in, uint32_t *out,
                                                      uint32_t n,
                                                      const uint32_t *__restrict__ block_pref,
                                                      uint32_t nblocks, uint32_t vec)
{
    __shared__ uint32_t ws[kWaves];
    const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kItems;
    const bool full = vec && (uint64_t)(blockIdx.x + 1) * kTile <= n;
    uint32_t v[kItems];
    uint32_t s = 0;
    if (full) {
#pragma unroll
        for (int i = 0; i < kItems / 4; i++) {
            const uint4 q = reinterpret_cast<const uint4 *>(in + base)[i];
            v[4 * i] = q.x;
            v[4 * i + 1] = q.y;
            v[4 * i + 2] = q.z;
            v[4 * i + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < kItems; i++) {
            uint64_t idx = base + i;
            v[i] = idx < n ? in[idx] : 0u;
        }
    }
#pragma unroll
    for (int i = 0; i < kItems; i++) s += v[i];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t incl = wave_incl_scan(s, lane);
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    uint32_t wpref = 0, btot = 0;
    for (int i = 0; i < kWaves; i++) {
        if (i < w) wpref += ws[i];
        btot += ws[i];
    }
    uint32_t run = block_pref[blockIdx.x] + wpref + incl - s;
    if (full) {
#pragma unroll
        for (int i = 0; i < kItems / 4; i++) {
            uint4 q;
            q.x = run; run += v[4 * i];
            q.y = run; run += v[4 * i + 1];
            q.z = run; run += v[4 * i + 2];
            q.w = run; run += v[4 * i + 3];
            reinterpret_cast<uint4 *>(out + base)[i] = q;
        }
    } else {
#pragma unroll
        for (int i = 0; i < kItems; i++) {
            uint64_t idx = base + i;
            if (idx < n) out[idx] = run;
            run += v[i];
        }
    }
    if (blockIdx.x == nblocks - 1 && threadIdx.x == 0) out[n] = block_pref[blockIdx.x] + btot;
}

// ---- radix sort -----------------------------------------------------------------
// Per-tile digit histogram.  Each quarter of a wave counts into its own LDS
// copy: an atomic add of 64 lanes into few buckets (a 7-bit digit has 128)
// serialises on equal addresses, and the copies split those collisions four
// ways.  Copies are one word longer than the digit range, so that one digit's
// counters in different copies sit in different LDS banks.  With `vec` (keys
// 16-byte aligned) a full tile is read as 16-byte loads, four keys each.
// tmax (may be null) receives the largest key of each tile.
constexpr int kHistCopies = kWaves * 4;
__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint32_t *__restrict__ keys,
                                                       uint32_t n, uint32_t shift,
                                                       uint32_t mask, uint32_t vec,
                                                       uint32_t *__restrict__ hist,
                                                       uint32_t *__restrict__ tmax)
{
    __shared__ uint32_t h[kHistCopies][kRadix + 1];
    __shared__ uint32_t wmax[kWaves];
    const int w = (threadIdx.x >> 6) * 4 + (threadIdx.x & 3);
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < kHistCopies; i++) h[i][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    if (vec && base + kTile <= n) {
        const uint4 *k4 = reinterpret_cast<const uint4 *>(keys + base);
        uint4 q[kItems / 4];
#pragma unroll
        for (int i = 0; i < kItems / 4; i++) q[i] = k4[i * kBlock + threadIdx.x];
#pragma unroll
        for (int i = 0; i < kItems / 4; i++) {
            atomicAdd(&h[w][(q[i].x >> shift) & mask], 1u);
            atomicAdd(&h[w][(q[i].y >> shift) & mask], 1u);
            atomicAdd(&h[w][(q[i].z >> shift) & mask], 1u);
            atomicAdd(&h[w][(q[i].w >> shift) & mask], 1u);
            m = max(m, max(max(q[i].x, q[i].y), max(q[i].z, q[i].w)));
        }
    } else {
#pragma unroll
        for (int i = 0; i < kItems; i++) {
            const uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
            if (idx < n) {
                const uint32_t key = keys[idx];
                atomicAdd(&h[w][(key >> shift) & mask], 1u);
                m = max(m, key);
            }
        }
    }
    if (tmax) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor(m, off, 64));
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    }
    __syncthreads();
    uint32_t t = 0;
#pragma unroll
    for (int i = 0; i < kHistCopies; i++) t += h[i][threadIdx.x];
    hist[(uint64_t)blockIdx.x * kRadix + threadIdx.x] = t;
    if (tmax && threadIdx.x == 0)
        tmax[blockIdx.x] = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
}

// Phase 1: per chunk of `tpc` tiles, per digit sums.
__global__ __launch_bounds__(kRadix) void k_hist_chunk_sums(const uint32_t *__restrict__ hist,
                                                            uint32_t ntiles, uint32_t tpc,
                                                            uint32_t *__restrict__ csum)
{
    const uint32_t d = threadIdx.x;
    const uint32_t t0 = blockIdx.x * tpc;
    const uint32_t t1 = min(ntiles, t0 + tpc);
    uint32_t s = 0;
#pragma unroll 8
    for (uint32_t t = t0; t < t1; t++) s += hist[(uint64_t)t * kRadix + d];
    csum[(uint64_t)blockIdx.x * kRadix + d] = s;
}

// Phase 2 (one block of kScanThreads): exclusive prefix over chunks per
// digit, then over digits.  Thread (part, d) sums chunks [part*64, +64) of
// digit d (32 loads in flight at a time), the four parts' sums and the digit
// totals are combined in LDS, and the thread reads its chunks again to write
// their prefixes.
constexpr int kScanParts = 4, kScanThreads = kScanParts * kRadix, kScanChunks = 64;
static_assert(kScanParts * kScanChunks >= 256, "the chunk scan covers 256 chunks");
__global__ __launch_bounds__(kScanThreads) void k_hist_chunk_scan(uint32_t *__restrict__ csum,
                                                                  uint32_t nchunks)
{
    __shared__ uint32_t psum[kScanParts][kRadix];
    __shared__ uint32_t ws[kRadix / 64];
    const uint32_t d = threadIdx.x % kRadix, part = threadIdx.x / kRadix;
    const uint32_t c0 = part * kScanChunks;
    constexpr int kHalf = kScanChunks / 2; // loads in flight per step (register budget)
    uint32_t tot = 0;
#pragma unroll 1
    for (int h = 0; h < 2; h++) {
        uint32_t x[kHalf];
        const uint32_t cb = c0 + h * kHalf;
#pragma unroll
        for (int i = 0; i < kHalf; i++) x[i] = cb + i < nchunks ? csum[(cb + i) * kRadix + d] : 0u;
#pragma unroll
        for (int i = 0; i < kHalf; i++) tot += x[i];
    }
    psum[part][d] = tot;
    __syncthreads();
    // digit totals and their exclusive scan over digits (the first 256 threads)
    const int lane = d & 63, w = d >> 6;
    uint32_t dtot = 0, incl = 0;
    if (part == 0) {
#pragma unroll
        for (int p = 0; p < kScanParts; p++) dtot += psum[p][d];
        incl = wave_incl_scan(dtot, lane);
        if (lane == 63) ws[w] = incl;
    }
    __syncthreads();
    if (part == 0) {
        uint32_t wpref = 0;
        for (int i = 0; i < w; i++) wpref += ws[i];
        // this digit's first output position, then each part's start
        uint32_t start = wpref + incl - dtot;
#pragma unroll
        for (int p = 0; p < kScanParts; p++) {
            const uint32_t t = psum[p][d];
            psum[p][d] = start;
            start += t;
        }
    }
    __syncthreads();
    uint32_t run = psum[part][d];
#pragma unroll 1
    for (int h = 0; h < 2; h++) {
        uint32_t x[kHalf];
        const uint32_t cb = c0 + h * kHalf;
#pragma unroll
        for (int i = 0; i < kHalf; i++) x[i] = cb + i < nchunks ? csum[(cb + i) * kRadix + d] : 0u;
#pragma unroll
        for (int i = 0; i < kHalf; i++) {
            if (cb + i < nchunks) csum[(cb + i) * kRadix + d] = run;
            run += x[i];
        }
    }
}

// Phase 3: rewrite each tile histogram entry as its global exclusive offset.
__global__ __launch_bounds__(kRadix) void k_hist_chunk_apply(uint32_t *__restrict__ hist,
                                                             uint32_t ntiles, uint32_t tpc,
                                                             const uint32_t *__restrict__ csum)
{
    const uint32_t d = threadIdx.x;
    const uint32_t t0 = blockIdx.x * tpc;
    const uint32_t t1 = min(ntiles, t0 + tpc);
    uint32_t run = csum[(uint64_t)blockIdx.x * kRadix + d];
    // 32 tiles' counts in flight, then their offsets written
    constexpr uint32_t kStep = 32;
    for (uint32_t tb = t0; tb < t1; tb += kStep) {
        uint32_t x[kStep];
#pragma unroll
        for (uint32_t k = 0; k < kStep; k++)
            x[k] = tb + k < t1 ? hist[(uint64_t)(tb + k) * kRadix + d] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < kStep; k++) {
            if (tb + k < t1) hist[(uint64_t)(tb + k) * kRadix + d] = run;
            run += x[k];
        }
    }
}

// Stable scatter of NB-bit digits.  Wave w of a tile owns the 1024 consecutive
// keys [tile*4096 + w*1024, +1024): it loads all of them (16 per lane) before
// ranking any, so that the wave has its whole share of the tile in flight at
// once, then walks them in 16 steps of 64.  In each step the lanes holding
// equal digits are found with NB ballots; the rank of a lane is the wave's
// running count for its digit plus the number of equal-digit lanes below it.
// The lowest lane of each equal-digit group then advances the running count.
// The tile is then reordered by digit in LDS (wave 0 < 1 < 2 < 3 inside a
// digit, i.e. input order) and written out run by run, so that keys of one
// digit leave as consecutive words instead of one scattered word each.
// Workgroups are dispatched round-robin over the 8 XCDs.  Giving each XCD a
// contiguous range of tiles puts the neighbouring runs of one digit (tile t's
// and tile t+1's) into the same L2, where they merge into whole lines before
// they are written back.  A bijection on [0, n): XCD x < n % 8 takes n / 8 + 1
// tiles, the others n / 8.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n)
{
    const uint32_t q = n / kXcds, r = n % kXcds;
    const uint32_t x = b % kXcds, j = b / kXcds;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + j;
}

// With V2 a second value array moves with the first (vin2 -> vout2).
template <uint32_t NB, bool V2>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
    const uint32_t *__restrict__ vin2, uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
    uint32_t *__restrict__ vout2, uint32_t n, uint32_t shift, const uint32_t *__restrict__ hist)
{
    __shared__ uint32_t wcnt[kWaves][kRadix]; // per-wave digit counts, then in-tile positions
    __shared__ uint32_t gbase[kRadix];        // global position of the tile's first key of digit d
    __shared__ uint32_t wsum[kWaves];
    __shared__ uint32_t sk[kTile], sv[kTile];
    __shared__ uint32_t sv2[V2 ? kTile : 1];
    constexpr uint32_t mask = (1u << NB) - 1u;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
    for (int i = 0; i < kWaves; i++) wcnt[i][threadIdx.x] = 0;
    __syncthreads();

    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint64_t tile0 = (uint64_t)tile * kTile;
    const uint64_t base = tile0 + (uint64_t)w * (kTile / kWaves);
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // the wave's keys from a wave-uniform base: lane offsets stay 32-bit
    const uint32_t *kw = kin + base;
    const uint32_t *vw = vin + base;
    const uint32_t *vw2 = V2 ? vin2 + base : nullptr;
    uint32_t k[kItems], v[kItems], r[kItems], v2[V2 ? kItems : 1];
    const bool full = tile0 + kTile <= n;
    if (full) {
#pragma unroll
        for (int i = 0; i < kItems; i++) {
            k[i] = kw[i * 64 + lane];
            v[i] = vw[i * 64 + lane];
            if constexpr (V2) v2[i] = vw2[i * 64 + lane];
        }
    } else {
#pragma unroll
        for (int i = 0; i < kItems; i++) {
            const bool valid = base + (uint64_t)i * 64 + lane < n;
            k[i] = valid ? kw[i * 64 + lane] : 0u;
            v[i] = valid ? vw[i * 64 + lane] : 0u;
            if constexpr (V2) v2[i] = valid ? vw2[i * 64 + lane] : 0u;
        }
    }
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        const bool valid = full || base + (uint64_t)i * 64 + lane < n;
        const uint32_t d = (k[i] >> shift) & mask;
        // peers: the valid lanes whose digit equals this lane's; per bit b a
        // lane keeps the ballot when its bit is set and its complement when
        // not (x is 0 or all ones)
        const uint64_t vb = __ballot(valid);
        uint32_t plo = (uint32_t)vb, phi = (uint32_t)(vb >> 32);
#pragma unroll
        for (uint32_t b = 0; b < NB; b++) {
            const uint32_t bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            const uint32_t x = bit - 1u;
            plo &= (uint32_t)m ^ x;
            phi &= (uint32_t)(m >> 32) ^ x;
        }
        const uint64_t peers = ((uint64_t)phi << 32) | plo;
        uint32_t rank = 0;
        if (valid) {
            const uint32_t before = wcnt[w][d];
            rank = before + (uint32_t)__popcll(peers & lt_mask);
            const int leader = __ffsll((unsigned long long)peers) - 1;
            if (lane == leader) wcnt[w][d] = before + (uint32_t)__popcll(peers);
        }
        r[i] = rank;
    }
    __syncthreads();
    // digit d (one thread each): tile count, then the exclusive scan over
    // digits gives the digit's first in-tile position
    const uint32_t d0 = threadIdx.x;
    uint32_t tot = 0;
#pragma unroll
    for (int i = 0; i < kWaves; i++) tot += wcnt[i][d0];
    const uint32_t incl = wave_incl_scan(tot, lane);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t loc = incl - tot;
    for (int i = 0; i < w; i++) loc += wsum[i];
    gbase[d0] = hist[(uint64_t)tile * kRadix + d0] - loc;
    {
        uint32_t run = loc;
#pragma unroll
        for (int i = 0; i < kWaves; i++) {
            const uint32_t c = wcnt[i][d0];
            wcnt[i][d0] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        const uint64_t idx = base + (uint64_t)i * 64 + lane;
        if (idx < n) {
            const uint32_t lp = wcnt[w][(k[i] >> shift) & mask] + r[i];
            sk[lp] = k[i];
            sv[lp] = v[i];
            if constexpr (V2) sv2[lp] = v2[i];
        }
    }
    __syncthreads();
    // in-tile position j of digit d goes to gbase[d] + j (consecutive per run)
    const uint32_t cnt = (uint32_t)min<uint64_t>(kTile, n - tile0);
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        const uint32_t j = (uint32_t)i * kBlock + threadIdx.x;
        if (j < cnt) {
            const uint32_t key = sk[j];
            const uint32_t pos = gbase[(key >> shift) & mask] + j;
            kout[pos] = key;
            vout[pos] = sv[j];
            if constexpr (V2) vout2[pos] = sv2[j];
        }
    }
}

// ---- vertex mapping -------------------------------------------------------------
// Segment starts -> edges_idx[v]; segment ends -> deg[v] = end (fixed up below).
// Keys >= N (a build whose ids are checked after the sort) are skipped.
__global__ void k_segment_bounds(const uint32_t *__restrict__ keys, uint32_t n, uint32_t N,
                                 uint32_t *__restrict__ idx, uint32_t *__restrict__ end)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = keys[i];
    if (v >= N) return;
    if (i == 0 || keys[i - 1] != v) idx[v] = (uint32_t)i;
    if (i == (uint64_t)n - 1 || keys[i + 1] != v) end[v] = (uint32_t)(i + 1);
}

// The same, four keys per thread from one 16-byte load (keys 16-byte aligned);
// the neighbours across the group come from the adjacent words.
__global__ void k_segment_bounds4(const uint32_t *__restrict__ keys, uint32_t n, uint32_t N,
                                  uint32_t *__restrict__ idx, uint32_t *__restrict__ end)
{
    const uint64_t i0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i0 >= n) return;
    const uint32_t m = (uint32_t)min<uint64_t>(4, n - i0);
    uint32_t k[4];
    if (m == 4) {
        const uint4 q = *reinterpret_cast<const uint4 *>(keys + i0);
        k[0] = q.x; k[1] = q.y; k[2] = q.z; k[3] = q.w;
    } else {
        for (uint32_t j = 0; j < 4; j++) k[j] = j < m ? keys[i0 + j] : 0u;
    }
    const bool has_prev = i0 > 0, has_next = i0 + m < n;
    const uint32_t prev = has_prev ? keys[i0 - 1] : 0u;
    const uint32_t next = has_next ? keys[i0 + m] : 0u;
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) {
        if (j >= m) break;
        const uint32_t v = k[j];
        if (v >= N) continue;
        const bool first = j == 0 ? (!has_prev || prev != v) : k[j - 1] != v;
        const bool last = j + 1 == m ? (!has_next || next != v) : k[j + 1] != v;
        if (first) idx[v] = (uint32_t)(i0 + j);
        if (last) end[v] = (uint32_t)(i0 + j + 1);
    }
}

__global__ void k_segment_degree(const uint32_t *__restrict__ idx, uint32_t *__restrict__ deg,
                                 uint32_t N)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const uint32_t e = deg[v];
    deg[v] = e ? e - idx[v] : 0u;
}

__device__ __forceinline__ uint32_t max4(uint4 q)
{
    return max(max(q.x, q.y), max(q.z, q.w));
}

// vec: a and b (when given) are 16-byte aligned; the first n / 4 * 4 words
// are read four at a time, the rest one at a time.
__global__ __launch_bounds__(kBlock) void k_max2(const uint32_t *__restrict__ a,
                                                 const uint32_t *__restrict__ b, uint32_t n,
                                                 uint32_t vec, uint32_t *__restrict__ out)
{
    uint32_t m = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t n4 = vec ? n / 4 : 0;
    for (uint64_t i = t; i < n4; i += stride) {
        m = max(m, max4(reinterpret_cast<const uint4 *>(a)[i]));
        if (b) m = max(m, max4(reinterpret_cast<const uint4 *>(b)[i]));
    }
    for (uint64_t i = n4 * 4 + t; i < n; i += stride) {
        m = max(m, a[i]);
        if (b) m = max(m, b[i]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor(m, off, 64));
    __shared__ uint32_t wm[kWaves];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kWaves; i++) m = max(m, wm[i]);
        atomicMax(out, m); // one atomic per workgroup
    }
}

} // namespace

// One level of the exclusive scan; block sums of level l live in scratch slot
// kScrScan0 + l (four levels cover 2^48 elements).
static void scan_level(ag_ctx *ctx, const uint32_t *in, uint32_t *out, uint32_t n, int level)
{
    hipStream_t s = ctx->stream;
    if (n == 0) {
        AG_HIP(hipMemsetAsync(out, 0, sizeof(uint32_t), s));
        return;
    }
    if (level > 3) fail(AG_ERR_INVALID, "scan of %u elements is too deep", n);
    const uint32_t nb = ceil_div(n, kTile);
    uint32_t *sums = static_cast<uint32_t *>(scratch(ctx, kScrScan0 + level, (size_t)(nb + 1) * 4));
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, s, in, n, sums);
    AG_HIP(hipGetLastError());
    if (nb == 1) {
        AG_HIP(hipMemsetAsync(sums, 0, sizeof(uint32_t), s));
    } else {
        scan_level(ctx, sums, sums, nb, level + 1);
    }
    const uint32_t vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0;
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kBlock), 0, s, in, out, n, sums, nb, vec);
    AG_HIP(hipGetLastError());
}

void scan_exclusive_u32(ag_ctx *ctx, const uint32_t *in, uint32_t *out, uint32_t n)
{
    std::lock_guard<std::recursive_mutex> lk(ctx->scratch_mu);
    scan_level(ctx, in, out, n, 0);
}

// Digit widths of the LSD passes over key_bits bits: as few passes of at most
// 8 bits as cover the key, the widths within one bit of each other.
static uint32_t digit_passes(uint32_t key_bits) { return (key_bits + 7) / 8; }
static uint32_t digit_shift(uint32_t key_bits, uint32_t p, uint32_t *nbits)
{
    const uint32_t passes = digit_passes(key_bits);
    const uint32_t base = key_bits / passes, extra = key_bits % passes;
    *nbits = base + (p < extra ? 1u : 0u);
    return p * base + std::min(p, extra);
}

struct SortPass {
    ag_ctx *ctx;
    uint32_t n, ntiles, tpc, nchunks;
    uint32_t *hist, *csum;
    SortPass(ag_ctx *c, uint32_t n_) : ctx(c), n(n_)
    {
        ntiles = ceil_div(n, kTile);
        tpc = ceil_div(ntiles, 256);
        nchunks = ceil_div(ntiles, tpc);
        hist = static_cast<uint32_t *>(scratch(ctx, kScrHist, (size_t)ntiles * kRadix * 4));
        csum = static_cast<uint32_t *>(scratch(ctx, kScrCsum, (size_t)nchunks * kRadix * 4));
    }
    // one stable pass: (ck, cv) -> (ok, ov) by the nbits-bit digit at shift;
    // with key_max, also atomicMax(*key_max, largest key)
    void run(const uint32_t *ck, const uint32_t *cv, uint32_t *ok, uint32_t *ov, uint32_t shift,
             uint32_t nbits, uint32_t *key_max = nullptr, const uint32_t *cv2 = nullptr,
             uint32_t *ov2 = nullptr) const
    {
        hipStream_t s = ctx->stream;
        const uint32_t vec = (reinterpret_cast<uintptr_t>(ck) & 15u) == 0;
        uint32_t *tmax = key_max ? scratch_as<uint32_t>(ctx, kScrTileMax, ntiles) : nullptr;
        hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(kBlock), 0, s, ck, n, shift,
                           (1u << nbits) - 1u, vec, hist, tmax);
        if (tmax)
            hipLaunchKernelGGL(k_max2, dim3(std::min<uint32_t>(ceil_div(ntiles, kBlock * 16), 4096)),
                               dim3(kBlock), 0, s, tmax, nullptr, ntiles, 1u, key_max);
        hipLaunchKernelGGL(k_hist_chunk_sums, dim3(nchunks), dim3(kRadix), 0, s, hist, ntiles, tpc,
                           csum);
        hipLaunchKernelGGL(k_hist_chunk_scan, dim3(1), dim3(kScanThreads), 0, s, csum, nchunks);
        hipLaunchKernelGGL(k_hist_chunk_apply, dim3(nchunks), dim3(kRadix), 0, s, hist, ntiles, tpc,
                           csum);
        switch (nbits) {
#define AG_SCATTER(B)                                                                              \
    case B:                                                                                        \
        if (cv2)                                                                                   \
            hipLaunchKernelGGL((k_radix_scatter<B, true>), dim3(ntiles), dim3(kBlock), 0, s, ck,   \
                               cv, cv2, ok, ov, ov2, n, shift, hist);                              \
        else                                                                                       \
            hipLaunchKernelGGL((k_radix_scatter<B, false>), dim3(ntiles), dim3(kBlock), 0, s, ck,  \
                               cv, nullptr, ok, ov, nullptr, n, shift, hist);                      \
        break;
            AG_SCATTER(1) AG_SCATTER(2) AG_SCATTER(3) AG_SCATTER(4)
            AG_SCATTER(5) AG_SCATTER(6) AG_SCATTER(7) AG_SCATTER(8)
#undef AG_SCATTER
        default: fail(AG_ERR_INVALID, "radix digit of %u bits", nbits);
        }
        AG_HIP(hipGetLastError());
    }
};

void radix_sort_pairs(ag_ctx *ctx, const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                      uint32_t *vout, uint32_t n, uint32_t key_bits)
{
    hipStream_t s = ctx->stream;
    if (n == 0) return;
    std::lock_guard<std::recursive_mutex> lk(ctx->scratch_mu);
    const uint32_t passes = digit_passes(key_bits);
    if (passes == 0) {
        AG_HIP(hipMemcpyAsync(kout, kin, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
        AG_HIP(hipMemcpyAsync(vout, vin, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
        return;
    }
    const SortPass sp(ctx, n);
    uint32_t *tk = nullptr, *tv = nullptr;
    if (passes > 1) {
        tk = static_cast<uint32_t *>(scratch(ctx, kScrSortK, (size_t)n * 4));
        tv = static_cast<uint32_t *>(scratch(ctx, kScrSortV, (size_t)n * 4));
    }
    // Ping-pong so that the last pass lands in (kout, vout).
    const uint32_t *ck = kin, *cv = vin;
    for (uint32_t p = 0; p < passes; p++) {
        uint32_t nbits;
        const uint32_t shift = digit_shift(key_bits, p, &nbits);
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint32_t *ok = to_out ? kout : tk;
        uint32_t *ov = to_out ? vout : tv;
        sp.run(ck, cv, ok, ov, shift, nbits);
        ck = ok;
        cv = ov;
    }
}

void radix_sort_triples(ag_ctx *ctx, const uint32_t *kin, const uint32_t *vin, const uint32_t *vin2,
                        uint32_t *kout, uint32_t *vout, uint32_t *vout2, uint32_t n,
                        uint32_t key_bits)
{
    hipStream_t s = ctx->stream;
    if (n == 0) return;
    std::lock_guard<std::recursive_mutex> lk(ctx->scratch_mu);
    const uint32_t passes = digit_passes(key_bits);
    if (passes == 0) {
        AG_HIP(hipMemcpyAsync(kout, kin, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
        AG_HIP(hipMemcpyAsync(vout, vin, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
        AG_HIP(hipMemcpyAsync(vout2, vin2, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
        return;
    }
    const SortPass sp(ctx, n);
    uint32_t *tk = nullptr, *tv = nullptr, *tv2 = nullptr;
    if (passes > 1) {
        tk = static_cast<uint32_t *>(scratch(ctx, kScrSortK, (size_t)n * 4));
        tv = static_cast<uint32_t *>(scratch(ctx, kScrSortV, (size_t)n * 4));
        tv2 = static_cast<uint32_t *>(scratch(ctx, kScrSort2V, (size_t)n * 4));
    }
    const uint32_t *ck = kin, *cv = vin, *cv2 = vin2;
    for (uint32_t p = 0; p < passes; p++) {
        uint32_t nbits;
        const uint32_t shift = digit_shift(key_bits, p, &nbits);
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint32_t *ok = to_out ? kout : tk;
        uint32_t *ov = to_out ? vout : tv;
        uint32_t *ov2 = to_out ? vout2 : tv2;
        sp.run(ck, cv, ok, ov, shift, nbits, nullptr, cv2, ov2);
        ck = ok;
        cv = ov;
        cv2 = ov2;
    }
}

void radix_sort_pairs_late(ag_ctx *ctx, const uint32_t *kin, const uint32_t *vin, uint32_t n,
                           uint32_t key_bits,
                           const std::function<void(uint32_t **kout, uint32_t **vout)> &out,
                           uint32_t *key_max)
{
    std::lock_guard<std::recursive_mutex> lk(ctx->scratch_mu);
    uint32_t *kout = nullptr, *vout = nullptr;
    const uint32_t passes = digit_passes(key_bits);
    if (n == 0 || passes <= 1) {
        if (key_max && n) device_max_async(ctx, kin, nullptr, n, key_max);
        out(&kout, &vout);
        radix_sort_pairs(ctx, kin, vin, kout, vout, n, key_bits);
        return;
    }
    const SortPass sp(ctx, n);
    uint32_t *bk[2] = {static_cast<uint32_t *>(scratch(ctx, kScrSortK, (size_t)n * 4)),
                       static_cast<uint32_t *>(scratch(ctx, kScrSort2K, (size_t)n * 4))};
    uint32_t *bv[2] = {static_cast<uint32_t *>(scratch(ctx, kScrSortV, (size_t)n * 4)),
                       static_cast<uint32_t *>(scratch(ctx, kScrSort2V, (size_t)n * 4))};
    uint32_t nbits;
    uint32_t shift = digit_shift(key_bits, 0, &nbits);
    sp.run(kin, vin, bk[0], bv[0], shift, nbits, key_max);
    out(&kout, &vout); // the caller's allocation overlaps the first pass
    int cur = 0;       // pair holding the last pass's output
    for (uint32_t p = 1; p < passes; p++) {
        shift = digit_shift(key_bits, p, &nbits);
        const bool last = p + 1 == passes;
        uint32_t *ok = last ? kout : bk[1 - cur];
        uint32_t *ov = last ? vout : bv[1 - cur];
        sp.run(bk[cur], bv[cur], ok, ov, shift, nbits);
        cur = 1 - cur;
    }
}

void vertex_map(ag_ctx *ctx, const uint32_t *sorted_keys, uint32_t n, uint32_t N,
                uint32_t *degree, uint32_t *edges_idx, uint32_t *offsets)
{
    hipStream_t s = ctx->stream;
    if (N == 0) {
        if (offsets) AG_HIP(hipMemsetAsync(offsets, 0, 4, s));
        return;
    }
    AG_HIP(hipMemsetAsync(degree, 0, (size_t)N * 4, s));
    AG_HIP(hipMemsetAsync(edges_idx, 0, (size_t)N * 4, s));
    if (n) {
        if ((reinterpret_cast<uintptr_t>(sorted_keys) & 15u) == 0)
            hipLaunchKernelGGL(k_segment_bounds4, dim3(ceil_div(ceil_div(n, 4), 256)), dim3(256), 0,
                               s, sorted_keys, n, N, edges_idx, degree);
        else
            hipLaunchKernelGGL(k_segment_bounds, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                               sorted_keys, n, N, edges_idx, degree);
        hipLaunchKernelGGL(k_segment_degree, dim3(ceil_div(N, 256)), dim3(256), 0, s, edges_idx,
                           degree, N);
        AG_HIP(hipGetLastError());
    }
    if (offsets) scan_exclusive_u32(ctx, degree, offsets, N);
}

void device_max_async(ag_ctx *ctx, const uint32_t *a, const uint32_t *b, uint32_t n,
                      uint32_t *out)
{
    if (n == 0) return;
    const uint32_t blocks = std::min<uint32_t>(ceil_div(n, kBlock * 16), 4096);
    const uint32_t vec = (reinterpret_cast<uintptr_t>(a) & 15u) == 0 &&
                         (!b || (reinterpret_cast<uintptr_t>(b) & 15u) == 0);
    hipLaunchKernelGGL(k_max2, dim3(blocks), dim3(kBlock), 0, ctx->stream, a, b, n, vec, out);
    AG_HIP(hipGetLastError());
}

static uint32_t device_max_impl(ag_ctx *ctx, const uint32_t *a, const uint32_t *b, uint32_t n)
{
    if (n == 0) return 0;
    std::lock_guard<std::recursive_mutex> lk(ctx->scratch_mu);
    uint32_t *out = static_cast<uint32_t *>(scratch(ctx, kScrMax, 4));
    AG_HIP(hipMemsetAsync(out, 0, 4, ctx->stream));
    const uint32_t blocks = std::min<uint32_t>(ceil_div(n, kBlock * 16), 4096);
    const uint32_t vec = (reinterpret_cast<uintptr_t>(a) & 15u) == 0 &&
                         (!b || (reinterpret_cast<uintptr_t>(b) & 15u) == 0);
    hipLaunchKernelGGL(k_max2, dim3(blocks), dim3(kBlock), 0, ctx->stream, a, b, n, vec, out);
    AG_HIP(hipGetLastError());
    uint32_t m = 0;
    AG_HIP(hipMemcpyAsync(&m, out, 4, hipMemcpyDeviceToHost, ctx->stream));
    AG_HIP(hipStreamSynchronize(ctx->stream));
    return m;
}

uint32_t device_max2(ag_ctx *ctx, const uint32_t *a, const uint32_t *b, uint32_t n)
{
    return device_max_impl(ctx, a, b, n);
}
uint32_t device_max1(ag_ctx *ctx, const uint32_t *a, uint32_t n)
{
    return device_max_impl(ctx, a, nullptr, n);
}

} // namespace ag
