// Batched personalized PageRank (multi-right-hand-side pull sweeps).
//
// Column q of a batch is one query with teleport vector s_q (uniform over its
// seed vertices):  x'_q = d * A^T D^-1 x_q + s_q * ((1 - d) + d * dangling(x_q)),
// x_q starting at s_q, stopped when sum |x'_q - x_q| <= tol_q.  The reference
// has no PPR; the gather is the same pull contract as pageRankPullGraphCSR
// (next[v] = sum of contrib[u] over v's inverse edges, pageRank.c:2900-2907),
// applied to every column at once so that one random row access per edge
// serves all columns.
//
// Layout: vertex-major [N][capacity] in the degree-sorted labels of the
// graph's plan: the FP32 iterate X and two ping-pong contribution buffers.
// FP32 storage keeps one FP32 word per column.  BF16 storage keeps the same
// FP32 word split into its high half (a truncated BF16, read as the midpoint
// of its truncation interval, so unbiased) and its low half; a row is cut in
// blocks of up to 32 columns, each block holding its columns' high halves and
// then their low halves, so both halves of a column share one 128-byte line.
// A coarse column reads the high halves only (half the gather bytes); once its
// residual stalls or meets the tolerance it also reads the low halves and
// continues in exact FP32, in place, in the same batch.
//
// A lane holds 16 bytes of one row (4 FP32 or 8 16-bit columns, VW); LPR
// lanes cover the columns of a column group, so one wave gathers 64 / LPR rows
// at a time.  LPR follows the highest used slot.
//   * rows with in-degree above the plan's low_max are cut into chunks of
//     kChunk edges; a chunk is read by kSlots lane groups (slot s takes edges
//     s, s + kSlots, ... in order) whose sums combine as (s0 + s1) + (s2 + s3).
//     A one-chunk row is updated at once; longer rows leave per-chunk partials
//     that k_ppr_rows sums in the same slot order over chunks;
//   * low rows (the plan's SELL-64 slices) take one lane group each and sum
//     their edges in order;
//   * rows without in-edges change only if they are seeds (zrows list).
// Every column's summation order is fixed by the graph alone (not by the batch
// width, the lanes per row or the other columns), so a query's result is the
// same alone and in any batch.  Residual and dangling sums are FP64.  A sweep
// is three launches: sweep, multi-chunk rows, one-workgroup column reduction;
// kernel boundaries order them, with no device-scope fence (an L2 write-back +
// invalidate on gfx950).  Columns retire individually: a done column is frozen
// (no writes), its slot is reused by the next admission, and a draining batch
// packs its columns into the low slots so that the sweeps narrow.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>

#include "ag_extrap.h"
#include "ag_ppr.h"

namespace ag {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kMaxCap = 64;
constexpr uint32_t kSlots = 4;     // lane groups per chunk (fixed summation order)
constexpr uint32_t kChunk = 256;   // edges per chunk of a row above low_max
constexpr int kUnroll = 8;         // gathers in flight per lane
constexpr uint32_t kMid = 0x80008000u; // low halves of two coarse columns: interval midpoints
// A coarse column turns fine once its L1 residual is below this, or stops
// falling within 8x of it: the BF16 rounding noise of a sweep is about
// d * 2^-10 * ||x||_1 ~ 1e-3, so further coarse sweeps would mostly measure
// noise.  (Early sweeps may grow the residual while the mass spreads; that is
// not a stall.)
#ifndef AG_PPR_COARSE_FLOOR
#define AG_PPR_COARSE_FLOOR 0.01
#endif
constexpr double kCoarseFloor = AG_PPR_COARSE_FLOOR;
// Largest relative error of a coarse read (the midpoint of a BF16 truncation
// interval): half a BF16 ulp, 2^-8 of the value.
constexpr double kMidRel = 0x1p-8;
// E * 2^-134 for E < 2^32 edges: the midpoint reads of zero contributions.
constexpr double kMidAbs = 0x1p-102;
// Fewest remaining sweeps (estimated at 0.6x residual per sweep for the
// slowest running column) for which a draining batch packs its columns.
constexpr uint32_t kCompactMinTail = 3;

struct PprArgs {
    uint32_t N, cap;
    uint32_t blk; // BF16 split block (32 or 64 columns)
    uint32_t n_nonlow, n_nonzero;
    uint32_t nch, nmulti;
    float damp;
    const uint32_t *__restrict__ roff;
    const uint32_t *__restrict__ cols;
    const uint32_t *__restrict__ slice_ptr;
    const uint32_t *__restrict__ outdeg;
    const uint32_t *__restrict__ ch_row;   // [nch] row of each chunk
    const uint32_t *__restrict__ ch_first; // [n_nonlow + 1] first chunk of each row
    const uint32_t *__restrict__ multi;    // [nmulti] rows of more than one chunk
    float *X;
    const void *Cin;
    void *Cout;
    float *part; // [nch][cap] chunk sums of multi-chunk rows
    const uint32_t *seed_bits, *seed_v, *seed_c, *zrows;
    const float *seed_w;
    PprCol *cs;
    PprBatchState *bs;
    double *blk_res, *blk_dang; // [partial rows][kMaxCap]
    uint32_t part_base;         // first partial row of this launch
    uint32_t extrap;            // 1: columns may take an extrapolated step
};

// Per-slot descriptor of one admission (or move: e0 = destination slot).
struct PprAdmit {
    uint32_t slot, e0, e1, max_iter;
    double tol;
};

// BF16 storage: byte offsets, within a row of 4 * cap bytes, of the high and
// the low half of column c.  Columns form blocks of up to `blk` (32 or 64):
// the block's high halves, then its low halves.
//   * blk 64: a full row (cap 64) is one 128-byte line of highs followed by
//     one of lows, so a coarse sweep reads one line per gathered row;
//   * blk 32: columns 0-31 keep both halves in one line, so a narrow fine
//     sweep (a drained batch packed into the low slots) reads one line.
// The batch switches between the two as its used slots widen and narrow.
__device__ __forceinline__ uint32_t split_hi(uint32_t cap, uint32_t blk, uint32_t c)
{
    const uint32_t b = c & ~(blk - 1);
    return b * 4 + (c - b) * 2;
}
__device__ __forceinline__ uint32_t split_lo(uint32_t cap, uint32_t blk, uint32_t c)
{
    const uint32_t b = c & ~(blk - 1);
    return split_hi(cap, blk, c) + min(blk, cap - b) * 2;
}

// Scalar store of one contribution (slot management kernels).
__device__ __forceinline__ void st(void *buf, uint32_t fmt, uint32_t cap, uint32_t blk, uint32_t row,
                                   uint32_t col, float c)
{
    if (fmt == AG_PREC_FP32) {
        static_cast<float *>(buf)[(uint64_t)row * cap + col] = c;
    } else {
        const uint32_t b = __float_as_uint(c);
        char *r = static_cast<char *>(buf) + (uint64_t)row * cap * 4;
        *reinterpret_cast<uint16_t *>(r + split_hi(cap, blk, col)) = (uint16_t)(b >> 16);
        *reinterpret_cast<uint16_t *>(r + split_lo(cap, blk, col)) = (uint16_t)(b & 0xffffu);
    }
}

__device__ __forceinline__ uint32_t ldnt(const uint32_t *p) { return __builtin_nontemporal_load(p); }

// ---- FP32 storage: 4 columns per lane ----
struct F32Rows {
    static constexpr int W = 4;
    static constexpr int U = kUnroll;
    const float *__restrict__ c;
    uint32_t cap, c0;
    __device__ __forceinline__ uint4 load(uint32_t u) const
    {
        return *reinterpret_cast<const uint4 *>(c + (uint64_t)u * cap + c0);
    }
    static __device__ __forceinline__ void add(float (&a)[4], const uint4 r)
    {
        a[0] += __uint_as_float(r.x);
        a[1] += __uint_as_float(r.y);
        a[2] += __uint_as_float(r.z);
        a[3] += __uint_as_float(r.w);
    }
};

// ---- BF16 storage: 8 columns per lane, high halves (+ low halves when FINE) ----
template <bool FINE>
struct SplitRows {
    static constexpr int W = 8;
    static constexpr int U = kUnroll;
    const char *__restrict__ base;
    uint32_t stride, hoff, loff; // row bytes; the lane's high / low offsets in a row
    uint32_t fm[4]; // per word: 0xffff on the halves of fine columns
    struct Raw {
        uint4 h, l;
    };
    __device__ __forceinline__ Raw load(uint32_t u) const
    {
        Raw r;
        const char *p = base + (uint64_t)u * stride;
        r.h = *reinterpret_cast<const uint4 *>(p + hoff);
        // a fine wave reads the low halves on every lane (no divergent loads);
        // lanes without fine columns mask them to the midpoints in add()
        r.l = FINE ? *reinterpret_cast<const uint4 *>(p + loff) : make_uint4(kMid, kMid, kMid, kMid);
        return r;
    }
    __device__ __forceinline__ void add2(float &a0, float &a1, uint32_t h, uint32_t l, uint32_t m) const
    {
        const uint32_t le = FINE ? ((l & m) | (kMid & ~m)) : kMid;
        a0 += __uint_as_float((h << 16) | (le & 0xffffu));
        a1 += __uint_as_float((h & 0xffff0000u) | (le >> 16));
    }
    __device__ __forceinline__ void add(float (&a)[8], const Raw &r) const
    {
        add2(a[0], a[1], r.h.x, r.l.x, fm[0]);
        add2(a[2], a[3], r.h.y, r.l.y, fm[1]);
        add2(a[4], a[5], r.h.z, r.l.z, fm[2]);
        add2(a[6], a[7], r.h.w, r.l.w, fm[3]);
    }
};

template <typename R> struct RawOf { using T = uint4; };
template <bool F> struct RawOf<SplitRows<F>> { using T = typename SplitRows<F>::Raw; };

template <typename R>
__device__ __forceinline__ typename RawOf<R>::T zero_raw()
{
    if constexpr (R::W == 4) {
        return make_uint4(0u, 0u, 0u, 0u);
    } else {
        typename RawOf<R>::T r;
        r.h = make_uint4(0u, 0u, 0u, 0u);
        r.l = make_uint4(0u, 0u, 0u, 0u);
        return r;
    }
}

template <typename R>
__device__ __forceinline__ void add_raw(const R &rows, float (&acc)[R::W], const typename RawOf<R>::T &r)
{
    if constexpr (R::W == 4) R::add(acc, r);
    else rows.add(acc, r);
}

// acc += contributions of sources p[0], p[stride], ..., p[(n-1)*stride], in
// that order, for the lane's columns.  Indices of the next group are loaded
// while the current group's rows are in flight.
template <typename R>
__device__ __forceinline__ void gather_seq(const R &rows, const uint32_t *__restrict__ p, uint32_t stride,
                                           uint32_t n, float (&acc)[R::W])
{
    using Raw = typename RawOf<R>::T;
    uint32_t k = 0;
    if (n >= (uint32_t)R::U) {
        uint32_t u[R::U];
#pragma unroll
        for (int i = 0; i < R::U; i++) u[i] = ldnt(p + (uint64_t)i * stride);
        for (; k + R::U <= n; k += R::U) {
            Raw r[R::U];
#pragma unroll
            for (int i = 0; i < R::U; i++) r[i] = rows.load(u[i]);
            if (k + 2 * R::U <= n) {
#pragma unroll
                for (int i = 0; i < R::U; i++) u[i] = ldnt(p + (uint64_t)(k + R::U + i) * stride);
            }
#pragma unroll
            for (int i = 0; i < R::U; i++) add_raw(rows, acc, r[i]);
        }
    }
    if (k < n) {
        // tail: masked loads; FP32 adds zero, BF16 adds 0x00008000 (4.6e-41)
        // per missing edge, below every value the sums can hold
        Raw r[R::U];
#pragma unroll
        for (int i = 0; i < R::U; i++) {
            r[i] = zero_raw<R>();
            if (k + i < n) r[i] = rows.load(ldnt(p + (uint64_t)(k + i) * stride));
        }
#pragma unroll
        for (int i = 0; i < R::U; i++) {
            if constexpr (R::W == 4) {
                add_raw(rows, acc, r[i]);
            } else {
                if (k + i < n) add_raw(rows, acc, r[i]);
            }
        }
    }
}

template <typename CT> struct Store;
template <> struct Store<float> {
    static constexpr int W = 4;
    static __device__ __forceinline__ void put(const PprArgs &a, uint32_t, uint32_t, uint64_t base,
                                               const float (&c)[4])
    {
        *reinterpret_cast<float4 *>(static_cast<float *>(a.Cout) + base) = make_float4(c[0], c[1], c[2], c[3]);
    }
};
template <> struct Store<bf16_t> {
    static constexpr int W = 8;
    static __device__ __forceinline__ void put(const PprArgs &a, uint32_t row, uint32_t c0, uint64_t,
                                               const float (&c)[8])
    {
        uint32_t h[4], l[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t b0 = __float_as_uint(c[2 * j]), b1 = __float_as_uint(c[2 * j + 1]);
            h[j] = (b0 >> 16) | (b1 & 0xffff0000u);
            l[j] = (b0 & 0xffffu) | (b1 << 16);
        }
        char *r = static_cast<char *>(a.Cout) + (uint64_t)row * a.cap * 4;
        *reinterpret_cast<uint4 *>(r + split_hi(a.cap, a.blk, c0)) = make_uint4(h[0], h[1], h[2], h[3]);
        *reinterpret_cast<uint4 *>(r + split_lo(a.cap, a.blk, c0)) = make_uint4(l[0], l[1], l[2], l[3]);
    }
};

// Column update of label `row` for the lane's VW columns (amask: active ones).
// res / dang: the lane's FP64 accumulators in LDS, column i at [i * kBlock].
template <typename CT>
__device__ __forceinline__ void update_row(const PprArgs &a, uint32_t row, uint32_t c0, uint32_t amask,
                                           const float (&acc)[Store<CT>::W], const float (&t)[Store<CT>::W],
                                           const float (&ex)[Store<CT>::W], double *res, double *dang)
{
#pragma clang fp contract(off)
    constexpr int VW = Store<CT>::W;
    float sv[VW];
#pragma unroll
    for (int i = 0; i < VW; i++) sv[i] = 0.0f;
    if ((a.seed_bits[row >> 5] >> (row & 31)) & 1u) {
        const uint32_t ns = a.bs->nseeds;
        for (uint32_t q = 0; q < ns; q++) {
            if (a.seed_v[q] != row) continue;
            const uint32_t c = a.seed_c[q];
            const float wq = a.seed_w[q];
#pragma unroll
            for (int i = 0; i < VW; i++)
                if (c == c0 + i) sv[i] += wq;
        }
    }
    const uint64_t base = (uint64_t)row * a.cap + c0;
    float x[VW];
#pragma unroll
    for (int i = 0; i < VW; i += 4) {
        const float4 v = *reinterpret_cast<const float4 *>(a.X + base + i);
        x[i] = v.x;
        x[i + 1] = v.y;
        x[i + 2] = v.z;
        x[i + 3] = v.w;
    }
    const uint32_t od = a.outdeg[row];
#pragma unroll
    for (int i = 0; i < VW; i++) {
        if (!((amask >> i) & 1u)) continue;
        const float gathered = a.damp * acc[i];
        const float nw = gathered + sv[i] * t[i];
        res[i * kBlock] += fabs((double)nw - (double)x[i]);
        // extrapolated step (w > 0): nw + w (nw - x) removes the error mode
        // that shrinks by rho = w / (1 + w) per sweep
        const float xo = ex[i] != 0.0f ? nw + ex[i] * (nw - x[i]) : nw;
        if (od == 0) dang[i * kBlock] += (double)xo;
        x[i] = xo;
    }
#pragma unroll
    for (int i = 0; i < VW; i += 4)
        *reinterpret_cast<float4 *>(a.X + base + i) = make_float4(x[i], x[i + 1], x[i + 2], x[i + 3]);
    // a row without out-edges is never gathered: its contribution stays 0
    if (od) {
        float c[VW];
#pragma unroll
        for (int i = 0; i < VW; i++) c[i] = x[i] / (float)od;
        Store<CT>::put(a, row, c0, base, c);
    }
}

// The lane's columns in a column group: active and fine masks, teleport factors.
template <int VW>
__device__ __forceinline__ uint32_t column_state(const PprArgs &a, uint32_t K, uint32_t c0, float (&t)[VW],
                                                 float (&ex)[VW], uint32_t &fmask)
{
    uint32_t amask = 0;
    fmask = 0;
#pragma unroll
    for (int i = 0; i < VW; i++) {
        t[i] = 0.0f;
        ex[i] = 0.0f;
        const uint32_t c = c0 + i;
        if (c < K) {
            const PprCol &col = a.cs[c];
            if (!col.done) {
                amask |= 1u << i;
                if (!col.coarse) fmask |= 1u << i;
                t[i] = col.t;
                ex[i] = col.w;
            }
        }
    }
    return amask;
}

// Per-column block sums of the lanes' FP64 partials in a fixed order (lanes of
// equal column, then waves), into s_res / s_dang at the group's columns.
template <int LPR, int VW>
__device__ __forceinline__ void flush_columns(uint32_t cg, const double *lres, const double *ldang,
                                              double (*w_res)[LPR * VW], double (*w_dang)[LPR * VW],
                                              double *s_res, double *s_dang)
{
    constexpr int CPW = LPR * VW;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double res[VW], dang[VW];
#pragma unroll
    for (int i = 0; i < VW; i++) {
        res[i] = lres[i * kBlock];
        dang[i] = ldang[i * kBlock];
    }
#pragma unroll
    for (int i = 0; i < VW; i++) {
#pragma unroll
        for (int off = LPR; off < 64; off <<= 1) {
            res[i] += __shfl_xor(res[i], off, 64);
            dang[i] += __shfl_xor(dang[i], off, 64);
        }
    }
    if (lane < LPR) {
#pragma unroll
        for (int i = 0; i < VW; i++) {
            w_res[w][lane * VW + i] = res[i];
            w_dang[w][lane * VW + i] = dang[i];
        }
    }
    __syncthreads();
    if (threadIdx.x < CPW) {
        const uint32_t c = cg * CPW + threadIdx.x;
        if (c < kMaxCap) {
            double r = 0.0, d = 0.0;
#pragma unroll
            for (int k = 0; k < kWaves; k++) {
                r += w_res[k][threadIdx.x];
                d += w_dang[k][threadIdx.x];
            }
            s_res[c] = r;
            s_dang[c] = d;
        }
    }
    __syncthreads();
}

// Work units of one column group for one lane: gathers through `rows`.
template <int LPR, typename CT, typename R>
__device__ __forceinline__ void sweep_units(const PprArgs &a, const R &rows, uint32_t c0, uint32_t amask,
                                            const float (&t)[R::W], const float (&ex)[R::W], double *res,
                                            double *dang, uint32_t nz)
{
    constexpr int VW = R::W;
    constexpr int RGL = 64 / LPR;    // low rows per wave
    constexpr int RG = RGL / kSlots; // chunks per wave
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int sub = lane / LPR;
    const int slot = sub % kSlots, rg = sub / kSlots;
    const uint32_t wu_ch = (a.nch + RG - 1) / RG;
    const uint32_t wu_low = (a.n_nonzero - a.n_nonlow + RGL - 1) / RGL;
    const uint32_t total = wu_ch + wu_low + (nz + RGL - 1) / RGL;
    const uint32_t gw = blockIdx.x * kWaves + w, nwav = gridDim.x * kWaves;
    for (uint32_t wu = gw; wu < total; wu += nwav) {
        if (wu < wu_ch) {
            const uint32_t k = wu * RG + rg;
            if (k < a.nch) {
                const uint32_t row = a.ch_row[k];
                const uint32_t f0 = a.ch_first[row];
                const uint32_t b0 = a.roff[row] + (k - f0) * kChunk;
                const uint32_t b1 = min(a.roff[row + 1], b0 + kChunk);
                float acc[VW];
#pragma unroll
                for (int i = 0; i < VW; i++) acc[i] = 0.0f;
                if (amask) {
                    const uint32_t j = b0 + slot;
                    const uint32_t n = j < b1 ? (b1 - j + kSlots - 1) / kSlots : 0u;
                    gather_seq(rows, a.cols + j, kSlots, n, acc);
                }
                // (s0 + s1) + (s2 + s3); partner lanes share row and columns
#pragma unroll
                for (int i = 0; i < VW; i++) acc[i] += __shfl_xor(acc[i], LPR, 64);
#pragma unroll
                for (int i = 0; i < VW; i++) acc[i] += __shfl_xor(acc[i], 2 * LPR, 64);
                if (slot == 0 && amask) {
                    if (a.ch_first[row + 1] - f0 == 1) {
                        update_row<CT>(a, row, c0, amask, acc, t, ex, res, dang);
                    } else {
                        float *pp = a.part + (uint64_t)k * a.cap + c0;
#pragma unroll
                        for (int i = 0; i < VW; i += 4)
                            *reinterpret_cast<float4 *>(pp + i) =
                                make_float4(acc[i], acc[i + 1], acc[i + 2], acc[i + 3]);
                    }
                }
            }
        } else if (wu < wu_ch + wu_low) {
            const uint32_t rl = (wu - wu_ch) * RGL + sub;
            const uint32_t row = a.n_nonlow + rl;
            if (row < a.n_nonzero && amask) {
                const uint32_t deg = a.roff[row + 1] - a.roff[row];
                float acc[VW];
#pragma unroll
                for (int i = 0; i < VW; i++) acc[i] = 0.0f;
                gather_seq(rows, a.cols + a.slice_ptr[rl >> 6] + (rl & 63), 64u, deg, acc);
                update_row<CT>(a, row, c0, amask, acc, t, ex, res, dang);
            }
        } else {
            // seeds without in-edges: x' = s_v * t
            const uint32_t zi = (wu - wu_ch - wu_low) * RGL + sub;
            if (zi < nz && amask) {
                float acc[VW];
#pragma unroll
                for (int i = 0; i < VW; i++) acc[i] = 0.0f;
                update_row<CT>(a, a.zrows[zi], c0, amask, acc, t, ex, res, dang);
            }
        }
    }
}

template <int LPR, typename CT>
__global__ __launch_bounds__(kBlock) void k_ppr_sweep(PprArgs a)
{
    constexpr int VW = Store<CT>::W;
    constexpr int CPW = LPR * VW;
    static_assert(64 / LPR / kSlots >= 1, "a chunk needs kSlots lane groups");
    __shared__ double w_res[kWaves][CPW], w_dang[kWaves][CPW];
    __shared__ double s_res[kMaxCap], s_dang[kMaxCap];
    __shared__ double l_res[VW][kBlock], l_dang[VW][kBlock]; // per-lane column sums
    const uint32_t K = a.bs->K;
    if (K == 0) return;
    const int lane = threadIdx.x & 63;
    const int l = lane % LPR;
    if (threadIdx.x < kMaxCap) {
        s_res[threadIdx.x] = 0.0;
        s_dang[threadIdx.x] = 0.0;
    }
    __syncthreads();
    const uint32_t ncg = (K + CPW - 1) / CPW;
    const uint32_t nz = a.bs->nzrows;
    for (uint32_t cg = 0; cg < ncg; cg++) {
        const uint32_t c0 = cg * CPW + l * VW;
        float t[VW], ex[VW];
        uint32_t fmask;
        const uint32_t amask = column_state<VW>(a, K, c0, t, ex, fmask);
        double *res = &l_res[0][threadIdx.x], *dang = &l_dang[0][threadIdx.x];
#pragma unroll
        for (int i = 0; i < VW; i++) { res[i * kBlock] = 0.0; dang[i * kBlock] = 0.0; }
        if constexpr (VW == 4) {
            F32Rows rows{static_cast<const float *>(a.Cin), a.cap, c0};
            sweep_units<LPR, CT>(a, rows, c0, amask, t, ex, res, dang, nz);
        } else {
            // the low halves are read by waves holding a fine column (wave-uniform path)
            const bool wave_fine = __builtin_amdgcn_ballot_w64((fmask & amask) != 0) != 0;
            if (wave_fine) {
                SplitRows<true> rows;
                rows.base = static_cast<const char *>(a.Cin);
                rows.stride = a.cap * 4;
                rows.hoff = split_hi(a.cap, a.blk, c0);
                rows.loff = split_lo(a.cap, a.blk, c0);
                const uint32_t fa = fmask & amask;
#pragma unroll
                for (int j = 0; j < 4; j++)
                    rows.fm[j] = (((fa >> (2 * j)) & 1u) ? 0x0000ffffu : 0u) |
                                 (((fa >> (2 * j + 1)) & 1u) ? 0xffff0000u : 0u);
                sweep_units<LPR, CT>(a, rows, c0, amask, t, ex, res, dang, nz);
            } else {
                SplitRows<false> rows;
                rows.base = static_cast<const char *>(a.Cin);
                rows.stride = a.cap * 4;
                rows.hoff = split_hi(a.cap, a.blk, c0);
                rows.loff = split_lo(a.cap, a.blk, c0);
#pragma unroll
                for (int j = 0; j < 4; j++) rows.fm[j] = 0u;
                sweep_units<LPR, CT>(a, rows, c0, amask, t, ex, res, dang, nz);
            }
        }
        flush_columns<LPR, VW>(cg, res, dang, w_res, w_dang, s_res, s_dang);
    }
    if (threadIdx.x < kMaxCap) {
        const uint64_t o = (uint64_t)(a.part_base + blockIdx.x) * kMaxCap + threadIdx.x;
        a.blk_res[o] = s_res[threadIdx.x];
        a.blk_dang[o] = s_dang[threadIdx.x];
    }
}

// Rows of more than one chunk, after the sweep (the kernel boundary orders the
// chunk partials): kSlots lane groups per row, slot s sums chunks s, s +
// kSlots, ... in order, combined as (s0 + s1) + (s2 + s3); then the update.
template <int LPR, typename CT>
__global__ __launch_bounds__(kBlock) void k_ppr_rows(PprArgs a)
{
    constexpr int VW = Store<CT>::W;
    constexpr int CPW = LPR * VW;
    constexpr int RG = 64 / LPR / kSlots;
    __shared__ double w_res[kWaves][CPW], w_dang[kWaves][CPW];
    __shared__ double s_res[kMaxCap], s_dang[kMaxCap];
    __shared__ double l_res[VW][kBlock], l_dang[VW][kBlock]; // per-lane column sums
    const uint32_t K = a.bs->K;
    if (K == 0) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int l = lane % LPR, sub = lane / LPR;
    const int slot = sub % kSlots, rg = sub / kSlots;
    if (threadIdx.x < kMaxCap) {
        s_res[threadIdx.x] = 0.0;
        s_dang[threadIdx.x] = 0.0;
    }
    __syncthreads();
    const uint32_t ncg = (K + CPW - 1) / CPW;
    const uint32_t total = (a.nmulti + RG - 1) / RG;
    const uint32_t gw = blockIdx.x * kWaves + w, nwav = gridDim.x * kWaves;
    for (uint32_t cg = 0; cg < ncg; cg++) {
        const uint32_t c0 = cg * CPW + l * VW;
        float t[VW], ex[VW];
        uint32_t fmask;
        const uint32_t amask = column_state<VW>(a, K, c0, t, ex, fmask);
        double *res = &l_res[0][threadIdx.x], *dang = &l_dang[0][threadIdx.x];
#pragma unroll
        for (int i = 0; i < VW; i++) { res[i * kBlock] = 0.0; dang[i * kBlock] = 0.0; }
        for (uint32_t wu = gw; wu < total; wu += nwav) {
            const uint32_t idx = wu * RG + rg;
            if (idx < a.nmulti) {
                const uint32_t row = a.multi[idx];
                const uint32_t f0 = a.ch_first[row], f1 = a.ch_first[row + 1];
                float acc[VW];
#pragma unroll
                for (int i = 0; i < VW; i++) acc[i] = 0.0f;
                if (amask) {
                    // RU chunk partials per step, 16 float4 loads in flight per
                    // lane; sums stay in chunk order.
                    constexpr int RU = 64 / VW;
                    uint32_t k = f0 + slot;
                    for (; k + (RU - 1) * kSlots < f1; k += RU * kSlots) {
                        float4 v[RU][VW / 4];
#pragma unroll
                        for (int e = 0; e < RU; e++)
#pragma unroll
                            for (int i = 0; i < VW / 4; i++)
                                v[e][i] = *reinterpret_cast<const float4 *>(
                                    a.part + (uint64_t)(k + e * kSlots) * a.cap + c0 + 4 * i);
#pragma unroll
                        for (int e = 0; e < RU; e++)
#pragma unroll
                            for (int i = 0; i < VW / 4; i++) {
                                acc[4 * i] += v[e][i].x;
                                acc[4 * i + 1] += v[e][i].y;
                                acc[4 * i + 2] += v[e][i].z;
                                acc[4 * i + 3] += v[e][i].w;
                            }
                    }
                    for (; k < f1; k += kSlots) {
#pragma unroll
                        for (int i = 0; i < VW / 4; i++) {
                            const float4 v = *reinterpret_cast<const float4 *>(
                                a.part + (uint64_t)k * a.cap + c0 + 4 * i);
                            acc[4 * i] += v.x;
                            acc[4 * i + 1] += v.y;
                            acc[4 * i + 2] += v.z;
                            acc[4 * i + 3] += v.w;
                        }
                    }
                }
#pragma unroll
                for (int i = 0; i < VW; i++) acc[i] += __shfl_xor(acc[i], LPR, 64);
#pragma unroll
                for (int i = 0; i < VW; i++) acc[i] += __shfl_xor(acc[i], 2 * LPR, 64);
                if (slot == 0 && amask) update_row<CT>(a, row, c0, amask, acc, t, ex, res, dang);
            }
        }
        flush_columns<LPR, VW>(cg, res, dang, w_res, w_dang, s_res, s_dang);
    }
    if (threadIdx.x < kMaxCap) {
        const uint64_t o = (uint64_t)(a.part_base + blockIdx.x) * kMaxCap + threadIdx.x;
        a.blk_res[o] = s_res[threadIdx.x];
        a.blk_dang[o] = s_dang[threadIdx.x];
    }
}

constexpr int kRedBlock = 256;

// Closes one sweep of a running column: residual R of the iterate it read,
// dangling mass D of the iterate it stored (k_ppr_reduce and the admission
// push share it).
__device__ void close_sweep(PprCol &col, double R, double D, float damp, uint32_t extrap)
{
    const double prev = col.residual;
    const double d_read = col.dangling; // dangling mass of the iterate this sweep read
    // R is |T(x) - x| for the x this sweep read; after an extrapolated
    // step the stored iterate is not T(x), so R certifies nothing and
    // the column cannot stop on it
    const bool stepped = col.w != 0.0f;
    col.w = 0.0f;
    col.iters += 1;
    col.residual = R;
    col.dangling = D;
    col.t = (float)((1.0 - (double)damp) + (double)damp * D);
    if (col.iters >= col.max_iter) {
        col.done = 1;
    } else if (col.coarse && !stepped && (!extrap || col.iters <= 4) &&
               R + kMidRel * fmax(0.0, 1.000001 - d_read) + kMidAbs <= col.tol) {
        // Certified on a coarse sweep.  A midpoint read is within 2^-8
        // of each contribution plus 2^-134 (the midpoint read of a
        // zero), so the coarse map T_c is within delta = d (2^-8 M +
        // E 2^-134) of T in L1, M the mass of x on rows with out-edges
        // (1 - dangling mass while x >= 0: no extrapolated step yet).
        // With R + delta / d <= tol the returned T_c(x) satisfies
        // ||T_c(x) - x*||_1 <= d / (1 - d) tol, the bound of a plain
        // stop.  It fires when x sits on rows without out-edges
        // (M ~ 0), e.g. a query whose seeds are isolated, which would
        // otherwise force one FP32-read sweep.
        col.done = 1;
    } else if (col.coarse && (R <= col.tol || R <= kCoarseFloor ||
                               (R > 0.97 * prev && R <= 8.0 * kCoarseFloor))) {
        col.coarse = 0; // storage floor reached: read full FP32 from now on
        col.fine_at = col.iters;
    } else if (R <= col.tol && !stepped) {
        col.done = 1;
    }
    if (!col.done && extrap) col.w = extrapolation_weight(R, prev, stepped, col.since, col.rho);
}

// One workgroup per column slot: residual and dangling sums over the partial
// rows of both kernels in a fixed order (thread t takes rows t, t + 256, ...,
// then a fixed tree), then convergence.  A coarse (BF16) column whose residual
// stalls or meets its tolerance turns fine: from the next sweep on it reads
// exact FP32 contributions, and it is done at the tolerance there.
__global__ __launch_bounds__(kRedBlock) void k_ppr_reduce(PprArgs a, uint32_t nparts)
{
    __shared__ double w_r[kRedBlock / 64], w_d[kRedBlock / 64];
    const uint32_t K = a.bs->K;
    const uint32_t c = blockIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0 && K) a.bs->sweeps += 1;
    if (c >= K) return;
    double r = 0.0, d = 0.0;
    for (uint32_t b = threadIdx.x; b < nparts; b += kRedBlock) {
        r += a.blk_res[(uint64_t)b * kMaxCap + c];
        d += a.blk_dang[(uint64_t)b * kMaxCap + c];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        r += __shfl_xor(r, off, 64);
        d += __shfl_xor(d, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        w_r[threadIdx.x >> 6] = r;
        w_d[threadIdx.x >> 6] = d;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        PprCol &col = a.cs[c];
        if (!col.done) {
            const double R = (w_r[0] + w_r[1]) + (w_r[2] + w_r[3]);
            const double D = (w_d[0] + w_d[1]) + (w_d[2] + w_d[3]);
            close_sweep(col, R, D, a.damp, a.extrap);
        }
    }
}

// ---- slot management kernels ----
// Column (re)initialisation of the slots admitted in one step: X and both
// contribution buffers cleared, one thread per (row, admitted slot).
__global__ void k_ppr_clear_cols(float *X, void *C0, void *C1, uint32_t N, uint32_t cap,
                                 uint32_t blk, const PprAdmit *adm, uint32_t na, uint32_t fmt)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)N * na) return;
    const uint32_t row = (uint32_t)(i / na), q = (uint32_t)(i % na);
    const uint64_t k = (uint64_t)row * cap + adm[q].slot;
    X[k] = 0.0f;
    st(C0, fmt, cap, blk, row, adm[q].slot, 0.0f);
    st(C1, fmt, cap, blk, row, adm[q].slot, 0.0f);
}

// Initial column state of an admission: x0 = s, teleport factor from the
// dangling seed mass d0.
__device__ PprCol fresh_column(const PprAdmit &d, double d0, float damp, uint32_t fmt)
{
    PprCol col;
    col.t = (float)((1.0 - (double)damp) + (double)damp * d0);
    col.done = 0;
    col.iters = 0;
    col.max_iter = d.max_iter;
    col.tol = d.tol;
    col.residual = INFINITY;
    col.dangling = d0;
    col.coarse = fmt != AG_PREC_FP32 ? 1u : 0u;
    col.fine_at = 0;
    col.w = 0.0f;
    col.since = 0;
    col.rho = 0.0;
    return col;
}

// One workgroup per admitted slot: x0 = s on the seed rows, contributions
// s / outdeg in the buffer the next sweep reads, column state.
__global__ void k_ppr_seed_cols(float *X, void *Cin, const uint32_t *outdeg, uint32_t cap,
                                uint32_t blk, uint32_t fmt, const PprAdmit *adm, const uint32_t *sv,
                                const float *sw, PprCol *cs, float damp)
{
    const PprAdmit d = adm[blockIdx.x];
    for (uint32_t q = d.e0 + threadIdx.x; q < d.e1; q += blockDim.x) {
        const uint32_t v = sv[q];
        const uint64_t k = (uint64_t)v * cap + d.slot;
        X[k] = sw[q];
        const uint32_t od = outdeg[v];
        st(Cin, fmt, cap, blk, v, d.slot, od ? sw[q] / (float)od : 0.0f);
    }
    if (threadIdx.x == 0) {
        double d0 = 0.0;
        for (uint32_t q = d.e0; q < d.e1; q++)
            if (outdeg[sv[q]] == 0) d0 += (double)sw[q];
        cs[d.slot] = fresh_column(d, d0, damp, fmt);
    }
}

// First sweep of the admitted columns by push, one workgroup per admitted
// slot (X and both contribution buffers of the slot cleared).  x1 = T(x0) for
// x0 = s is nonzero only on the seeds and their out-neighbours, so walking the
// seeds' out-edges (forward CSR) replaces a sweep over all E in-edges:
//  1. acc[v] = sum of s_u / outdeg(u) over the edges u -> v, seeds in table
//     order with a barrier between seeds.  The adds of one seed carry equal
//     values, so every sum is formed in one order, fixed by the query: the
//     column's iterate does not depend on its batch.
//  2. Each touched row is claimed once through the sign bit of its X word:
//     x1 = d acc + s_v t, contribution x1 / outdeg, and the residual and
//     dangling sums in 2^-61 fixed point (integer adds, hence independent of
//     the claim order).  The claimed word holds -x1 until step 3.
//  3. Sign bits cleared; the column state closes sweep 1 as k_ppr_reduce does.
// The contributions are exact FP32 (no midpoint reads), for BF16 columns too.
constexpr int kPushBlock = 1024;
// A step per seed (~10 us with its barrier): queries of more seeds take the
// pull first sweep, whose cost does not grow with the seed count.  The choice
// depends on the query alone, so a column's result still does not depend on
// its batch.
constexpr uint32_t kPushMaxSeeds = 64;
constexpr double kFix = 0x1p61; // sums <= 2 (residual) and <= 1 (dangling) fit in 64 bits

__device__ __forceinline__ void push_claim(float *X, void *Cin, const uint32_t *outdeg, uint32_t cap,
                                           uint32_t blk, uint32_t fmt, uint32_t slot, uint32_t v,
                                           float s, float t, float damp, unsigned long long *s_r,
                                           unsigned long long *s_d)
{
    uint32_t *word = reinterpret_cast<uint32_t *>(X) + ((uint64_t)v * cap + slot);
    const uint32_t old = atomicOr(word, 0x80000000u);
    if (old & 0x80000000u) return;
    const float nw = damp * __uint_as_float(old) + s * t;
    atomicExch(word, __float_as_uint(nw) | 0x80000000u);
    const uint32_t od = outdeg[v];
    if (od) st(Cin, fmt, cap, blk, v, slot, nw / (float)od);
    atomicAdd(s_r, (unsigned long long)(fabs((double)nw - (double)s) * kFix));
    if (od == 0) atomicAdd(s_d, (unsigned long long)((double)nw * kFix));
}

__global__ __launch_bounds__(kPushBlock) void k_ppr_admit_push(
    float *X, void *Cin, const uint32_t *outdeg, const uint32_t *order, const uint32_t *rank,
    const uint32_t *fwd_off, const uint32_t *fwd_dst, uint32_t cap, uint32_t blk, uint32_t fmt,
    const PprAdmit *adm, const uint32_t *sv, const float *sw, PprCol *cs, float damp, uint32_t extrap)
{
    __shared__ unsigned long long s_r, s_d;
    __shared__ float s_t;
    __shared__ double s_d0;
    const PprAdmit d = adm[blockIdx.x];
    const uint32_t slot = d.slot;
    if (threadIdx.x == 0) {
        double d0 = 0.0;
        for (uint32_t q = d.e0; q < d.e1; q++)
            if (outdeg[sv[q]] == 0) d0 += (double)sw[q];
        s_d0 = d0;
        s_t = (float)((1.0 - (double)damp) + (double)damp * d0);
        s_r = 0;
        s_d = 0;
    }
    __syncthreads();
    // 1. gathered sums, one seed at a time
    for (uint32_t q = d.e0; q < d.e1; q++) {
        const uint32_t u = sv[q];
        const uint32_t od = outdeg[u];
        if (od == 0) continue; // uniform: every thread reads the same seed
        const float c = sw[q] / (float)od;
        const uint32_t vid = order[u];
        const uint32_t j1 = fwd_off[vid + 1];
        for (uint32_t j = fwd_off[vid] + threadIdx.x; j < j1; j += kPushBlock)
            atomicAdd(&X[(uint64_t)rank[fwd_dst[j]] * cap + slot], c);
        __threadfence();
        __syncthreads();
    }
    // 2. seeds (their s_v), then the other touched rows (s_v = 0)
    const float t = s_t;
    for (uint32_t q = d.e0 + threadIdx.x; q < d.e1; q += kPushBlock)
        push_claim(X, Cin, outdeg, cap, blk, fmt, slot, sv[q], sw[q], t, damp, &s_r, &s_d);
    __threadfence();
    __syncthreads();
    for (uint32_t q = d.e0; q < d.e1; q++) {
        const uint32_t vid = order[sv[q]];
        const uint32_t j1 = fwd_off[vid + 1];
        for (uint32_t j = fwd_off[vid] + threadIdx.x; j < j1; j += kPushBlock)
            push_claim(X, Cin, outdeg, cap, blk, fmt, slot, rank[fwd_dst[j]], 0.0f, t, damp, &s_r, &s_d);
    }
    __threadfence();
    __syncthreads();
    // 3. sign bits cleared
    for (uint32_t q = d.e0 + threadIdx.x; q < d.e1; q += kPushBlock)
        atomicAnd(reinterpret_cast<uint32_t *>(X) + ((uint64_t)sv[q] * cap + slot), 0x7fffffffu);
    for (uint32_t q = d.e0; q < d.e1; q++) {
        const uint32_t vid = order[sv[q]];
        const uint32_t j1 = fwd_off[vid + 1];
        for (uint32_t j = fwd_off[vid] + threadIdx.x; j < j1; j += kPushBlock)
            atomicAnd(reinterpret_cast<uint32_t *>(X) + ((uint64_t)rank[fwd_dst[j]] * cap + slot),
                      0x7fffffffu);
    }
    if (threadIdx.x == 0) {
        PprCol col = fresh_column(d, s_d0, damp, fmt);
        close_sweep(col, (double)s_r / kFix, (double)s_d / kFix, damp, extrap);
        cs[slot] = col;
    }
}

// Stops columns released while running (a cancelled query).
__global__ void k_ppr_stop_cols(PprCol *cs, const PprAdmit *adm, uint32_t n)
{
    const uint32_t q = threadIdx.x;
    if (q < n) cs[adm[q].slot].done = 1;
}


// Re-lays every row of a split-BF16 buffer from `from`-column blocks to
// `to`-column blocks, in place: a workgroup stages kRelayRows rows in LDS,
// then writes each 32-bit word of the new layout from its two halves.
constexpr uint32_t kRelayBlock = 256, kRewriteBytes = 16384;
// Packs running columns into lower slots of the same batch: iterate and both
// contribution buffers (both halves) copied, so rows a sweep never rewrites
// (no in-edges, not a seed) keep the moved column's values, not those of the
// slot's previous column.  adm[q].slot = source, adm[q].e0 = destination; one
// thread per (row, move).
__global__ void k_ppr_move_cols(float *X, void *C0, void *C1, uint32_t cap, uint32_t blk,
                                uint32_t fmt, const PprAdmit *adm, uint32_t na, uint32_t N)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)N * na) return;
    const uint32_t row = (uint32_t)(i / na), q = (uint32_t)(i % na);
    const uint64_t s = (uint64_t)row * cap + adm[q].slot, d = (uint64_t)row * cap + adm[q].e0;
    X[d] = X[s];
    if (fmt == AG_PREC_FP32) {
        float *c0 = static_cast<float *>(C0), *c1 = static_cast<float *>(C1);
        c0[d] = c0[s];
        c1[d] = c1[s];
    } else {
        const uint32_t hs = split_hi(cap, blk, adm[q].slot), ls = split_lo(cap, blk, adm[q].slot);
        const uint32_t hd = split_hi(cap, blk, adm[q].e0), ld = split_lo(cap, blk, adm[q].e0);
        for (void *buf : {C0, C1}) {
            char *r = static_cast<char *>(buf) + (uint64_t)row * cap * 4;
            *reinterpret_cast<uint16_t *>(r + hd) = *reinterpret_cast<const uint16_t *>(r + hs);
            *reinterpret_cast<uint16_t *>(r + ld) = *reinterpret_cast<const uint16_t *>(r + ls);
        }
    }
}

// Destination slot -> source slot of a row rewrite (identity for a plain
// layout switch; a compaction maps the packed slots to the columns they take).
struct SlotMap {
    uint8_t src[kMaxCap];
};

inline SlotMap identity_map()
{
    SlotMap m;
    for (uint32_t c = 0; c < kMaxCap; c++) m.src[c] = (uint8_t)c;
    return m;
}

// Rows rewritten in place through LDS, kRewriteBytes of rows per workgroup
// with 16-byte loads and stores.  split = 0: 32-bit words (FP32 iterate or
// contributions) move with the slot map.  split = 1: split-BF16 rows are
// re-laid from `from`-column blocks to `to`-column blocks while the halves
// move with the slot map.  A table of source positions per output position,
// built once per workgroup, keeps divisions out of the copy loop.
__global__ __launch_bounds__(kRelayBlock) void k_ppr_rewrite_rows(void *A, uint32_t N, uint32_t cap,
                                                                 uint32_t split, uint32_t from,
                                                                 uint32_t to, SlotMap m)
{
    __shared__ uint4 h4[kRewriteBytes / 16];
    __shared__ uint16_t tab[2 * kMaxCap];
    const uint32_t rows = kRewriteBytes / (cap * 4);
    const uint64_t r0 = (uint64_t)blockIdx.x * rows;
    const uint32_t nr = (uint32_t)min<uint64_t>(rows, N - r0);
    uint4 *g = reinterpret_cast<uint4 *>(static_cast<char *>(A) + r0 * cap * 4);
    const uint32_t nq = nr * cap / 4; // cap is a multiple of 8
    for (uint32_t i = threadIdx.x; i < nq; i += kRelayBlock) h4[i] = g[i];
    const uint32_t j = threadIdx.x;
    if (split && j < 2 * cap) {
        // 16-bit position j of the new layout -> column and half -> old position
        const uint32_t b = (j / (2 * to)) * to; // first column of j's block
        const uint32_t wb = min(to, cap - b), r = j - 2 * b;
        const bool hi = r < wb;
        const uint32_t col = m.src[hi ? b + r : b + r - wb];
        tab[j] = (uint16_t)((hi ? split_hi(cap, from, col) : split_lo(cap, from, col)) / 2);
    } else if (!split && j < cap) {
        tab[j] = m.src[j];
    }
    __syncthreads();
    const uint32_t *h32 = reinterpret_cast<const uint32_t *>(h4);
    const uint16_t *h16 = reinterpret_cast<const uint16_t *>(h4);
    for (uint32_t i = threadIdx.x; i < nq; i += kRelayBlock) {
        const uint32_t w0 = 4 * i, row = w0 / cap, c0 = w0 - row * cap; // 4 words of one row
        uint32_t o[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t c = c0 + k;
            o[k] = split ? (uint32_t)h16[row * 2 * cap + tab[2 * c]] |
                               ((uint32_t)h16[row * 2 * cap + tab[2 * c + 1]] << 16)
                         : h32[row * cap + tab[c]];
        }
        g[i] = make_uint4(o[0], o[1], o[2], o[3]);
    }
}

__global__ void k_ppr_move_state(PprCol *cs, const PprAdmit *adm, uint32_t na)
{
    const uint32_t q = threadIdx.x;
    if (q < na) cs[adm[q].e0] = cs[adm[q].slot];
}

__global__ void k_ppr_extract(const float *X, const uint32_t *order, uint32_t N, uint32_t cap,
                              uint32_t c, float *out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    out[order[i]] = X[(uint64_t)i * cap + c];
}

// Vertex-order writer (coalesced writes, suited to host memory over the bus).
__global__ void k_ppr_extract_vo(const float *X, const uint32_t *rank, uint32_t N, uint32_t cap,
                                 uint32_t c, float *out)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    out[v] = X[(uint64_t)rank[v] * cap + c];
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_set_bits(uint32_t *bits, const uint32_t *v, uint32_t n)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) atomicOr(&bits[v[q] >> 5], 1u << (v[q] & 31));
}

__global__ void k_map_labels(const uint32_t *rank, const uint32_t *ids, uint32_t n, uint32_t *out)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = rank[ids[q]];
}

uint32_t vec_width(uint32_t fmt) { return fmt == AG_PREC_FP32 ? 4u : 8u; }

// Launch of the sweep pair for one (lanes per row, storage type).
template <int LPR, typename CT>
void launch_lpr(const PprArgs &a, uint32_t g1, uint32_t g2, hipStream_t s)
{
    hipLaunchKernelGGL((k_ppr_sweep<LPR, CT>), dim3(g1), dim3(kBlock), 0, s, a);
    if (a.nmulti) {
        PprArgs b = a;
        b.part_base = g1;
        hipLaunchKernelGGL((k_ppr_rows<LPR, CT>), dim3(g2), dim3(kBlock), 0, s, b);
    }
}

template <typename CT>
void launch_ct(uint32_t lpr, const PprArgs &a, uint32_t g1, uint32_t g2, hipStream_t s)
{
    switch (lpr) {
    case 1: launch_lpr<1, CT>(a, g1, g2, s); break;
    case 2: launch_lpr<2, CT>(a, g1, g2, s); break;
    case 4: launch_lpr<4, CT>(a, g1, g2, s); break;
    case 8: launch_lpr<8, CT>(a, g1, g2, s); break;
    default:
        if constexpr (sizeof(CT) == 4) launch_lpr<16, CT>(a, g1, g2, s);
        else fail(AG_ERR_INVALID, "BF16 columns: at most 64 columns per wave");
        break;
    }
}

template <int LPR, typename CT>
int sweep_blocks_per_cu()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_ppr_sweep<LPR, CT>, kBlock, 0) != hipSuccess ||
        nb <= 0)
        nb = 2;
    return std::min(nb, 16);
}

template <typename CT>
int blocks_per_cu(uint32_t lpr)
{
    switch (lpr) {
    case 1: return sweep_blocks_per_cu<1, CT>();
    case 2: return sweep_blocks_per_cu<2, CT>();
    case 4: return sweep_blocks_per_cu<4, CT>();
    case 8: return sweep_blocks_per_cu<8, CT>();
    default:
        if constexpr (sizeof(CT) == 4) return sweep_blocks_per_cu<16, CT>();
        return 2;
    }
}

uint32_t log2u(uint32_t v)
{
    uint32_t r = 0;
    while ((1u << (r + 1)) <= v) r++;
    return r;
}

} // namespace

// Chunk table of the rows [0, n_nonlow) (in-degree above the plan's low_max):
// kChunk edges per chunk, and the list of rows of more than one chunk.
void ensure_ppr_chunks(ag_ctx *ctx, Plan *p)
{
    std::lock_guard<std::mutex> lk(p->mu);
    if (p->ppr_ready) return;
    hipStream_t s = ctx->stream;
    const uint32_t nr = p->n_nonlow;
    std::vector<uint32_t> roff(nr + 1, 0u);
    AG_HIP(hipMemcpyAsync(roff.data(), p->roff.p, (size_t)(nr + 1) * 4, hipMemcpyDeviceToHost, s));
    AG_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> first(nr + 1), rows, multi;
    uint32_t nch = 0;
    for (uint32_t r = 0; r < nr; r++) {
        first[r] = nch;
        const uint32_t deg = roff[r + 1] - roff[r];
        const uint32_t c = std::max<uint32_t>(1, ceil_div(deg, kChunk));
        for (uint32_t k = 0; k < c; k++) rows.push_back(r);
        if (c > 1) multi.push_back(r);
        nch += c;
    }
    first[nr] = nch;
    p->ppr_ch_first.alloc((size_t)(nr + 1) * 4);
    p->ppr_ch_row.alloc((size_t)std::max<uint32_t>(nch, 1) * 4);
    p->ppr_multi.alloc((size_t)std::max<size_t>(multi.size(), 1) * 4);
    AG_HIP(hipMemcpyAsync(p->ppr_ch_first.p, first.data(), first.size() * 4, hipMemcpyHostToDevice, s));
    if (nch) AG_HIP(hipMemcpyAsync(p->ppr_ch_row.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice, s));
    if (!multi.empty())
        AG_HIP(hipMemcpyAsync(p->ppr_multi.p, multi.data(), multi.size() * 4, hipMemcpyHostToDevice, s));
    AG_HIP(hipStreamSynchronize(s));
    p->ppr_nch = nch;
    p->ppr_nmulti = (uint32_t)multi.size();
    p->ppr_ready = true;
}

const std::vector<uint32_t> &rank_host(ag_ctx *ctx, Plan *p)
{
    std::lock_guard<std::mutex> lk(p->mu);
    if (p->rank_host.size() != p->N) {
        std::vector<uint32_t> r(p->N);
        AG_HIP(hipMemcpyAsync(r.data(), p->rank.p, (size_t)p->N * 4, hipMemcpyDeviceToHost, ctx->stream));
        AG_HIP(hipStreamSynchronize(ctx->stream));
        p->rank_host.swap(r);
    }
    return p->rank_host;
}

// Bytes of each buffer a batch of `cap` columns allocates (the order of
// PprBatch's `mine` table): X, C0, C1, column state, batch state, residual
// and dangling partials, chunk partials, seed bitmap, output staging,
// admission descriptors.
static void batch_need(const ag_ctx *ctx, const Plan *p, uint32_t cap, size_t need[kKeepBufs])
{
    const size_t N = p->N;
    const size_t nparts = (size_t)ctx->num_cus * 16 * 2;
    const size_t v[kKeepBufs] = {
        N * cap * 4, N * cap * 4, N * cap * 4, sizeof(PprCol) * kMaxCap, sizeof(PprBatchState),
        nparts * kMaxCap * 8, nparts * kMaxCap * 8,
        (size_t)std::max<uint32_t>(p->ppr_nch, 1) * cap * 4, ((N + 31) / 32) * 4, N * 4,
        sizeof(PprAdmit) * kMaxCap};
    for (int i = 0; i < kKeepBufs; i++) need[i] = v[i];
}

uint64_t PprBatch::bytes_for(ag_ctx *ctx, ag_graph *g, uint32_t cap, uint32_t fmt)
{
    if (fmt != AG_PREC_FP32 && fmt != AG_PREC_BF16)
        fail(AG_ERR_INVALID, "batched columns store FP32 or BF16 (precision %u)", fmt);
    Plan *p = ensure_plan(ctx, g, nullptr);
    ensure_ppr_chunks(ctx, p);
    size_t need[kKeepBufs];
    batch_need(ctx, p, cap, need);
    uint64_t b = 1u << 20; // seed tables (labels, weights, rows without in-edges)
    for (size_t x : need) b += x;
    return b;
}

void ppr_release_kept(ag_graph *g)
{
    if (!g || !g->plan) return;
    std::lock_guard<std::mutex> lk(g->plan->mu);
    for (DevBuf &k : g->plan->ppr_keep) k.release();
}

PprBatch::PprBatch(ag_ctx *ctx, ag_graph *g, uint32_t cap, uint32_t cpw, uint32_t fmt, float damp,
                   bool extrapolate, bool exact_keep, hipStream_t stream)
    : ctx_(ctx), s_(stream ? stream : ctx->stream), cap_(cap), fmt_(fmt), damp_(damp),
      extrap_(extrapolate)
{
    if (cap == 0 || cap > kMaxCap || cap % 8) fail(AG_ERR_INVALID, "batch capacity must be 8..64, multiple of 8");
    if (fmt != AG_PREC_FP32 && fmt != AG_PREC_BF16)
        fail(AG_ERR_INVALID, "batched columns store FP32 or BF16 (precision %u); FP16 storage is "
                             "available to the single-query path", fmt);
    if (!(damp > 0.0f && damp < 1.0f)) fail(AG_ERR_INVALID, "damping must be in (0, 1)");
    p_ = ensure_plan(ctx, g, nullptr);
    ensure_ppr_chunks(ctx, p_);
    N_ = p_->N;
    if (N_ == 0) fail(AG_ERR_INVALID, "graph has no vertices");
    // Extrapolating columns keep the pull first sweep: with the exact first
    // iterate of the push their residual ratios settle differently, and on
    // R-MAT s22 the slowest of 64 columns took up to 3 more sweeps.
    if (g->has_full_fwd && !extrapolate) {
        fwd_off_ = g->fwd_off.as<uint32_t>();
        fwd_dst_ = g->fwd_dst.as<uint32_t>();
    }
    set_cpw(cpw);
    hipStream_t s = s_;
    {
        // The buffers of the last finished batch on this graph are lent to
        // this one, so a repeated batch call makes no hipMalloc / hipFree (the
        // column buffers alone are ~12 bytes per vertex and column).  Every
        // admitted column is cleared before use, and every other buffer is
        // (re)written before it is read.
        size_t need[kKeepBufs];
        batch_need(ctx, p_, cap, need);
        DevBuf *mine[kKeepBufs] = {&X_,        &C_[0], &C_[1], &cols_,      &bstate_, &blk_res_,
                                   &blk_dang_, &part_, &seed_bits_, &out_, &desc_};
        {
            std::lock_guard<std::mutex> lk(p_->mu);
            for (int i = 0; i < kKeepBufs; i++) {
                DevBuf &k = p_->ppr_keep[i];
                if (exact_keep ? k.bytes == need[i] : k.bytes >= need[i]) *mine[i] = std::move(k);
                else if (exact_keep) k.release();
            }
        }
        for (int i = 0; i < kKeepBufs; i++)
            if (!mine[i]->p) mine[i]->alloc(need[i]);
    }
    AG_HIP(hipMemsetAsync(bstate_.p, 0, sizeof(PprBatchState), s));
    AG_HIP(hipMemsetAsync(cols_.p, 0, cols_.bytes, s));
    q_.resize(cap);
    tags_.assign(cap, ~0ull);
    used_.assign(cap, 0);
    upload_seed_table();
}

PprBatch::~PprBatch()
{
    if (gexec_) (void)hipGraphExecDestroy(gexec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    (void)hipStreamSynchronize(s_);
    // keep the larger buffers on the graph for the next batch
    DevBuf *mine[kKeepBufs] = {&X_,        &C_[0], &C_[1], &cols_,      &bstate_, &blk_res_,
                               &blk_dang_, &part_, &seed_bits_, &out_, &desc_};
    std::lock_guard<std::mutex> lk(p_->mu);
    for (int i = 0; i < kKeepBufs; i++)
        if (mine[i]->bytes > p_->ppr_keep[i].bytes) p_->ppr_keep[i] = std::move(*mine[i]);
}

uint64_t PprBatch::bytes() const
{
    return X_.bytes + C_[0].bytes + C_[1].bytes + blk_res_.bytes + blk_dang_.bytes + part_.bytes +
           seed_bits_.bytes + out_.bytes + seed_v_.bytes + seed_c_.bytes + seed_w_.bytes +
           zrows_.bytes + desc_.bytes;
}

// Widest column group: VW columns (16 bytes of one row) per lane times a
// power of two of lanes, at most 64 columns.
void PprBatch::set_cpw(uint32_t cpw)
{
    if (cpw != 4 && cpw != 8 && cpw != 16 && cpw != 32 && cpw != 64)
        fail(AG_ERR_INVALID, "cols_per_wave must be 4, 8, 16, 32 or 64");
    cpw_ = std::max(cpw, vec_width(fmt_));
    graph_lpr_ = 0; // recapture
    if (gexec_) { (void)hipGraphExecDestroy(gexec_); gexec_ = nullptr; }
    if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
}

// Lanes per row for the current slots: the column group covers the used
// slots, up to cols_per_wave.
uint32_t PprBatch::lanes_per_row() const
{
    const uint32_t vw = vec_width(fmt_);
    uint32_t cols = vw;
    while (cols < hi_ && cols < cpw_) cols <<= 1;
    return std::min(cols, cpw_) / vw;
}

PprSeedSet PprBatch::make_seeds(const uint32_t *seeds, uint32_t n, double tol, uint32_t max_iter)
{
    if (n == 0 || !seeds) fail(AG_ERR_INVALID, "a query needs at least one seed vertex");
    std::map<uint32_t, uint32_t> cnt;
    for (uint32_t i = 0; i < n; i++) {
        if (seeds[i] >= N_) fail(AG_ERR_INVALID, "seed vertex %u >= num_vertices %u", seeds[i], N_);
        cnt[seeds[i]]++;
    }
    PprSeedSet q;
    // vertex ids -> labels through a host copy of the plan's rank array
    const std::vector<uint32_t> &rank = rank_host(ctx_, p_);
    for (auto &kv : cnt) {
        q.labels.push_back(rank[kv.first]);
        q.w.push_back((float)((double)kv.second / (double)n));
    }
    q.tol = tol;
    q.max_iter = max_iter ? max_iter : 1000;
    return q;
}

void PprBatch::upload_seed_table()
{
    hipStream_t s = s_;
    std::vector<uint32_t> v, c, z;
    std::vector<float> w;
    for (uint32_t k = 0; k < cap_; k++) {
        if (!used_[k]) continue;
        for (size_t q = 0; q < q_[k].labels.size(); q++) {
            v.push_back(q_[k].labels[q]);
            c.push_back(k);
            w.push_back(q_[k].w[q]);
            if (q_[k].labels[q] >= p_->n_nonzero) z.push_back(q_[k].labels[q]);
        }
    }
    // seed rows without in-edges, each once (one lane group updates a row)
    std::sort(z.begin(), z.end());
    z.erase(std::unique(z.begin(), z.end()), z.end());
    const size_t n = v.size();
    const size_t need = std::max<size_t>(n, 1) * 4;
    if (seed_v_.bytes < need) {
        const size_t nb = std::max<size_t>(need * 2, (size_t)cap_ * 64 * 4);
        seed_v_.alloc(nb);
        seed_c_.alloc(nb);
        seed_w_.alloc(nb);
        zrows_.alloc(nb);
        graph_lpr_ = 0; // the captured sweeps hold the old pointers
        if (gexec_) { (void)hipGraphExecDestroy(gexec_); gexec_ = nullptr; }
        if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
    }
    if (n) {
        AG_HIP(hipMemcpyAsync(seed_v_.p, v.data(), n * 4, hipMemcpyHostToDevice, s));
        AG_HIP(hipMemcpyAsync(seed_c_.p, c.data(), n * 4, hipMemcpyHostToDevice, s));
        AG_HIP(hipMemcpyAsync(seed_w_.p, w.data(), n * 4, hipMemcpyHostToDevice, s));
    }
    if (!z.empty()) AG_HIP(hipMemcpyAsync(zrows_.p, z.data(), z.size() * 4, hipMemcpyHostToDevice, s));
    const uint64_t words = (N_ + 31) / 32;
    hipLaunchKernelGGL(k_fill_u32, dim3(ceil_div(words, 256)), dim3(256), 0, s,
                       seed_bits_.as<uint32_t>(), words, 0u);
    if (n)
        hipLaunchKernelGGL(k_set_bits, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                           seed_bits_.as<uint32_t>(), seed_v_.as<uint32_t>(), (uint32_t)n);
    // K (slot high-water mark) and the table sizes live on the device so that
    // captured sweeps stay valid.
    PprBatchState bs{hi_, (uint32_t)z.size(), 0u, (uint32_t)n};
    AG_HIP(hipMemcpyAsync(bstate_.p, &bs, sizeof(bs), hipMemcpyHostToDevice, s));
    AG_HIP(hipGetLastError());
    // host vectors above are temporaries: wait for the copies
    AG_HIP(hipStreamSynchronize(s));
}

uint32_t PprBatch::admit(const PprSeedSet &q, uint64_t tag)
{
    uint32_t c = 0;
    while (c < cap_ && used_[c]) c++;
    if (c >= cap_) fail(AG_ERR_ADMISSION, "no free column slot");
    used_[c] = 1;
    used_count_++;
    hi_ = std::max(hi_, c + 1);
    q_[c] = q;
    tags_[c] = tag;
    pending_.push_back(c);
    dirty_ = true;
    return c;
}

void PprBatch::release(uint32_t slot)
{
    if (!used(slot)) fail(AG_ERR_INVALID, "slot %u not in use", slot);
    used_[slot] = 0;
    used_count_--;
    q_[slot] = PprSeedSet();
    tags_[slot] = ~0ull;
    pending_.erase(std::remove(pending_.begin(), pending_.end(), slot), pending_.end());
    stop_.push_back(slot);
    while (hi_ && !used_[hi_ - 1]) hi_--;
    dirty_ = true;
}

bool PprBatch::cancel(uint64_t tag)
{
    for (uint32_t k = 0; k < cap_; k++)
        if (used_[k] && tags_[k] == tag) {
            release(k);
            return true;
        }
    return false;
}

void PprBatch::compact()
{
    if (!pending_.empty() || used_count_ == 0 || used_count_ == hi_) return;
    // Moving columns costs about half a wide sweep (s22: ~1 ms against 2.4 ms
    // fine sweeps that narrow to 0.9 ms): not worth it for the last two sweeps.
    if (tail_ < kCompactMinTail) return;
    const uint32_t vw = vec_width(fmt_);
    auto width = [&](uint32_t h) {
        uint32_t c = vw;
        while (c < h) c <<= 1;
        return c;
    };
    // only when the packed slots need narrower column groups
    if (width(used_count_) >= width(hi_)) return;
    std::vector<uint32_t> src, dst;
    for (uint32_t k = hi_; k-- > used_count_;)
        if (used_[k]) src.push_back(k);
    for (uint32_t k = 0; k < used_count_ && dst.size() < src.size(); k++)
        if (!used_[k]) dst.push_back(k);
    std::vector<PprAdmit> mv;
    for (size_t i = 0; i < src.size(); i++) mv.push_back({src[i], dst[i], 0, 0, 0.0});
    const uint32_t to = layout_for(used_count_);
    hipStream_t s = s_;
    AG_HIP(hipMemcpyAsync(desc_.p, mv.data(), mv.size() * sizeof(PprAdmit), hipMemcpyHostToDevice, s));
    if (to != blk_) {
        // The packed width switches the split layout: one in-place rewrite
        // per buffer moves the columns and re-lays the rows in the same pass
        // (rows without in-edges are never rewritten by a sweep, so both
        // contribution buffers are rewritten whole).
        SlotMap m = identity_map();
        for (size_t i = 0; i < src.size(); i++) m.src[dst[i]] = (uint8_t)src[i];
        const uint32_t grid = ceil_div(N_, kRewriteBytes / (cap_ * 4));
        hipLaunchKernelGGL(k_ppr_rewrite_rows, dim3(grid), dim3(kRelayBlock), 0, s, X_.p, N_, cap_, 0u,
                           blk_, blk_, m);
        for (int b = 0; b < 2; b++)
            hipLaunchKernelGGL(k_ppr_rewrite_rows, dim3(grid), dim3(kRelayBlock), 0, s, C_[b].p, N_, cap_,
                               1u, blk_, to, m);
        blk_ = to;
    } else {
        hipLaunchKernelGGL(k_ppr_move_cols, dim3(ceil_div((uint64_t)N_ * mv.size(), 256)), dim3(256), 0, s,
                           X_.as<float>(), C_[0].p, C_[1].p, cap_, blk_, fmt_, desc_.as<PprAdmit>(),
                           (uint32_t)mv.size(), N_);
    }
    hipLaunchKernelGGL(k_ppr_move_state, dim3(1), dim3(kMaxCap), 0, s, cols_.as<PprCol>(),
                       desc_.as<PprAdmit>(), (uint32_t)mv.size());
    AG_HIP(hipGetLastError());
    AG_HIP(hipStreamSynchronize(s)); // mv is a temporary
    for (size_t i = 0; i < src.size(); i++) {
        const uint32_t a = src[i], b = dst[i];
        used_[b] = 1;
        q_[b] = q_[a];
        tags_[b] = tags_[a];
        used_[a] = 0;
        q_[a] = PprSeedSet();
        tags_[a] = ~0ull;
        // b now holds a running column: it must not be stopped by an earlier release
        stop_.erase(std::remove(stop_.begin(), stop_.end(), b), stop_.end());
    }
    while (hi_ && !used_[hi_ - 1]) hi_--;
    dirty_ = true;
}

// Split-BF16 block for the used slots: 64-column blocks while more than 48
// slots are in use (coarse sweeps read one line per row), 32-column blocks
// once at most 32 are (fine sweeps of the packed columns read one line); in
// between the current layout stays.  Running columns are re-laid in place in
// both contribution buffers (rows no sweep rewrites keep their values).
uint32_t PprBatch::layout_for(uint32_t hi) const
{
    if (fmt_ != AG_PREC_BF16 || cap_ <= 32) return blk_; // one block either way
    if (blk_ == 64 && hi <= 32) return 32;
    if (blk_ == 32 && hi > 48) return 64;
    return blk_;
}

void PprBatch::choose_layout()
{
    const uint32_t want = layout_for(hi_);
    if (want == blk_) return;
    if (used_count_ > pending_.size()) { // running columns hold data
        hipStream_t s = s_;
        const uint32_t rows = kRewriteBytes / (cap_ * 4);
        for (int b = 0; b < 2; b++)
            hipLaunchKernelGGL(k_ppr_rewrite_rows, dim3(ceil_div(N_, rows)), dim3(kRelayBlock), 0, s,
                               C_[b].p, N_, cap_, 1u, blk_, want, identity_map());
        AG_HIP(hipGetLastError());
    }
    blk_ = want;
}

// Applies the step's slot changes: stops released columns, rebuilds the seed
// table, and initialises all admitted columns in one pass over the rows.
void PprBatch::flush()
{
    if (!dirty_) return;
    choose_layout();
    hipStream_t s = s_;
    std::vector<PprAdmit> adm;
    if (!stop_.empty()) {
        std::sort(stop_.begin(), stop_.end());
        stop_.erase(std::unique(stop_.begin(), stop_.end()), stop_.end());
        for (uint32_t c : stop_) adm.push_back({c, 0, 0, 0, 0.0});
        AG_HIP(hipMemcpyAsync(desc_.p, adm.data(), adm.size() * sizeof(PprAdmit),
                              hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_ppr_stop_cols, dim3(1), dim3(kMaxCap), 0, s, cols_.as<PprCol>(),
                           desc_.as<PprAdmit>(), (uint32_t)adm.size());
        AG_HIP(hipStreamSynchronize(s));
        stop_.clear();
    }
    upload_seed_table();
    if (!pending_.empty()) {
        // entry ranges of each admitted slot in the table just uploaded:
        // columns whose first sweep is a push first, then the others
        adm.clear();
        std::vector<PprAdmit> pulled;
        uint32_t off = 0;
        for (uint32_t k = 0; k < cap_; k++) {
            if (!used_[k]) continue;
            const uint32_t n = (uint32_t)q_[k].labels.size();
            if (std::find(pending_.begin(), pending_.end(), k) != pending_.end())
                (fwd_off_ && n <= kPushMaxSeeds ? adm : pulled)
                    .push_back({k, off, off + n, q_[k].max_iter, q_[k].tol});
            off += n;
        }
        const uint32_t npush = (uint32_t)adm.size();
        adm.insert(adm.end(), pulled.begin(), pulled.end());
        AG_HIP(hipMemcpyAsync(desc_.p, adm.data(), adm.size() * sizeof(PprAdmit),
                              hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_ppr_clear_cols, dim3(ceil_div((uint64_t)N_ * adm.size(), 256)), dim3(256),
                           0, s, X_.as<float>(), C_[0].p, C_[1].p, N_, cap_, blk_,
                           desc_.as<PprAdmit>(), (uint32_t)adm.size(), fmt_);
        if (npush)
            hipLaunchKernelGGL(k_ppr_admit_push, dim3(npush), dim3(kPushBlock), 0, s,
                               X_.as<float>(), C_[cur_].p, p_->outdeg.as<uint32_t>(),
                               p_->order.as<uint32_t>(), p_->rank.as<uint32_t>(), fwd_off_, fwd_dst_,
                               cap_, blk_, fmt_, desc_.as<PprAdmit>(), seed_v_.as<uint32_t>(),
                               seed_w_.as<float>(), cols_.as<PprCol>(), damp_, extrap_ ? 1u : 0u);
        if (!pulled.empty())
            hipLaunchKernelGGL(k_ppr_seed_cols, dim3((uint32_t)pulled.size()), dim3(256), 0, s,
                               X_.as<float>(), C_[cur_].p, p_->outdeg.as<uint32_t>(), cap_, blk_, fmt_,
                               desc_.as<PprAdmit>() + npush, seed_v_.as<uint32_t>(),
                               seed_w_.as<float>(), cols_.as<PprCol>(), damp_);
        AG_HIP(hipGetLastError());
        AG_HIP(hipStreamSynchronize(s)); // adm is a temporary
        pending_.clear();
    }
    dirty_ = false;
}

void PprBatch::launch_sweep(bool from0)
{
    PprArgs a;
    std::memset(&a, 0, sizeof(a));
    a.N = N_;
    a.cap = cap_;
    a.blk = blk_;
    a.n_nonlow = p_->n_nonlow;
    a.n_nonzero = p_->n_nonzero;
    a.nch = p_->ppr_nch;
    a.nmulti = p_->ppr_nmulti;
    a.damp = damp_;
    a.extrap = extrap_ ? 1u : 0u;
    a.roff = p_->roff.as<uint32_t>();
    a.cols = p_->cols.as<uint32_t>();
    a.slice_ptr = p_->slice_ptr.as<uint32_t>();
    a.outdeg = p_->outdeg.as<uint32_t>();
    a.ch_row = p_->ppr_ch_row.as<uint32_t>();
    a.ch_first = p_->ppr_ch_first.as<uint32_t>();
    a.multi = p_->ppr_multi.as<uint32_t>();
    a.X = X_.as<float>();
    a.Cin = C_[from0 ? 0 : 1].p;
    a.Cout = C_[from0 ? 1 : 0].p;
    a.part = part_.as<float>();
    a.seed_bits = seed_bits_.as<uint32_t>();
    a.seed_v = seed_v_.as<uint32_t>();
    a.seed_c = seed_c_.as<uint32_t>();
    a.seed_w = seed_w_.as<float>();
    a.zrows = zrows_.as<uint32_t>();
    a.cs = cols_.as<PprCol>();
    a.bs = bstate_.as<PprBatchState>();
    a.blk_res = blk_res_.as<double>();
    a.blk_dang = blk_dang_.as<double>();
    a.part_base = 0;
    const uint32_t lpr = lanes_per_row();
    int &nb = blocks_[log2u(lpr)];
    if (!nb) nb = fmt_ == AG_PREC_FP32 ? blocks_per_cu<float>(lpr) : blocks_per_cu<bf16_t>(lpr);
    const uint32_t g1 = (uint32_t)ctx_->num_cus * (uint32_t)nb;
    const uint32_t rows_per_block = (64 / lpr / kSlots) * kWaves;
    const uint32_t g2 = std::max<uint32_t>(1, std::min<uint32_t>(g1, ceil_div(p_->ppr_nmulti, rows_per_block)));
    hipStream_t s = s_;
    if (fmt_ == AG_PREC_FP32) launch_ct<float>(lpr, a, g1, g2, s);
    else launch_ct<bf16_t>(lpr, a, g1, g2, s);
    const uint32_t nparts = g1 + (a.nmulti ? g2 : 0u);
    hipLaunchKernelGGL(k_ppr_reduce, dim3(kMaxCap), dim3(kRedBlock), 0, s, a, nparts);
    AG_HIP(hipGetLastError());
}

void PprBatch::sweep(uint32_t n, bool replay, bool timed)
{
    flush();
    if (used_count_ == 0 || n == 0) return;
    hipStream_t s = s_;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timed) {
        AG_HIP(hipEventCreate(&e0));
        AG_HIP(hipEventCreate(&e1));
        AG_HIP(hipEventRecord(e0, s));
    }
    uint32_t done = 0;
    if (replay) {
        // A captured pair of sweeps (C0 -> C1 -> C0), replayed while the
        // current buffer is C0; recaptured when the lanes per row or the
        // split layout change.
        if (cur_ == 1) { launch_sweep(false); cur_ = 0; done++; }
        const uint32_t lpr = lanes_per_row();
        if (gexec_ && (graph_lpr_ != lpr || graph_blk_ != blk_)) {
            (void)hipGraphExecDestroy(gexec_);
            (void)hipGraphDestroy(graph_);
            gexec_ = nullptr;
            graph_ = nullptr;
        }
        if (n - done >= 2 && !gexec_) {
            AG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
            launch_sweep(true);
            launch_sweep(false);
            AG_HIP(hipStreamEndCapture(s, &graph_));
            AG_HIP(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
            graph_lpr_ = lpr;
            graph_blk_ = blk_;
        }
        while (n - done >= 2) {
            AG_HIP(hipGraphLaunch(gexec_, s));
            done += 2;
        }
    }
    while (done < n) {
        launch_sweep(cur_ == 0);
        cur_ ^= 1;
        done++;
    }
    if (!timed) return;
    AG_HIP(hipEventRecord(e1, s));
    AG_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    AG_HIP(hipEventElapsedTime(&ms, e0, e1));
    last_ms_ = ms / n;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

std::vector<PprBatch::Done> PprBatch::poll()
{
    std::vector<Done> out;
    if (hi_ == 0) return out;
    flush();
    std::vector<PprCol> cs(hi_);
    AG_HIP(hipMemcpyAsync(cs.data(), cols_.p, hi_ * sizeof(PprCol), hipMemcpyDeviceToHost,
                          s_));
    AG_HIP(hipStreamSynchronize(s_));
    uint32_t hint = 8, tail = 0;
    bool known = false;
    for (uint32_t k = 0; k < hi_; k++) {
        if (!used_[k]) continue;
        if (cs[k].done) {
            out.push_back({k, tags_[k], cs[k].iters, cs[k].residual, cs[k].fine_at});
        } else if (std::isfinite(cs[k].residual) && cs[k].residual > 0.0 && cs[k].tol > 0.0) {
            const double need = std::ceil(std::log(cs[k].tol / cs[k].residual) / std::log(0.6));
            hint = std::min<uint32_t>(hint, (uint32_t)std::max(1.0, std::min(8.0, need)));
            tail = std::max<uint32_t>(tail, (uint32_t)std::max(1.0, std::min(64.0, need)));
            known = true;
        } else {
            hint = std::min<uint32_t>(hint, 4);
            tail = 64;
        }
    }
    hint_ = known || hint < 8 ? hint : 4;
    tail_ = tail;
    return out;
}

void PprBatch::extract_to(uint32_t slot, float *dst)
{
    if (!used(slot)) fail(AG_ERR_INVALID, "slot %u not in use", slot);
    if (!dst) return;
    hipLaunchKernelGGL(k_ppr_extract_vo, dim3(ceil_div(N_, 256)), dim3(256), 0, s_,
                       X_.as<float>(), p_->rank.as<uint32_t>(), N_, cap_, slot, dst);
    AG_HIP(hipGetLastError());
}

void PprBatch::extract(uint32_t slot, float *host_out)
{
    if (!used(slot)) fail(AG_ERR_INVALID, "slot %u not in use", slot);
    if (!host_out) return;
    hipStream_t s = s_;
    hipLaunchKernelGGL(k_ppr_extract, dim3(ceil_div(N_, 256)), dim3(256), 0, s, X_.as<float>(),
                       p_->order.as<uint32_t>(), N_, cap_, slot, out_.as<float>());
    AG_HIP(hipGetLastError());
    AG_HIP(hipMemcpyAsync(host_out, out_.p, (size_t)N_ * 4, hipMemcpyDeviceToHost, s));
    AG_HIP(hipStreamSynchronize(s));
}

uint32_t ppr_autotune(ag_ctx *ctx, ag_graph *g, uint32_t fmt, uint32_t capacity, bool exact_keep)
{
    if (fmt != AG_PREC_FP32 && fmt != AG_PREC_BF16)
        fail(AG_ERR_INVALID, "batched columns store FP32 or BF16 (precision %u)", fmt);
    if (capacity == 0 || capacity > kMaxCap || capacity % 8)
        fail(AG_ERR_INVALID, "autotune capacity must be 8..64, multiple of 8");
    const bool cache = capacity == kMaxCap;
    if (cache && g->ppr_cpw[fmt]) return g->ppr_cpw[fmt];
    Plan *p = ensure_plan(ctx, g, nullptr);
    const uint32_t N = p->N;
    // a full batch of single-seed queries at the capacity the caller will run
    PprBatch b(ctx, g, capacity, 64, fmt, 0.85f, false, exact_keep);
    for (uint32_t q = 0; q < capacity; q++) {
        const uint32_t seed = (uint32_t)(((uint64_t)q * 2654435761u) % N);
        b.admit(b.make_seeds(&seed, 1, 0.0, 1u << 30), q);
    }
    // each width, widest first: one warm sweep, then the fastest of three
    // timed ones (one timed sweep let run-to-run noise of a few percent decide
    // between widths within that distance).  Narrower widths re-read the
    // index stream once per column group, so the search stops at the first
    // width whose first timing is 1.3x the best (those narrower only slow on).
    uint32_t best = 64;
    double best_ms = 1e30;
    for (uint32_t cpw : {64u, 32u, 16u, 8u, 4u}) {
        if (cpw < vec_width(fmt)) continue;
        b.set_cpw(cpw);
        b.sweep(1, false);
        b.sweep(2, false);
        double ms = b.last_sweep_ms();
        if (ms > 1.3 * best_ms) break;
        for (int k = 0; k < 2; k++) {
            b.sweep(2, false);
            ms = std::min(ms, b.last_sweep_ms());
        }
        if (ms < best_ms) {
            best_ms = ms;
            best = cpw;
        }
    }
    if (cache) g->ppr_cpw[fmt] = best;
    return best;
}

} // namespace ag

using namespace ag;

extern "C" {

void ag_ppr_options_default(ag_ppr_options *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->precision = AG_PREC_BF16;
    o->cols_per_wave = 0;
    o->damp = 0.85f;
    o->use_graph_replay = 1;
}

int ag_ppr_batch_bytes(ag_ctx *ctx, ag_graph *g, uint32_t capacity, uint32_t precision,
                       uint64_t *bytes)
{
    return guard([&] {
        check_ctx(ctx);
        if (!g || !bytes) fail(AG_ERR_INVALID, "null argument");
        if (capacity == 0 || capacity > kMaxCap || capacity % 8)
            fail(AG_ERR_INVALID, "batch capacity must be 8..64, multiple of 8");
        *bytes = PprBatch::bytes_for(ctx, g, capacity, precision);
    });
}

int ag_ppr_autotune(ag_ctx *ctx, ag_graph *g, uint32_t precision, uint32_t *cpw)
{
    return guard([&] {
        PhaseRange phase_("ag_ppr_autotune");
        check_ctx(ctx);
        if (!g || !cpw) fail(AG_ERR_INVALID, "null argument");
        *cpw = ppr_autotune(ctx, g, precision);
    });
}

int ag_ppr_batch(ag_ctx *ctx, ag_graph *g, const ag_ppr_query *q, uint32_t nq,
                 const ag_ppr_options *opt, float *x, ag_ppr_stat *stats, double *elapsed_ms)
{
    return guard([&] {
        PhaseRange phase_("ag_ppr_batch");
        check_ctx(ctx);
        if (!g || (nq && !q)) fail(AG_ERR_INVALID, "null argument");
        ag_ppr_options o;
        if (opt) o = *opt; else ag_ppr_options_default(&o);
        if (nq == 0) return;
        const uint32_t cpw = o.cols_per_wave ? o.cols_per_wave : ppr_autotune(ctx, g, o.precision);
        const uint32_t cap = std::min<uint32_t>(kMaxCap, ((nq + 7) / 8) * 8);
        PprBatch b(ctx, g, cap, cpw, o.precision, o.damp, o.extrapolate != 0);
        const uint32_t N = g->N;
        std::vector<PprSeedSet> sets(nq);
        for (uint32_t i = 0; i < nq; i++)
            sets[i] = b.make_seeds(q[i].seeds, q[i].nseeds, q[i].tol, q[i].max_iter);
        AG_HIP(hipStreamSynchronize(ctx->stream));
        const auto t0 = std::chrono::steady_clock::now();
        uint32_t next = 0, finished = 0;
        while (next < nq && b.free_slots()) { b.admit(sets[next], next); next++; }
        while (finished < nq) {
            b.sweep(b.sweeps_hint(), o.use_graph_replay != 0, false);
            auto fin = b.retire([&](uint64_t tag) { return x ? x + (size_t)tag * N : nullptr; },
                                next >= nq);
            for (auto &f : fin) {
                if (stats) {
                    stats[f.tag].iterations = f.iters;
                    stats[f.tag].residual = f.residual;
                    stats[f.tag].fine_sweep = f.fine_at;
                    stats[f.tag].status = AG_OK;
                    stats[f.tag].latency_ms = std::chrono::duration<double, std::milli>(
                        std::chrono::steady_clock::now() - t0).count();
                }
                finished++;
            }
            while (next < nq && b.free_slots()) { b.admit(sets[next], next); next++; }
        }
        AG_HIP(hipStreamSynchronize(ctx->stream));
        if (elapsed_ms)
            *elapsed_ms = std::chrono::duration<double, std::milli>(
                std::chrono::steady_clock::now() - t0).count();
    });
}

int ag_ppr_single(ag_ctx *ctx, ag_graph *g, const ag_ppr_query *q, uint32_t precision, float *x,
                  ag_ppr_stat *stat)
{
    ag_ppr_options o;
    ag_ppr_options_default(&o);
    o.precision = precision;
    return ag_ppr_single_opts(ctx, g, q, &o, x, stat);
}

int ag_ppr_single_opts(ag_ctx *ctx, ag_graph *g, const ag_ppr_query *q, const ag_ppr_options *opt,
                       float *x, ag_ppr_stat *stat)
{
    return guard([&] {
        PhaseRange phase_("ag_ppr_single_opts");
        check_ctx(ctx);
        if (!opt) fail(AG_ERR_INVALID, "null argument");
        const uint32_t precision = opt->precision;
        if (!g || !q) fail(AG_ERR_INVALID, "null argument");
        Plan *p = ensure_plan(ctx, g, nullptr);
        const uint32_t N = p->N;
        if (N == 0) fail(AG_ERR_INVALID, "graph has no vertices");
        if (!q->nseeds || !q->seeds) fail(AG_ERR_INVALID, "a query needs at least one seed vertex");
        std::map<uint32_t, uint32_t> cnt;
        for (uint32_t i = 0; i < q->nseeds; i++) {
            if (q->seeds[i] >= N) fail(AG_ERR_INVALID, "seed vertex %u >= num_vertices %u", q->seeds[i], N);
            cnt[q->seeds[i]]++;
        }
        std::vector<uint32_t> ids;
        std::vector<float> w;
        for (auto &kv : cnt) {
            ids.push_back(kv.first);
            w.push_back((float)((double)kv.second / (double)q->nseeds));
        }
        hipStream_t s = ctx->stream;
        const uint32_t n = (uint32_t)ids.size();
        if (p->ppr_ids.bytes < (size_t)n * 4) {
            const size_t cap = std::max<size_t>((size_t)n * 4, 4096);
            p->ppr_ids.alloc(cap);
            p->ppr_labels.alloc(cap);
            p->ppr_w.alloc(cap);
        }
        if (!p->ppr_out.p) {
            p->ppr_bits.alloc((size_t)((N + 31) / 32) * 4);
            p->ppr_out.alloc((size_t)N * 4);
        }
        DevBuf &din = p->ppr_ids, &lab = p->ppr_labels, &dw = p->ppr_w, &bits = p->ppr_bits,
               &out = p->ppr_out;
        AG_HIP(hipMemcpyAsync(din.p, ids.data(), n * 4, hipMemcpyHostToDevice, s));
        AG_HIP(hipMemcpyAsync(dw.p, w.data(), n * 4, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_map_labels, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                           p->rank.as<uint32_t>(), din.as<uint32_t>(), n, lab.as<uint32_t>());
        const uint64_t words = (N + 31) / 32;
        hipLaunchKernelGGL(k_fill_u32, dim3(ceil_div(words, 256)), dim3(256), 0, s,
                           bits.as<uint32_t>(), words, 0u);
        hipLaunchKernelGGL(k_set_bits, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                           bits.as<uint32_t>(), lab.as<uint32_t>(), n);
        AG_HIP(hipGetLastError());
        SeedTable st;
        st.bits = bits.as<uint32_t>();
        st.v = lab.as<uint32_t>();
        st.w = dw.as<float>();
        st.n = n;
        ag_pr_options o;
        ag_pr_options_default(&o);
        o.mode = kModePPR;
        o.iterations = q->max_iter ? q->max_iter : 1000;
        o.epsilon = q->tol;
        o.precision = precision;
        o.damp = opt->damp;
        o.flags = (precision != AG_PREC_FP32 ? AG_PR_FLAG_FINISH_FP32 : 0u) |
                  (opt->extrapolate ? AG_PR_FLAG_EXTRAPOLATE : 0u);
        ag_pr_result r;
        const auto t0 = std::chrono::steady_clock::now();
        run_pagerank_seeded(ctx, g, o, st, out.as<float>(), &r);
        if (x) AG_HIP(hipMemcpyAsync(x, out.p, (size_t)N * 4, hipMemcpyDeviceToHost, s));
        AG_HIP(hipStreamSynchronize(s));
        if (stat) {
            stat->iterations = r.iterations;
            stat->status = AG_OK;
            stat->residual = r.error_total;
            stat->fine_sweep = r.fp32_switch_sweep;
            stat->latency_ms = std::chrono::duration<double, std::milli>(
                std::chrono::steady_clock::now() - t0).count();
        }
    });
}

} // extern "C"
