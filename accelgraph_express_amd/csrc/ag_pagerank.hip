// Pull PageRank on the device: the gather of the reference's PageRank CSR PULL
// compute unit (03_capi_integration/accelerator_rtl/cu_control/cu_PageRank/
// CSR/PULL: next[v] = sum of auxiliary1[inverse_dest[j]] over v's inverse
// edge range, cu_edge_job_control.sv:373-449, cu_sum_kernel_control.sv:188-228)
// fused with the host loops of pageRankPullGraphCSR (pageRank.c:1024-1085).
//
// Layout ("plan"), built once per graph on the device:
//   * vertices are relabelled by descending in-degree (stable by id), so the
//     in-degree bins are contiguous label ranges and the hottest sources sit
//     together in the contribution vector;
//   * hub rows   (in-degree >= hub_min): CSR, split in fixed chunks of
//                  hub_chunk edges, one workgroup per chunk; a one-workgroup
//                  finalize kernel launched after the sweep sums each row's
//                  chunk partials in chunk order and updates the vertex;
//   * mid rows   (low_max < in-degree < hub_min): CSR, one wave per row,
//                  per-lane partial sums and a shuffle reduction;
//   * low rows   (1 <= in-degree <= low_max): SELL-64 slices, 64 rows per
//                  wave, one lane per row, column-major so each step of the
//                  wave is one coalesced 256-byte load; each lane sums its row
//                  sequentially in the reference's edge order;
//   * zero rows: no gather.
// One sweep is two launches: the sweep kernel (gather + update + next
// contribution + FP64 residual partials per workgroup) and the finalize kernel
// (hub rows, fixed-order reduction, convergence decided on the device).  No
// device-scope fence is used: on gfx950 one compiles to an L2 write-back plus
// invalidate (buffer_wbl2 / buffer_inv), which evicts the cached contribution
// vector; the kernel boundary orders the two launches instead.
#include <hip/hip_fp16.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "ag_extrap.h"
#include "ag_plan.h"

namespace ag {

namespace {

constexpr int kBlock = 256;
constexpr int kSweepBlock = 1024; // persistent sweep workgroup (16 waves)
constexpr int kSweepWaves = kSweepBlock / 64;
constexpr uint32_t kWaveRowMin = 1024; // in-degree >= this: 64 lanes per row
constexpr uint32_t kHalfRowMin = 256;  // >= this: 32 lanes per row; else 16
constexpr int kZeroRowsPerUnit = kSweepBlock * 4;
constexpr uint32_t kLdsBytes = 160 * 1024;   // LDS per CU on gfx950
constexpr uint32_t kStaticLds = 2048;        // static shared arrays of the sweep
constexpr uint32_t kFmtF32 = 0, kFmtBF16 = 1; // 2 = FP16
#ifndef AG_PR_DEPTH
#define AG_PR_DEPTH 8
#endif
#ifndef AG_PR_SELL
#define AG_PR_SELL 8
#endif
constexpr uint32_t kDdFlag = 0x80000000u; // DATA_DRIVEN: "updated" mark on a contribution
constexpr int kDepth = AG_PR_DEPTH; // loads in flight per lane, CSR rows (power of two)
constexpr int kSell = AG_PR_SELL;   // loads in flight per lane, SELL rows

} // namespace

void plan_destroy(Plan *p) { delete p; }

// Device-side sweep state.
struct SweepState {
    uint32_t done;
    uint32_t ticket;
    uint32_t iterations;
    uint32_t sweeps;
    uint32_t last_active;
    uint32_t fp32_switch;
    uint32_t since; // personalized, extrapolating: plain sweeps since the last extrapolated step
    uint32_t pad1;
    double last_err;
    double prev_err;
    double dangling;
    float tele;
    float w;        // > 0: the next sweep stores x + w (T(x) - x) instead of T(x)
    double rho;     // last residual ratio
};

struct SweepArgs {
    uint32_t N;
    uint32_t n_hub, n_w, n_h, n_nonlow, n_nonzero; // bin boundaries (labels)
    uint32_t nb_hub, nb_w, nb_h, nb_q, nb_low, nb_zero, nvblocks; // work units per bin
    uint32_t hub_chunk, hub_chunks;
    // FP32 REFERENCE mode (and DATA_DRIVEN with epsilon > 0): zero in-degree
    // rows keep pr = base and contribution base/outdeg in both buffers
    // (written at init); their |delta| is 0, so they count as active exactly
    // when epsilon <= 0.
    uint32_t skip_zero;
    float f16_scale, f16_inv; // FP16 storage scale (power of two) and its inverse
    uint32_t mode;
    uint32_t max_iter;
    uint32_t finish_fp32;
    uint32_t extrap; // personalized mode: Aitken-type extrapolated steps (st->w)
    float damp, base;
    double epsilon;
    double inv_n;
    const uint32_t *__restrict__ roff;
    const uint32_t *__restrict__ cols;
    const uint32_t *__restrict__ slice_ptr;
    const uint32_t *__restrict__ outdeg;
    const uint32_t *__restrict__ chunk_row;
    const uint32_t *__restrict__ chunk_first;
    float *hub_partial;
    uint32_t nparts;    // sweep workgroups (partials written per sweep)
    uint32_t hot_bytes; // LDS bytes of the contribution head per workgroup
    float *pr;
    void *c0, *c1; // contribution ping-pong buffers (sized for FP32)
    double *blk_err, *blk_dang;
    uint32_t *blk_act;
    SweepState *st;
    double *err_iter;
    uint32_t *act_iter;
    uint32_t *fmt; // fmt[j] = storage format of the contributions read by sweep j
    const uint32_t *seed_bits; // kModePPR: seed bitmap / labels / weights
    const uint32_t *seed_v;
    const float *seed_w;
    uint32_t nseeds;
    // partitioned-table layout (TabPlan; see k_pr_tables / k_pr_combine)
    uint32_t tab_T, tab_nt, tab_NT, tab_cold_base4, tab_nwaves, tab_lane_max4;
    uint32_t tab_cold_unit0, tab_nunits; // units [cold_unit0, nunits) are the cold table's
    uint32_t tab_cold_lane_u0, tab_cold_lane_s0, tab_cold_lane_s1; // cold lane units: unit index
                                         // of the first, its first sub-row, the cold table's end
    uint32_t tab_row_blocks;             // combine blocks over the non-hub rows
    uint32_t tab_hub_blocks;             // combine blocks over the hub rows (one wave each)
    uint32_t fin_hubs;                   // k_pr_finalize sums the hub rows (row layout)
    const uint16_t *tab_idx16;
    const uint32_t *tab_idx32;
    const uint32_t *tab_sub_start;
    float *tab_part;
    const uint4 *tab_units;
    const uint32_t *tab_wg_begin;
    const uint4 *tab_items;
    const uint4 *tab_hub_items;
    const uint32_t *tab_wtab;
};

namespace {

__device__ __forceinline__ float load_c(const float *p, uint32_t i) { return p[i]; }

// Element i of a device array addressed by a 32-bit byte offset from the
// array's base: with the base in scalar registers a load or store then takes
// one 32-bit offset per lane instead of a 64-bit address.  Callers keep every
// such array below 4 GiB (the partitioned-table layout checks its arrays).
template <typename T> __device__ __forceinline__ const T &at32(const T *base, uint32_t i)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + i * (uint32_t)sizeof(T));
}
template <typename T> __device__ __forceinline__ T &at32(T *base, uint32_t i)
{
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + i * (uint32_t)sizeof(T));
}
__device__ __forceinline__ float load_c(const bf16_t *p, uint32_t i)
{
    return __uint_as_float((uint32_t)p[i].x << 16);
}
__device__ __forceinline__ float load_c(const __half *p, uint32_t i) { return __half2float(p[i]); }
// The same through a 32-bit byte offset (contribution buffers stay below 4 GiB).
__device__ __forceinline__ float load_c32(const float *p, uint32_t i) { return at32(p, i); }
__device__ __forceinline__ float load_c32(const bf16_t *p, uint32_t i)
{
    return __uint_as_float((uint32_t)at32(p, i).x << 16);
}
__device__ __forceinline__ float load_c32(const __half *p, uint32_t i) { return __half2float(at32(p, i)); }

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f)
{
    const uint32_t b = __float_as_uint(f);
    return (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}

// FP16 words hold c * scale (a power of two, so scaling is exact), which moves
// contributions of size ~1/N into FP16's normal range.
__device__ __forceinline__ void store_c(void *buf, uint32_t fmt, uint32_t i, float c,
                                        float f16_scale)
{
    if (fmt == kFmtF32)
        static_cast<float *>(buf)[i] = c;
    else if (fmt == kFmtBF16)
        static_cast<uint16_t *>(buf)[i] = f32_to_bf16_rne(c);
    else
        static_cast<__half *>(buf)[i] = __float2half_rn(c * f16_scale);
}

__device__ __forceinline__ uint32_t ldnt(const uint32_t *p)
{
    return __builtin_nontemporal_load(p);
}

__device__ __forceinline__ float wave_sum(float x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}
__device__ __forceinline__ double wave_sum(double x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// Sums by DPP lane moves (in registers, no LDS), each in a fixed order:
// quad_sum / row_sum leave the sum over the lane's group of 4 / 16 lanes in
// every lane of the group; wave_sum63 leaves the sum over all 64 lanes in
// lane 63 (row sums, then row 0+1 into row 1 and row 2+3 into row 3 by
// row_bcast:15, then rows 0+1 into row 3 by row_bcast:31).  Lanes a DPP move
// does not write (outside its row mask) add 0.  *_or: the same for OR.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWMASK, 0xf, false);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_f32(float x)
{
    return __uint_as_float(dpp_u32<CTRL, ROWMASK>(__float_as_uint(x)));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double x)
{
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = dpp_u32<CTRL, ROWMASK>((uint32_t)b), hi = dpp_u32<CTRL, ROWMASK>((uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
constexpr int kDppQuad1 = 0xB1, kDppQuad2 = 0x4E, kDppRor4 = 0x124, kDppRor8 = 0x128,
              kDppBcast15 = 0x142, kDppBcast31 = 0x143;
template <int CTRL, int ROWMASK, typename T> __device__ __forceinline__ T dpp_move(T x)
{
    if constexpr (sizeof(T) == 8) return dpp_f64<CTRL, ROWMASK>(x);
    else if constexpr (std::is_same<T, float>::value) return dpp_f32<CTRL, ROWMASK>(x);
    else return dpp_u32<CTRL, ROWMASK>(x);
}
template <typename T> __device__ __forceinline__ T quad_sum(T x)
{
    x += dpp_move<kDppQuad1, 0xf>(x);
    x += dpp_move<kDppQuad2, 0xf>(x);
    return x;
}
template <typename T> __device__ __forceinline__ T row_sum(T x)
{
    x = quad_sum(x);
    x += dpp_move<kDppRor4, 0xf>(x);
    x += dpp_move<kDppRor8, 0xf>(x);
    return x;
}
template <typename T> __device__ __forceinline__ T wave_sum63(T x)
{
    x = row_sum(x);
    x += dpp_move<kDppBcast15, 0xa>(x);
    x += dpp_move<kDppBcast31, 0xc>(x);
    return x;
}
__device__ __forceinline__ uint32_t quad_or(uint32_t x)
{
    x |= dpp_u32<kDppQuad1, 0xf>(x);
    x |= dpp_u32<kDppQuad2, 0xf>(x);
    return x;
}
__device__ __forceinline__ uint32_t row_or(uint32_t x)
{
    x = quad_or(x);
    x |= dpp_u32<kDppRor4, 0xf>(x);
    x |= dpp_u32<kDppRor8, 0xf>(x);
    return x;
}
__device__ __forceinline__ uint32_t wave_or63(uint32_t x)
{
    x = row_or(x);
    x |= dpp_u32<kDppBcast15, 0xa>(x);
    x |= dpp_u32<kDppBcast31, 0xc>(x);
    return x;
}

// Vertex update for label i with gathered sum s (pageRank.c:1055-1070 in
// REFERENCE mode; x' = tele + d*s in STANDARD mode), then the contribution the
// next sweep reads (pageRank.c:1027-1032).  Written without contraction so
// that base + damp*next rounds twice, as the reference's C does.
//
// DATA_DRIVEN (pageRank.c:2886-2926): `flag` has bit 31 set when the row is
// on this sweep's worklist, i.e. some in-neighbour took a new rank in the
// previous sweep (or this is the first sweep).  Such a row adds |delta|/N to
// the error; it takes the new rank only when |delta| >= epsilon, and then its
// contribution carries bit 31 (the sign bit; contributions are >= 0) so that
// the next sweep's gathers find its out-neighbours.  That is the reference's
// workListNext marking over forward edges, read from the pull side.
// update_vertex_v takes the old rank pr[i] and the out-degree outdeg[i] as
// arguments (a caller can load them before its gather); update_vertex loads
// them itself.
__device__ __forceinline__ void update_vertex_v(const SweepArgs &a, uint32_t i, float s,
                                                float old, uint32_t od, float tele, void *cout,
                                                uint32_t wfmt, double &err, uint32_t &act,
                                                double &dang, uint32_t flag = 0u)
{
#pragma clang fp contract(off)
    if (a.mode == kModeSpmv) { // SPMV pull: y[i] = gathered sum
        a.pr[i] = s;
        return;
    }
    if (a.mode == AG_PR_DATA_DRIVEN) {
        float keep = old;
        uint32_t mark = 0u;
        if (flag & kDdFlag) {
            const float nw = a.base + a.damp * s;
            const double e = fabs((double)(nw - old));
            err += e; // divided by N once, in k_pr_finalize
            if (e >= a.epsilon) {
                a.pr[i] = nw;
                keep = nw;
                act += 1u;
                mark = kDdFlag;
            }
        }
        const float c = od ? keep / (float)od : 0.0f;
        static_cast<uint32_t *>(cout)[i] = __float_as_uint(c) | mark;
        return;
    }
    float nw;
    if (a.mode == AG_PR_REFERENCE) {
        nw = a.base + a.damp * s;
    } else if (a.mode == kModePPR) {
        // x' = d * gathered + s_v * ((1 - d) + d * dangling); tele holds the
        // parenthesised factor, s_v is 0 off the seed set.
        float sv = 0.0f;
        if ((a.seed_bits[i >> 5] >> (i & 31)) & 1u)
            for (uint32_t q = 0; q < a.nseeds; q++)
                if (a.seed_v[q] == i) sv += a.seed_w[q];
        const float gathered = a.damp * s;
        nw = gathered + sv * tele;
    } else {
        nw = tele + a.damp * s;
    }
    // stored iterate: T(x), or on an extrapolated sweep x + w (T(x) - x)
    float xo = nw;
    if (a.extrap) {
        const float w = a.st->w;
        if (w != 0.0f) xo = nw + w * (nw - old);
    }
    a.pr[i] = xo;
    if (a.mode == AG_PR_REFERENCE) {
        const double e = fabs((double)(nw - old));
        err += e; // divided by N once, in k_pr_finalize
        act += (e >= a.epsilon) ? 1u : 0u;
    } else {
        err += fabs((double)nw - (double)old);
        if (od == 0) dang += (double)xo;
    }
    const float c = od ? xo / (float)od : 0.0f;
    store_c(cout, wfmt, i, c, a.f16_scale);
}

__device__ __forceinline__ void update_vertex(const SweepArgs &a, uint32_t i, float s,
                                              float tele, void *cout, uint32_t wfmt,
                                              double &err, uint32_t &act, double &dang,
                                              uint32_t flag = 0u)
{
    const bool spmv = a.mode == kModeSpmv;
    update_vertex_v(a, i, s, spmv ? 0.0f : a.pr[i], spmv ? 0u : a.outdeg[i], tele, cout, wfmt, err,
                    act, dang, flag);
}

template <typename CT> struct FmtOf;
template <> struct FmtOf<float> { static constexpr uint32_t v = kFmtF32; };
template <> struct FmtOf<bf16_t> { static constexpr uint32_t v = kFmtBF16; };
template <> struct FmtOf<__half> { static constexpr uint32_t v = 2u; };

// Gather source: labels below `hot` read the workgroup's LDS copy of the
// contribution vector's head, the rest read global memory.
template <typename CT>
__device__ __forceinline__ float gather(const CT *__restrict__ g, uint32_t u, uint32_t hot)
{
    extern __shared__ __align__(16) unsigned char hot_raw[];
    const CT *lt = reinterpret_cast<const CT *>(hot_raw);
    return u < hot ? load_c(lt, u) : load_c(g, u);
}

// Adds one gathered contribution.  DATA_DRIVEN contributions carry the
// "updated" mark in the sign bit: the sum takes |x| (a free source modifier)
// and the marks are ORed.
template <bool DD>
__device__ __forceinline__ void acc(float &s, uint32_t &fl, float x)
{
    if (DD) {
        fl |= __float_as_uint(x);
        s += __builtin_fabsf(x);
    } else {
        s += x;
    }
}

// Rows [.., row_end) with G lanes per row (G = 16, 32 or 64): lanes of a
// group stride over the row's edges, kDepth loads in flight each, then a
// shuffle reduction inside the group and one lane updates the vertex.
// `first` is kDdFlag on a DATA_DRIVEN run's first sweep (every row is on the
// worklist), else 0.
template <int G, typename CT, bool DD>
__device__ __forceinline__ void group_rows(const SweepArgs &a, uint32_t row, uint32_t row_end,
                                           const CT *__restrict__ cin, void *cout, uint32_t wfmt,
                                           float tele, uint32_t hot, float rscale, double &err,
                                           uint32_t &act, double &dang, int lane, uint32_t first)
{
    const int gl = lane & (G - 1);
    const bool valid = row < row_end;
    uint32_t fl = first;
    float sp[kDepth];
#pragma unroll
    for (int t = 0; t < kDepth; t++) sp[t] = 0.f;
    if (valid) {
        const uint32_t rb = a.roff[row], re = a.roff[row + 1];
        // indices of the next batch are in flight while the current batch gathers
        uint32_t j = rb + gl;
        uint32_t c[kDepth];
#pragma unroll
        for (int t = 0; t < kDepth; t++) c[t] = j + t * G < re ? ldnt(a.cols + j + t * G) : 0u;
        for (; j < re; j += kDepth * G) {
            const uint32_t jn = j + kDepth * G;
            uint32_t n[kDepth];
#pragma unroll
            for (int t = 0; t < kDepth; t++) n[t] = jn + t * G < re ? ldnt(a.cols + jn + t * G) : 0u;
#pragma unroll
            for (int t = 0; t < kDepth; t++)
                if (j + t * G < re) acc<DD>(sp[t], fl, gather(cin, c[t], hot));
#pragma unroll
            for (int t = 0; t < kDepth; t++) c[t] = n[t];
        }
    }
    // pairwise over the partial sums: ((s0 + s1) + (s2 + s3)) + ...
#pragma unroll
    for (int w = 1; w < kDepth; w <<= 1)
#pragma unroll
        for (int t = 0; t + w < kDepth; t += 2 * w) sp[t] += sp[t + w];
    float s = sp[0];
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (DD) {
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) fl |= __shfl_xor(fl, off, 64);
    }
    if (valid && gl == 0)
        update_vertex(a, row, s * rscale, tele, cout, wfmt, err, act, dang, fl);
}

template <typename CT, bool DD>
__device__ void sweep_body(const SweepArgs &a, uint32_t wfmt, const CT *__restrict__ cin,
                           void *cout, float tele, uint32_t hot, double &err, uint32_t &act,
                           double &dang, uint32_t first)
{
    // FP16 words hold contributions times f16_scale: gathered sums scale back
    const float rscale = FmtOf<CT>::v == 2u ? a.f16_inv : 1.0f;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t u_w = a.nb_hub, u_h = u_w + a.nb_w, u_q = u_h + a.nb_h, u_l = u_q + a.nb_q,
                   u_z = u_l + a.nb_low;

    for (uint32_t vb = blockIdx.x; vb < a.nvblocks; vb += gridDim.x) {
        if (vb < u_w) {
            // ---- hub rows: one wave per chunk of hub_chunk edges ----
            const uint32_t c = vb * kSweepWaves + w;
            if (c < a.hub_chunks) {
                const uint32_t row = a.chunk_row[c];
                const uint32_t k = c - a.chunk_first[row];
                const uint32_t b0 = a.roff[row] + k * a.hub_chunk;
                const uint32_t b1 = min(a.roff[row + 1], b0 + a.hub_chunk);
                float sp[kDepth];
#pragma unroll
                for (int t = 0; t < kDepth; t++) sp[t] = 0.f;
                uint32_t fl = 0u;
                uint32_t j = b0 + lane;
                uint32_t ci[kDepth];
#pragma unroll
                for (int t = 0; t < kDepth; t++) ci[t] = j + t * 64 < b1 ? ldnt(a.cols + j + t * 64) : 0u;
                for (; j < b1; j += kDepth * 64) {
                    const uint32_t jn = j + kDepth * 64;
                    uint32_t n[kDepth];
#pragma unroll
                    for (int t = 0; t < kDepth; t++) n[t] = jn + t * 64 < b1 ? ldnt(a.cols + jn + t * 64) : 0u;
#pragma unroll
                    for (int t = 0; t < kDepth; t++)
                        if (j + t * 64 < b1) acc<DD>(sp[t], fl, gather(cin, ci[t], hot));
#pragma unroll
                    for (int t = 0; t < kDepth; t++) ci[t] = n[t];
                }
#pragma unroll
                for (int w2 = 1; w2 < kDepth; w2 <<= 1)
#pragma unroll
                    for (int t = 0; t + w2 < kDepth; t += 2 * w2) sp[t] += sp[t + w2];
                const float s = wave_sum(sp[0]);
                // chunk partial; k_pr_finalize sums a row's chunks in chunk order
                // (DATA_DRIVEN: the chunk's OR of marks rides in the sign bit)
                if (DD) {
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) fl |= __shfl_xor(fl, off, 64);
                    if (lane == 0)
                        a.hub_partial[c] = __uint_as_float(__float_as_uint(s) | (fl & kDdFlag));
                } else if (lane == 0) {
                    a.hub_partial[c] = s;
                }
            }
        } else if (vb < u_h) {
            const uint32_t row = a.n_hub + (vb - u_w) * kSweepWaves + w;
            group_rows<64, CT, DD>(a, row, a.n_w, cin, cout, wfmt, tele, hot, rscale, err, act, dang,
                                    lane, first);
        } else if (vb < u_q) {
            const uint32_t row = a.n_w + (vb - u_h) * (2 * kSweepWaves) + w * 2 + (lane >> 5);
            group_rows<32, CT, DD>(a, row, a.n_h, cin, cout, wfmt, tele, hot, rscale, err, act, dang,
                                    lane, first);
        } else if (vb < u_l) {
            const uint32_t row = a.n_h + (vb - u_q) * (4 * kSweepWaves) + w * 4 + (lane >> 4);
            group_rows<16, CT, DD>(a, row, a.n_nonlow, cin, cout, wfmt, tele, hot, rscale, err, act, dang,
                                    lane, first);
        } else if (vb < u_z) {
            // ---- low rows: SELL-64 slice per wave, lane per row, edges in order ----
            const uint32_t sl = (vb - u_l) * kSweepWaves + w;
            const uint32_t nlow = a.n_nonzero - a.n_nonlow;
            if (sl * 64 < nlow) {
                const uint32_t row = a.n_nonlow + sl * 64 + lane;
                const bool valid = row < a.n_nonzero;
                const uint32_t deg = valid ? a.roff[row + 1] - a.roff[row] : 0u;
                const uint32_t width = __shfl(deg, 0, 64); // rows sorted: lane 0 is the widest
                const uint32_t *p = a.cols + a.slice_ptr[sl] + lane;
                float s = 0.0f;
                uint32_t fl = first;
                // batch of kSell indices in flight while the previous batch
                // gathers; each lane adds in the reference's edge order
                uint32_t u[kSell];
#pragma unroll
                for (int t = 0; t < kSell; t++) u[t] = (uint32_t)t < deg ? ldnt(p + t * 64) : 0u;
                for (uint32_t k = 0; k < width; k += kSell) {
                    uint32_t nx[kSell];
#pragma unroll
                    for (int t = 0; t < kSell; t++)
                        nx[t] = k + kSell + t < deg ? ldnt(p + (k + kSell + t) * 64) : 0u;
                    float x[kSell];
#pragma unroll
                    for (int t = 0; t < kSell; t++) x[t] = k + t < deg ? gather(cin, u[t], hot) : 0.f;
#pragma unroll
                    for (int t = 0; t < kSell; t++)
                        if (k + t < deg) acc<DD>(s, fl, x[t]);
#pragma unroll
                    for (int t = 0; t < kSell; t++) u[t] = nx[t];
                }
                if (valid) update_vertex(a, row, s * rscale, tele, cout, wfmt, err, act, dang, fl);
            }
        } else if (!a.skip_zero) {
            // ---- zero in-degree rows ----
            const uint32_t r0 = a.n_nonzero + (vb - u_z) * kZeroRowsPerUnit;
#pragma unroll
            for (int q = 0; q < kZeroRowsPerUnit / kSweepBlock; q++) {
                const uint32_t row = r0 + q * kSweepBlock + threadIdx.x;
                if (row < a.N) update_vertex(a, row, 0.0f, tele, cout, wfmt, err, act, dang, first);
            }
        }
    }
}

// Copies the first `hot` contributions into LDS (16-byte loads).
template <typename CT>
__device__ __forceinline__ void load_hot(const CT *__restrict__ cin, uint32_t hot)
{
    extern __shared__ __align__(16) unsigned char hot_raw[];
    const uint32_t words = (uint32_t)(((uint64_t)hot * sizeof(CT) + 3) / 4);
    const uint4 *src = reinterpret_cast<const uint4 *>(cin);
    uint4 *dst = reinterpret_cast<uint4 *>(hot_raw);
    for (uint32_t i = threadIdx.x; i < words / 4; i += kSweepBlock) dst[i] = src[i];
    const uint32_t *sw = reinterpret_cast<const uint32_t *>(cin);
    uint32_t *dw = reinterpret_cast<uint32_t *>(hot_raw);
    for (uint32_t i = (words / 4) * 4 + threadIdx.x; i < words; i += kSweepBlock) dw[i] = sw[i];
    __syncthreads();
}

// Persistent sweep: one 1024-thread workgroup per CU slot walks the work units
// (hub chunks, mid rows, low slices, zero rows) with a grid stride.
// One kernel per storage format of the contributions it reads (keeps each
// kernel's register footprint to one instantiation).  A mixed-precision run
// launches the low-precision and the FP32 kernel every sweep; the one whose
// format does not match the sweep's exits at once.
template <typename CT, bool DD>
__global__ __launch_bounds__(kSweepBlock) void k_pr_sweep(SweepArgs a, uint32_t it)
{
    __shared__ double s_err[kSweepWaves], s_dang[kSweepWaves];
    __shared__ uint32_t s_act[kSweepWaves];
    SweepState *st = a.st;
    if (st->done) return; // converged in an earlier sweep: uniform early exit
    if (a.fmt[it] != FmtOf<CT>::v) return;
    const uint32_t wfmt = a.fmt[it + 1];
    const CT *cin = static_cast<const CT *>((it & 1) ? a.c1 : a.c0);
    void *cout = (it & 1) ? a.c0 : a.c1;
    const float tele = st->tele;
    double err = 0.0, dang = 0.0;
    uint32_t act = 0;
    const uint32_t hot = min(a.N, a.hot_bytes / (uint32_t)sizeof(CT));
    load_hot(cin, hot);
    sweep_body<CT, DD>(a, wfmt, cin, cout, tele, hot, err, act, dang, (DD && it == 0) ? kDdFlag : 0u);

    // ---- block partials; k_pr_finalize reduces them in a fixed order ----
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    err = wave_sum(err);
    dang = wave_sum(dang);
    act = wave_sum(act);
    if (lane == 0) { s_err[w] = err; s_dang[w] = dang; s_act[w] = act; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double e = 0.0, d = 0.0;
        uint32_t c = 0;
        for (int i = 0; i < kSweepWaves; i++) { e += s_err[i]; d += s_dang[i]; c += s_act[i]; }
        a.blk_err[blockIdx.x] = e;
        a.blk_dang[blockIdx.x] = d;
        a.blk_act[blockIdx.x] = c;
    }
}

// ---- partitioned-table sweep ------------------------------------------------
//
// k_pr_tables: each persistent workgroup walks its range of work units; a
// unit belongs to one table, and the workgroup (re)loads that table's
// contributions into LDS (converted to FP32, the sentinel slot T set to 0)
// when the table changes.  Unit kinds, chosen so that every sub-row's index
// words are in flight at once (one memory round trip per sub-row):
//   kUnitLane: up to 4096 consecutive sub-rows, one lane each, kLaneRows per
//              lane (sub-rows longer than lane_max entries are skipped: their
//              own units cover them);
//   kUnitG4/G16/G64: sub-rows from the items list, 4 / 16 / 64 lanes each
//              (up to 64 / 256 / 1024 entries in one batch of loads);
//   kUnitHub: hub-row chunks, one wave each, partial to hub_partial.
// A lane sums its sub-row sequentially in edge order (4 indices per load);
// lane groups keep per-lane partials over a fixed stride and reduce them
// pairwise, then by shuffles.  Every partial sum's order is fixed by the
// layout alone.  DATA_DRIVEN partials carry the OR of the marks in the sign bit.
constexpr uint32_t kUnitLane = 0, kUnitG4 = 1, kUnitG16 = 2, kUnitG64 = 3, kUnitHub = 4;
constexpr int kFinBlock = 1024; // threads of the finalize workgroup
static_assert(kFinBlock == kSweepBlock, "the table kernel's workgroups run the finalize body");
__device__ void finalize_body(const SweepArgs &a, uint32_t it);
constexpr uint32_t kMaxTabNT = 63; // LDS tables + the cold table (tuning allows 62 LDS tables)
constexpr uint32_t kLaneRows = 4;                           // sub-rows per lane of a lane unit
constexpr uint32_t kLaneUnitRows = kSweepBlock * kLaneRows; // sub-rows per lane unit
constexpr int kTabDepth = 4;                    // index loads in flight per lane
constexpr uint32_t kLongItems = 2;              // long sub-rows per lane group of a unit

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Index words at position p (units of 4 entries): four u16 table-local
// indices packed in two words (LDS tables), or four u32 labels (cold).
template <bool COLD> struct IdxVec { typedef u32x2 T; };
template <> struct IdxVec<true> { typedef u32x4 T; };

template <bool COLD>
__device__ __forceinline__ typename IdxVec<COLD>::T tab_raw(const SweepArgs &a, uint32_t p)
{
    if constexpr (COLD)
        return __builtin_nontemporal_load(
            &at32(reinterpret_cast<const u32x4 *>(a.tab_idx32), p - a.tab_cold_base4));
    else
        return __builtin_nontemporal_load(&at32(reinterpret_cast<const u32x2 *>(a.tab_idx16), p));
}

template <bool COLD>
__device__ __forceinline__ uint32_t tab_idx(const typename IdxVec<COLD>::T &v, int q)
{
    if constexpr (COLD)
        return q == 0 ? v.x : q == 1 ? v.y : q == 2 ? v.z : v.w;
    else
        return q == 0 ? (v.x & 0xffffu) : q == 1 ? (v.x >> 16) : q == 2 ? (v.y & 0xffffu) : (v.y >> 16);
}

template <typename CT, bool COLD>
__device__ __forceinline__ float tab_gather(const CT *__restrict__ cin, const float *tab, uint32_t u)
{
    return COLD ? load_c32(cin, u) : tab[u];
}

// Adds the four entries of one index word, in order.
template <typename CT, bool DD, bool COLD>
__device__ __forceinline__ void tab_add4(float &s, uint32_t &fl, const CT *__restrict__ cin,
                                         const float *tab, const typename IdxVec<COLD>::T &v)
{
#pragma unroll
    for (int q = 0; q < 4; q++) acc<DD>(s, fl, tab_gather<CT, COLD>(cin, tab, tab_idx<COLD>(v, q)));
}

// Sum of positions [p0, p1) with stride `step` (units of 4 entries), kTabDepth
// partials per lane, combined pairwise.
template <typename CT, bool DD, bool COLD>
__device__ __forceinline__ float tab_strided_sum(const SweepArgs &a, const CT *__restrict__ cin,
                                                 const float *tab, uint32_t p0, uint32_t p1,
                                                 uint32_t step, uint32_t &fl)
{
    float sp[kTabDepth];
#pragma unroll
    for (int t = 0; t < kTabDepth; t++) sp[t] = 0.f;
    for (uint32_t p = p0; p < p1; p += kTabDepth * step) {
        typename IdxVec<COLD>::T v[kTabDepth];
#pragma unroll
        for (int t = 0; t < kTabDepth; t++)
            if (p + t * step < p1) v[t] = tab_raw<COLD>(a, p + t * step);
#pragma unroll
        for (int t = 0; t < kTabDepth; t++)
            if (p + t * step < p1) tab_add4<CT, DD, COLD>(sp[t], fl, cin, tab, v[t]);
    }
#pragma unroll
    for (int w = 1; w < kTabDepth; w <<= 1)
#pragma unroll
        for (int t = 0; t + w < kTabDepth; t += 2 * w) sp[t] += sp[t + w];
    return sp[0];
}

__device__ __forceinline__ float tab_mark(float s, uint32_t fl)
{
    return __uint_as_float(__float_as_uint(s) | (fl & kDdFlag));
}

// Lane sub-rows: sub-row first + (j*16 + w)*64 + lane of [first, end) for
// j < kLaneRows.  lane_bounds loads their index-word bounds; lane_sums adds
// each one sequentially in edge order (sub-rows longer than lane_max entries
// are skipped: their own units cover them) with the first kTabDepth index
// words of all of them in flight together.
__device__ __forceinline__ void lane_bounds(const SweepArgs &a, uint32_t first, uint32_t end,
                                            uint32_t (&p0)[kLaneRows], uint32_t (&p1)[kLaneRows])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (uint32_t j = 0; j < kLaneRows; j++) {
        const uint32_t k = first + (j * kSweepWaves + (uint32_t)w) * 64 + lane;
        p0[j] = p1[j] = 0;
        if (k < end) {
            p0[j] = at32(a.tab_sub_start, k);
            p1[j] = at32(a.tab_sub_start, k + 1);
        }
    }
}

template <typename CT, bool DD, bool COLD>
__device__ __forceinline__ void lane_sums(const SweepArgs &a, const CT *__restrict__ cin,
                                          const float *tab, uint32_t first,
                                          uint32_t (&p0)[kLaneRows], uint32_t (&p1)[kLaneRows],
                                          uint32_t dd0)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    bool mine[kLaneRows];
#pragma unroll
    for (uint32_t j = 0; j < kLaneRows; j++) {
        mine[j] = p1[j] > p0[j] && p1[j] - p0[j] <= a.tab_lane_max4;
        if (!mine[j]) p1[j] = p0[j]; // long sub-rows belong to their own units
    }
    typename IdxVec<COLD>::T v[kLaneRows][kTabDepth];
#pragma unroll
    for (uint32_t j = 0; j < kLaneRows; j++)
#pragma unroll
        for (int t = 0; t < kTabDepth; t++)
            if (p0[j] + t < p1[j]) v[j][t] = tab_raw<COLD>(a, p0[j] + t);
    float s[kLaneRows];
    uint32_t fl[kLaneRows];
#pragma unroll
    for (uint32_t j = 0; j < kLaneRows; j++) {
        s[j] = 0.0f;
        fl[j] = dd0;
#pragma unroll
        for (int t = 0; t < kTabDepth; t++)
            if (p0[j] + t < p1[j]) tab_add4<CT, DD, COLD>(s[j], fl[j], cin, tab, v[j][t]);
    }
    // the rest of the longer lane sub-rows, sequentially in edge order
#pragma unroll
    for (uint32_t j = 0; j < kLaneRows; j++) {
        for (uint32_t p = p0[j] + kTabDepth; p < p1[j]; p += kTabDepth) {
            typename IdxVec<COLD>::T x[kTabDepth];
#pragma unroll
            for (int t = 0; t < kTabDepth; t++)
                if (p + t < p1[j]) x[t] = tab_raw<COLD>(a, p + t);
#pragma unroll
            for (int t = 0; t < kTabDepth; t++)
                if (p + t < p1[j]) tab_add4<CT, DD, COLD>(s[j], fl[j], cin, tab, x[t]);
        }
    }
#pragma unroll
    for (uint32_t j = 0; j < kLaneRows; j++) {
        const uint32_t k = (j * kSweepWaves + (uint32_t)w) * 64 + lane;
        if (mine[j]) at32(a.tab_part, first + k) = DD ? tab_mark(s[j], fl[j]) : s[j];
    }
}

// A run of an LDS table's lane units: sub-rows [first, end) in chunks of
// kLaneUnitRows, each chunk's index-word bounds loaded while the previous
// chunk's index words are in flight.
template <typename CT, bool DD>
__device__ void lane_run(const SweepArgs &a, const CT *__restrict__ cin, const float *tab,
                         uint32_t first, uint32_t end, uint32_t dd0)
{
    uint32_t p0[kLaneRows], p1[kLaneRows];
    lane_bounds(a, first, end, p0, p1);
    for (uint32_t c = first; c < end; c += kLaneUnitRows) {
        uint32_t q0[kLaneRows], q1[kLaneRows];
        lane_bounds(a, c + kLaneUnitRows, end, q0, q1);
        lane_sums<CT, DD, false>(a, cin, tab, c, p0, p1, dd0);
#pragma unroll
        for (uint32_t j = 0; j < kLaneRows; j++) {
            p0[j] = q0[j];
            p1[j] = q1[j];
        }
    }
}

// Long sub-rows, G lanes each, kLongItems per lane group: item = (sub-row,
// first position, end position); group grp of the workgroup takes items
// first + r*ngrp + grp of [first, end).  long_items loads them (zero items
// past `end`); long_sums has the first kTabDepth words per lane of every item
// in flight together; lane gl of a group adds positions gl + t*G
// (+ k*kTabDepth*G) into its partial t, the partials are combined pairwise
// and then over the group by shuffles.
__device__ __forceinline__ void long_items(const SweepArgs &a, uint32_t first, uint32_t end,
                                           uint32_t G, uint4 (&it)[kLongItems])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t per_wave = 64u / G;
    const uint32_t grp = (uint32_t)w * per_wave + (uint32_t)lane / G;
    const uint32_t ngrp = kSweepWaves * per_wave;
#pragma unroll
    for (uint32_t r = 0; r < kLongItems; r++) {
        const uint32_t k = first + r * ngrp + grp;
        it[r] = k < end ? at32(a.tab_items, k) : make_uint4(0u, 0u, 0u, 0u);
    }
}

template <typename CT, bool DD, bool COLD>
__device__ __forceinline__ void long_sums(const SweepArgs &a, const CT *__restrict__ cin,
                                          const float *tab, uint32_t G,
                                          const uint4 (&it)[kLongItems], uint32_t dd0)
{
    const int lane = threadIdx.x & 63;
    const uint32_t gl = (uint32_t)lane & (G - 1);
    typename IdxVec<COLD>::T v[kLongItems][kTabDepth];
#pragma unroll
    for (uint32_t r = 0; r < kLongItems; r++)
#pragma unroll
        for (uint32_t t = 0; t < kTabDepth; t++) {
            const uint32_t p = it[r].y + gl + t * G;
            if (p < it[r].z) v[r][t] = tab_raw<COLD>(a, p);
        }
#pragma unroll
    for (uint32_t r = 0; r < kLongItems; r++) {
        float sp[kTabDepth];
        uint32_t fl = dd0;
#pragma unroll
        for (uint32_t t = 0; t < kTabDepth; t++) {
            sp[t] = 0.0f;
            if (it[r].y + gl + t * G < it[r].z) tab_add4<CT, DD, COLD>(sp[t], fl, cin, tab, v[r][t]);
        }
        // the rest of a sub-row longer than kTabDepth * G words
        for (uint32_t p = it[r].y + gl + kTabDepth * G; p < it[r].z; p += kTabDepth * G) {
            typename IdxVec<COLD>::T x[kTabDepth];
#pragma unroll
            for (uint32_t t = 0; t < kTabDepth; t++)
                if (p + t * G < it[r].z) x[t] = tab_raw<COLD>(a, p + t * G);
#pragma unroll
            for (uint32_t t = 0; t < kTabDepth; t++)
                if (p + t * G < it[r].z) tab_add4<CT, DD, COLD>(sp[t], fl, cin, tab, x[t]);
        }
#pragma unroll
        for (int q = 1; q < kTabDepth; q <<= 1)
#pragma unroll
            for (int t = 0; t + q < kTabDepth; t += 2 * q) sp[t] += sp[t + q];
        float sum = sp[0];
        for (uint32_t off = G / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, (int)off, 64);
        if (DD)
            for (uint32_t off = G / 2; off > 0; off >>= 1) fl |= __shfl_xor(fl, (int)off, 64);
        // an item past the range is all zero: it.z == 0 stores nothing
        if (it[r].z != 0u && gl == 0) at32(a.tab_part, it[r].x) = DD ? tab_mark(sum, fl) : sum;
    }
}

template <typename CT, bool DD, bool COLD>
__device__ void tab_unit(const SweepArgs &a, const CT *__restrict__ cin, const float *tab,
                         uint32_t kind, uint32_t first, uint32_t count, uint32_t dd0)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (kind == kUnitLane) {
        uint32_t p0[kLaneRows], p1[kLaneRows];
        lane_bounds(a, first, first + count, p0, p1);
        lane_sums<CT, DD, COLD>(a, cin, tab, first, p0, p1, dd0);
        return;
    }
    if (kind == kUnitHub) {
        // one wave per chunk: item = (first position, end position, slot)
        if ((uint32_t)w >= count) return;
        const uint4 hi = at32(a.tab_hub_items, first + w);
        uint32_t fl = 0u;
        float s = tab_strided_sum<CT, DD, COLD>(a, cin, tab, hi.x + lane, hi.y, 64, fl);
        s = wave_sum(s);
        if (DD) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) fl |= __shfl_xor(fl, off, 64);
        }
        if (lane == 0) at32(a.hub_partial, hi.z) = DD ? tab_mark(s, fl) : s;
        return;
    }
    // long sub-rows: G lanes each, kLongItems per lane group
    const uint32_t G = kind == kUnitG4 ? 4u : kind == kUnitG16 ? 16u : 64u;
    uint4 it[kLongItems];
    long_items(a, first, first + count, G, it);
    long_sums<CT, DD, COLD>(a, cin, tab, G, it, dd0);
}

// A run of an LDS table's long units of one class: items [first, end) in
// chunks of one unit, each chunk's items loaded while the previous chunk's
// index words are in flight.
template <typename CT, bool DD>
__device__ void long_run(const SweepArgs &a, const CT *__restrict__ cin, const float *tab,
                         uint32_t kind, uint32_t first, uint32_t end, uint32_t dd0)
{
    const uint32_t G = kind == kUnitG4 ? 4u : kind == kUnitG16 ? 16u : 64u;
    const uint32_t chunk = kSweepWaves * (64u / G) * kLongItems;
    uint4 it[kLongItems];
    long_items(a, first, end, G, it);
    for (uint32_t c = first; c < end; c += chunk) {
        uint4 nx[kLongItems];
        long_items(a, c + chunk, end, G, nx);
        long_sums<CT, DD, false>(a, cin, tab, G, it, dd0);
#pragma unroll
        for (uint32_t r = 0; r < kLongItems; r++) it[r] = nx[r];
    }
}

// Loads table contributions [base, base + n) into LDS as FP32: every thread
// issues all of its 16-byte loads before the first LDS store (one global
// latency per table load).  base is a multiple of 8 entries.
constexpr uint32_t kTabEntries = 40440; // sources per LDS table (multiple of 8)
template <typename CT>
__device__ __forceinline__ void tab_load(float *tab, const CT *__restrict__ cin, uint32_t base,
                                         uint32_t n)
{
    constexpr uint32_t per = 16 / sizeof(CT); // entries per 16 bytes
    constexpr uint32_t kU = (kTabEntries / per + kSweepBlock - 1) / kSweepBlock;
    const uint32_t nvec = n / per;
    const u32x4 *src = reinterpret_cast<const u32x4 *>(cin + base);
    for (uint32_t i0 = threadIdx.x; i0 < nvec; i0 += kU * kSweepBlock) {
        u32x4 v[kU];
#pragma unroll
        for (uint32_t k = 0; k < kU; k++)
            if (i0 + k * kSweepBlock < nvec) v[k] = src[i0 + k * kSweepBlock];
#pragma unroll
        for (uint32_t k = 0; k < kU; k++) {
            const uint32_t i = i0 + k * kSweepBlock;
            if (i >= nvec) continue;
            if (sizeof(CT) == 4) {
                *reinterpret_cast<u32x4 *>(tab + i * per) = v[k];
            } else {
                const CT *h = reinterpret_cast<const CT *>(&v[k]);
#pragma unroll
                for (uint32_t q = 0; q < per; q++) tab[i * per + q] = load_c(h, q);
            }
        }
    }
    for (uint32_t i = nvec * per + threadIdx.x; i < n; i += kSweepBlock) tab[i] = load_c(cin, base + i);
}

template <typename CT, bool DD>
__global__ __launch_bounds__(kSweepBlock) void k_pr_tables(SweepArgs a, uint32_t it, uint32_t fin_prev)
{
    extern __shared__ __align__(16) unsigned char tab_lds[];
    float *tab = reinterpret_cast<float *>(tab_lds);
    SweepState *st = a.st;
    if (st->done) return;
    if (a.fmt[it] != FmtOf<CT>::v) return;
    const CT *cin = static_cast<const CT *>((it & 1) ? a.c1 : a.c0);
    const uint32_t dd0 = (DD && it == 0) ? kDdFlag : 0u;
    // the workgroup's static range of LDS units, then cold units (global
    // gathers: latency-bound, with costs the static split cannot predict)
    // taken one at a time from the sweep's counter (st->ticket: zeroed by
    // init and, for the next sweep, by the combine kernel).  With fin_prev,
    // ticket 0 is the previous sweep's finalize (its combine kernel has
    // ended: the kernel boundary orders the two), run by the first workgroup
    // done with its static range; this sweep's combine kernel starts after
    // it, so the state it sets (convergence, teleport, formats) is in place.
    uint32_t cur = 0xffffffffu;
    const uint32_t u0 = a.tab_wg_begin[blockIdx.x], u1 = a.tab_wg_begin[blockIdx.x + 1];
    for (uint32_t ui = u0; ui < u1; ui++) {
        const uint4 un = a.tab_units[ui];
        const uint32_t t = un.x >> 8, kind = un.x & 0xffu;
        if (t != cur) {
            // load table t: labels [t*T, min((t+1)*T, N)), then the zero sentinel
            __syncthreads();
            const uint32_t base = t * a.tab_T;
            tab_load<CT>(tab, cin, base, min(a.tab_T, a.N - base));
            if (threadIdx.x == 0) tab[a.tab_T] = 0.0f;
            __syncthreads();
            cur = t;
        }
        if (kind == kUnitLane) {
            // the run of this table's lane units inside [ui, u1): consecutive
            // units cover consecutive kLaneUnitRows-chunks of sub-rows up to
            // un.w, the end of the table's sub-rows
            const uint32_t n = min(u1 - ui, (un.w - un.y + kLaneUnitRows - 1) / kLaneUnitRows);
            lane_run<CT, DD>(a, cin, tab, un.y, min(un.w, un.y + n * kLaneUnitRows), dd0);
            ui += n - 1;
            continue;
        }
        if (kind != kUnitHub) {
            // likewise a run of one class's long units (items up to un.w)
            const uint32_t G = kind == kUnitG4 ? 4u : kind == kUnitG16 ? 16u : 64u;
            const uint32_t chunk = kSweepWaves * (64u / G) * kLongItems;
            const uint32_t n = min(u1 - ui, (un.w - un.y + chunk - 1) / chunk);
            long_run<CT, DD>(a, cin, tab, kind, un.y, min(un.w, un.y + n * chunk), dd0);
            ui += n - 1;
            continue;
        }
        tab_unit<CT, DD, false>(a, cin, tab, kind, un.y, un.z, dd0);
    }
    // cold phase, one unit per ticket (taking the next ticket before a unit's
    // work was slower: a workgroup then holds a unit it has not started
    // while others run out of work)
    __shared__ uint32_t s_next;
    for (;;) {
        __syncthreads(); // s_next is free again
        if (threadIdx.x == 0) s_next = atomicAdd(&st->ticket, 1u);
        __syncthreads();
        const uint32_t tk = s_next;
        if (fin_prev && tk == 0) {
            finalize_body(a, it - 1);
            continue;
        }
        const uint32_t ui = a.tab_cold_unit0 + tk - fin_prev;
        if (ui >= a.tab_nunits) break;
        if (ui >= a.tab_cold_lane_u0) {
            // cold lane units are consecutive kSweepBlock-sub-row ranges: no
            // descriptor load on the ticket's critical path
            const uint32_t f = a.tab_cold_lane_s0 + (ui - a.tab_cold_lane_u0) * kSweepBlock;
            tab_unit<CT, DD, true>(a, cin, tab, kUnitLane, f, min(kSweepBlock, a.tab_cold_lane_s1 - f), dd0);
        } else {
            const uint4 un = a.tab_units[ui];
            tab_unit<CT, DD, true>(a, cin, tab, un.x & 0xffu, un.y, un.z, dd0);
        }
    }
}

// k_pr_combine: one lane per non-hub row with in-edges (labels n_hub ..
// n_nonzero), one 64-row wave per hardware wave.  The wave's record for table
// t (the lanes with a sub-row in t, and the slot of the first of them) is a
// scalar load; a lane's slot is that base plus the count of lower lanes in
// the mask.  The row's sub-row partials are added in table order, then the
// vertex is updated as in the row layout.  The next blocks take the hub rows
// (one wave per row, over its chunk partials), the last ones the zero
// in-degree rows (unless skipped).  Each block writes its residual partials.
constexpr int kCombBlock = 256;
// Combine records read through the constant address space: uniform loads of
// memory the kernel never writes, so they become scalar loads.
typedef __attribute__((address_space(4))) const uint32_t ConstU32;
template <bool DD, uint32_t NTP, uint32_t GRP>
__global__ __launch_bounds__(kCombBlock) void k_pr_combine(SweepArgs a, uint32_t it)
{
    __shared__ double s_err[kCombBlock / 64], s_dang[kCombBlock / 64];
    __shared__ uint32_t s_act[kCombBlock / 64];
    SweepState *st = a.st;
    if (st->done) return;
    const uint32_t wfmt = a.fmt[it + 1];
    const float rscale = a.fmt[it] == 2u ? a.f16_inv : 1.0f;
    void *cout = (it & 1) ? a.c0 : a.c1;
    const float tele = st->tele;
    const uint32_t dd0 = (DD && it == 0) ? kDdFlag : 0u;
    double err = 0.0, dang = 0.0;
    uint32_t act = 0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (blockIdx.x < a.tab_row_blocks) {
        const uint32_t wv = __builtin_amdgcn_readfirstlane(blockIdx.x * (kCombBlock / 64) + w);
        if (wv < a.tab_nwaves) {
            const uint32_t r = a.n_hub + wv * 64 + lane;
            // the update's loads go out with the partial loads below
            const bool live = r < a.n_nonzero;
            const float old = live && a.mode != kModeSpmv ? a.pr[r] : 0.0f;
            const uint32_t od = live && a.mode != kModeSpmv ? a.outdeg[r] : 0u;
            // the wave's records, in groups of GRP tables: GRP lane masks (two
            // words each), then GRP slot bases.  One group's records at a time
            // (a loop, not unrolled: their scalar registers stay within
            // budget); inside a group every partial load is issued before the
            // first add.  A lane without a sub-row in table t reads the slot
            // after the previous lane's (inside the table's range, or the
            // slack slot after the last table) and takes 0 instead.
            float s = 0.0f;
            uint32_t fl = dd0;
#pragma unroll 1
            for (uint32_t g0 = 0; g0 < NTP; g0 += GRP) {
                const ConstU32 *rec = (const ConstU32 *)(a.tab_wtab + (size_t)wv * (3 * NTP) + 3 * g0);
                uint64_t m[GRP];
#pragma unroll
                for (uint32_t t = 0; t < GRP; t++) m[t] = ((uint64_t)rec[2 * t + 1] << 32) | rec[2 * t];
                float pv[GRP];
#pragma unroll
                for (uint32_t t = 0; t < GRP; t++) {
                    const uint32_t slot = __builtin_amdgcn_mbcnt_hi(
                        (uint32_t)(m[t] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m[t], 0u));
                    const char *pt = reinterpret_cast<const char *>(a.tab_part + rec[2 * GRP + t]);
                    pv[t] = *reinterpret_cast<const float *>(pt + slot * 4u);
                }
#pragma unroll
                for (uint32_t t = 0; t < GRP; t++)
                    acc<DD>(s, fl, __builtin_amdgcn_inverse_ballot_w64(m[t]) ? pv[t] : 0.0f);
            }
            if (live) update_vertex_v(a, r, s * rscale, old, od, tele, cout, wfmt, err, act, dang, fl);
        }
    } else if (blockIdx.x < a.tab_row_blocks + a.tab_hub_blocks) {
        // hub rows, one wave each: lane k adds the row's chunk partials k,
        // k + 64, ... in order, then a DPP sum into lane 63 (a fixed order)
        const uint32_t row = (blockIdx.x - a.tab_row_blocks) * (kCombBlock / 64) + (uint32_t)w;
        if (row < a.n_hub) {
            const uint32_t f0 = a.chunk_first[row], f1 = a.chunk_first[row + 1];
            float sp = 0.0f;
            uint32_t fl = dd0;
            for (uint32_t q = f0 + lane; q < f1; q += 64) acc<DD>(sp, fl, a.hub_partial[q]);
            sp = wave_sum63(sp);
            if (DD) fl = wave_or63(fl);
            if (lane == 63) update_vertex(a, row, sp * rscale, tele, cout, wfmt, err, act, dang, fl);
        }
    } else if (!a.skip_zero) {
        const uint32_t i = (blockIdx.x - a.tab_row_blocks - a.tab_hub_blocks) * kCombBlock + threadIdx.x;
        if (i < a.N - a.n_nonzero)
            update_vertex(a, a.n_nonzero + i, 0.0f, tele, cout, wfmt, err, act, dang, dd0);
    }
    err = wave_sum63(err);
    dang = wave_sum63(dang);
    act = wave_sum63(act);
    if (lane == 63) { s_err[w] = err; s_dang[w] = dang; s_act[w] = act; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double e = 0.0, d = 0.0;
        uint32_t c = 0;
        for (int i = 0; i < kCombBlock / 64; i++) { e += s_err[i]; d += s_dang[i]; c += s_act[i]; }
        a.blk_err[blockIdx.x] = e;
        a.blk_dang[blockIdx.x] = d;
        a.blk_act[blockIdx.x] = c;
        if (blockIdx.x == 0) st->ticket = 0; // the next sweep's cold units
    }
}

// The end of a sweep, in one workgroup that a kernel boundary orders after
// every sweep workgroup (so no device-scope fence is needed): k_pr_finalize
// after each row-layout sweep and after the last sweep of a partitioned-layout
// batch, else inside the next sweep's table kernel.  In the row layout it sums
// each hub row's chunk partials in chunk order and updates the hub vertices
// (the partitioned layout's combine kernel does that); then it reduces all
// residual / active / dangling partials in a fixed order and decides
// convergence.
// Runs in one kFinBlock-thread workgroup: k_pr_finalize, or the table-kernel
// workgroup that takes the first ticket of the next sweep's cold phase.
__device__ void finalize_body(const SweepArgs &a, uint32_t it)
{
    __shared__ double s_err[kFinBlock / 64], s_dang[kFinBlock / 64];
    __shared__ uint32_t s_act[kFinBlock / 64];
    SweepState *st = a.st;
    const uint32_t wfmt = a.fmt[it + 1];
    const float rscale = a.fmt[it] == 2u ? a.f16_inv : 1.0f;
    void *cout = (it & 1) ? a.c0 : a.c1;
    const float tele = st->tele;
    double err = 0.0, dang = 0.0;
    uint32_t act = 0;
    const bool dd = a.mode == AG_PR_DATA_DRIVEN;
    // loads in batches of kFinBatch, added in order (one memory round trip
    // per batch instead of one per value)
    constexpr uint32_t kFinBatch = 8;
    for (uint32_t row = threadIdx.x; a.fin_hubs && row < a.n_hub; row += kFinBlock) {
        const uint32_t f0 = a.chunk_first[row], f1 = a.chunk_first[row + 1];
        float sum = 0.0f;
        uint32_t fl = (dd && it == 0) ? kDdFlag : 0u;
        for (uint32_t q0 = f0; q0 < f1; q0 += kFinBatch) {
            float v[kFinBatch];
#pragma unroll
            for (uint32_t k = 0; k < kFinBatch; k++) v[k] = q0 + k < f1 ? a.hub_partial[q0 + k] : 0.0f;
#pragma unroll
            for (uint32_t k = 0; k < kFinBatch; k++) {
                if (q0 + k >= f1) break;
                if (dd) acc<true>(sum, fl, v[k]);
                else sum += v[k];
            }
        }
        update_vertex(a, row, sum * rscale, tele, cout, wfmt, err, act, dang, fl);
    }
    for (uint32_t b0 = threadIdx.x; b0 < a.nparts; b0 += kFinBatch * kFinBlock) {
        double e[kFinBatch], d[kFinBatch];
        uint32_t c[kFinBatch];
#pragma unroll
        for (uint32_t k = 0; k < kFinBatch; k++) {
            const uint32_t b = b0 + k * kFinBlock;
            e[k] = b < a.nparts ? a.blk_err[b] : 0.0;
            d[k] = b < a.nparts ? a.blk_dang[b] : 0.0;
            c[k] = b < a.nparts ? a.blk_act[b] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kFinBatch; k++) {
            if (b0 + k * kFinBlock >= a.nparts) break;
            err += e[k];
            dang += d[k];
            act += c[k];
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    err = wave_sum(err);
    dang = wave_sum(dang);
    act = wave_sum(act);
    if (lane == 0) { s_err[w] = err; s_dang[w] = dang; s_act[w] = act; }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double E_ = 0.0, D_ = 0.0;
    uint32_t A_ = 0;
    for (int i = 0; i < kFinBlock / 64; i++) { E_ += s_err[i]; D_ += s_dang[i]; A_ += s_act[i]; }
    // REFERENCE and DATA_DRIVEN: error_total is the mean |delta|
    // (pageRank.c:1062-1063); the sum is divided by N once here
    if (a.mode == AG_PR_REFERENCE || dd) E_ = E_ / (double)a.N;
    if (a.skip_zero && a.epsilon <= 0.0) A_ += a.N - a.n_nonzero;
    a.err_iter[it] = E_;
    a.act_iter[it] = A_;
    st->prev_err = st->last_err;
    st->last_err = E_;
    st->last_active = A_;
    st->sweeps = it + 1;
    // Storage format of the contributions written by the next sweep.  A
    // low-precision run turns FP32 once its residual reaches the storage
    // format's rounding noise, d * 2^-(m+1) with m explicit mantissa bits
    // (BF16 7, FP16 10): round-to-nearest storage is a deterministic map, so
    // the rounded iteration converges to a fixed point of its own and its
    // residual keeps falling below that noise while the iterate stays up to
    // ~1e-4 from the true fixed point.  A stall (residual not falling) also
    // switches.
    uint32_t nf = a.fmt[it + 1];
    const double noise = (double)a.damp * (nf == kFmtBF16 ? 0x1p-8 : 0x1p-11);
    if (nf != kFmtF32 && a.finish_fp32 && it >= 2 &&
        (E_ <= noise || E_ > 0.97 * st->prev_err)) {
        nf = kFmtF32;
        st->fp32_switch = it + 2;
    }
    if (it + 2 < a.max_iter + 2) a.fmt[it + 2] = nf;
    if (a.mode == AG_PR_REFERENCE || dd) {
        if (A_ == 0) { st->done = 1; st->iterations = it; }
        else st->iterations = it + 1;
    } else if (a.mode == kModePPR) {
        st->dangling = D_;
        st->tele = (float)((1.0 - (double)a.damp) + (double)a.damp * D_);
        st->iterations = it + 1;
        // after an extrapolated step the stored iterate is not T(x): its
        // residual E_ certifies nothing, so no stop on it
        const bool stepped = st->w != 0.0f;
        st->w = 0.0f;
        if (E_ <= a.epsilon && !stepped) {
            st->done = 1;
        } else if (a.extrap) {
            st->w = extrapolation_weight(E_, it > 0 ? st->prev_err : 0.0, stepped, st->since, st->rho);
        }
    } else {
        st->dangling = D_;
        st->tele = (float)((1.0 - (double)a.damp) * a.inv_n + (double)a.damp * D_ * a.inv_n);
        st->iterations = it + 1;
        // extrapolated steps as in personalized mode: stop on a plain sweep's
        // residual only
        const bool stepped = st->w != 0.0f;
        st->w = 0.0f;
        if (a.epsilon > 0.0 && E_ <= a.epsilon && !stepped)
            st->done = 1;
        else if (a.extrap)
            st->w = extrapolation_weight(E_, it > 0 ? st->prev_err : 0.0, stepped, st->since, st->rho);
    }
}

__global__ __launch_bounds__(kFinBlock) void k_pr_finalize(SweepArgs a, uint32_t it)
{
    if (a.st->done) return;
    finalize_body(a, it);
}

// Sweep set-up: state, per-sweep formats, the zero after each contribution
// vector and (unless the caller's vector is the start, `warm`) the uniform or
// seeded start in label space.
__global__ void k_pr_init(SweepArgs a, float x0, float tele0, uint32_t fmt0, uint32_t warm)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        SweepState *st = a.st;
        st->done = 0; st->ticket = 0; st->iterations = 0; st->sweeps = 0;
        st->last_active = 0; st->fp32_switch = 0;
        st->last_err = 0.0; st->prev_err = 0.0; st->dangling = 0.0; st->tele = tele0;
        st->since = 0; st->w = 0.0f; st->rho = 0.0;
        if (a.mode == kModePPR) {
            // dangling mass of x0 = s: seed weight on vertices without out-edges
            double d0 = 0.0;
            for (uint32_t q = 0; q < a.nseeds; q++)
                if (a.outdeg[a.seed_v[q]] == 0) d0 += (double)a.seed_w[q];
            st->dangling = d0;
            st->tele = (float)((1.0 - (double)a.damp) + (double)a.damp * d0);
        }
    }
    if (i <= a.max_iter + 1) a.fmt[i] = fmt0;
    if (i == a.N) {
        // the zero after each contribution vector (read by the cold table's
        // padding): in the FP32 view, and in the 16-bit view when the run
        // starts in BF16 / FP16
        static_cast<float *>(a.c0)[a.N] = 0.0f;
        static_cast<float *>(a.c1)[a.N] = 0.0f;
        if (fmt0 != kFmtF32) {
            static_cast<uint16_t *>(a.c0)[a.N] = 0;
            static_cast<uint16_t *>(a.c1)[a.N] = 0;
        }
    }
    if (i >= a.N || warm) return;
    float x = x0;
    if (a.mode == kModePPR) {
        x = 0.0f;
        if ((a.seed_bits[i >> 5] >> (i & 31)) & 1u)
            for (uint32_t q = 0; q < a.nseeds; q++)
                if (a.seed_v[q] == i) x += a.seed_w[q];
    }
    a.pr[i] = x;
    const uint32_t od = a.outdeg[i];
    const float c = od ? x / (float)od : 0.0f;
    store_c(a.c0, fmt0, i, c, a.f16_scale);
    if (a.skip_zero && i >= a.n_nonzero) store_c(a.c1, fmt0, i, c, a.f16_scale);
}

// Warm start, after k_pr_init: vertex v's start xin[v] * xscale goes to its
// label rank[v] (rank and contribution), and each block writes the dangling
// mass of its vertices (FP64, fixed order) to part[block].  xscale is N in
// REFERENCE and DATA_DRIVEN modes (the caller's vector is pr / N) and 1 in
// STANDARD mode.
__global__ __launch_bounds__(kBlock) void k_pr_init_from(SweepArgs a, const float *__restrict__ xin,
                                                         const uint32_t *__restrict__ rank,
                                                         float xscale, uint32_t fmt0,
                                                         double *__restrict__ part)
{
#pragma clang fp contract(off)
    __shared__ double s[kBlock / 64];
    double dang = 0.0;
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v < a.N) {
        const uint32_t l = rank[v];
        const float x = xin[v] * xscale;
        a.pr[l] = x;
        const uint32_t od = a.outdeg[l];
        store_c(a.c0, fmt0, l, od ? x / (float)od : 0.0f, a.f16_scale);
        if (od == 0) dang = (double)x;
    }
    dang = wave_sum(dang);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = dang;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
}

// STANDARD warm start: the first sweep's teleport term from the start's
// dangling mass D (summed by k_sum_parts), in the form k_pr_finalize uses
// for later sweeps.
__global__ void k_pr_warm_tele(SweepArgs a, const double *__restrict__ dsum)
{
    if (threadIdx.x != 0 || a.mode != AG_PR_STANDARD) return;
    const double D = *dsum;
    a.st->tele = (float)((1.0 - (double)a.damp) * a.inv_n + (double)a.damp * D * a.inv_n);
}

// SPMV pull set-up: resets the sweep state, marks every sweep FP32, and loads
// the caller's input vector into label space (c0 = x[order[i]]).
__global__ void k_spmv_init(SweepArgs a, const float *__restrict__ x,
                            const uint32_t *__restrict__ order)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        SweepState *st = a.st;
        st->done = 0; st->ticket = 0; st->iterations = 0; st->sweeps = 0;
        st->last_active = 0; st->fp32_switch = 0;
        st->last_err = 0.0; st->prev_err = 0.0; st->dangling = 0.0; st->tele = 0.0f;
        st->since = 0; st->w = 0.0f; st->rho = 0.0;
    }
    if (i <= a.max_iter + 1) a.fmt[i] = kFmtF32;
    if (i < a.N) static_cast<float *>(a.c0)[i] = x[order[i]];
    if (i == a.N) static_cast<float *>(a.c0)[a.N] = 0.0f; // read by the cold table's padding
}

// Un-relabel (and, in REFERENCE mode, divide by N: pageRank.c:1079-1085) into
// the caller's order; per-block FP64 partial sums of the returned ranks.
// Output in vertex order: out[v] = pr[rank[v]] (coalesced writes; the
// label-order ranks are read at random, from cache), optionally / N, and
// per-block FP64 partial sums in vertex order.
__global__ __launch_bounds__(kBlock) void k_pr_finish(const float *__restrict__ pr,
                                                      const uint32_t *__restrict__ rank,
                                                      uint32_t N, uint32_t divide,
                                                      float *__restrict__ out,
                                                      double *__restrict__ part)
{
    __shared__ double s[kBlock / 64];
    double acc = 0.0;
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v < N) {
        float x = pr[rank[v]];
        if (divide) x = x / (float)N;
        out[v] = x;
        acc = (double)x;
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
}

// One workgroup: every thread adds its strided partials (kSumBatch loads in
// flight at a time), then the waves' sums in order.
constexpr uint32_t kSumBlock = 1024, kSumBatch = 8;
__global__ __launch_bounds__(kSumBlock) void k_sum_parts(const double *part, uint32_t n, double *out)
{
    __shared__ double s[kSumBlock / 64];
    double acc = 0.0;
    for (uint32_t b0 = threadIdx.x; b0 < n; b0 += kSumBatch * kSumBlock) {
        double v[kSumBatch];
#pragma unroll
        for (uint32_t k = 0; k < kSumBatch; k++) {
            const uint32_t b = b0 + k * kSumBlock;
            v[k] = b < n ? part[b] : 0.0;
        }
#pragma unroll
        for (uint32_t k = 0; k < kSumBatch; k++) acc += v[k];
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (uint32_t i = 0; i < kSumBlock / 64; i++) t += s[i];
        *out = t;
    }
}

// ---- plan construction kernels ----
__global__ void k_degree_keys(const uint32_t *__restrict__ indeg, uint32_t N, uint32_t maxdeg,
                              uint32_t *__restrict__ keys, uint32_t *__restrict__ ids)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    keys[v] = maxdeg - indeg[v];
    ids[v] = v;
}

// Relabels vertex order[i] as i and counts the bins.  Grid-strided; each
// wave counts with ballots, a workgroup sums its waves in LDS and adds its six
// counts with one atomic each (per-wave atomics on six shared counters
// serialise at the L2).
__global__ __launch_bounds__(256) void k_relabel_vertex(const uint32_t *__restrict__ order,
                                 const uint32_t *__restrict__ skeys, uint32_t maxdeg,
                                 const uint32_t *__restrict__ fwd_outdeg, uint32_t N,
                                 uint32_t *__restrict__ rank, uint32_t *__restrict__ rdeg,
                                 uint32_t *__restrict__ outdeg_r, uint32_t low_max,
                                 uint32_t hub_min, uint32_t *__restrict__ counts)
{
    __shared__ uint32_t s_cnt[4][6];
    uint32_t c[6] = {0, 0, 0, 0, 0, 0};
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < N; base += stride) {
        const uint32_t i = base + threadIdx.x;
        uint32_t d = 0, od = 1;
        if (i < N) {
            const uint32_t v = order[i];
            rank[v] = i;
            d = maxdeg - skeys[i];
            rdeg[i] = d;
            od = fwd_outdeg[v];
            outdeg_r[i] = od;
        }
        c[0] += (uint32_t)__popcll(__ballot(i < N && d >= hub_min));
        c[1] += (uint32_t)__popcll(__ballot(i < N && d > low_max));
        c[2] += (uint32_t)__popcll(__ballot(i < N && d > 0));
        c[3] += (uint32_t)__popcll(__ballot(i < N && od == 0));
        c[4] += (uint32_t)__popcll(__ballot(i < N && d >= kWaveRowMin));
        c[5] += (uint32_t)__popcll(__ballot(i < N && d >= kHalfRowMin));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; k++) s_cnt[w][k] = c[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const uint32_t t = s_cnt[0][threadIdx.x] + s_cnt[1][threadIdx.x] + s_cnt[2][threadIdx.x] +
                           s_cnt[3][threadIdx.x];
        // counts[]: 0 hub, 1 nonlow, 2 nonzero, 3 dangling, 4 wide, 5 half
        if (t) atomicAdd(&counts[threadIdx.x], t);
    }
}

__global__ void k_slice_width(const uint32_t *__restrict__ rdeg, uint32_t low_begin,
                              uint32_t nslices, uint32_t *__restrict__ wid)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nslices) wid[s] = 64u * rdeg[low_begin + s * 64];
}

__global__ void k_add_base(uint32_t *__restrict__ x, uint32_t n, uint32_t base)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += base;
}

// Edge-slot fills over rows [0, nrows) (every such row has an in-edge):
// slots [first, total), total <= roff[nrows], one thread per slot, whose row
// is the last r with roff[r] <= e.  Block b covers slots first + 256 b ..;
// k_block_rows gives each block its first row (the row holding its first
// slot: one writer per block), the block stages the offsets of the <= 256
// rows it spans in LDS, and each thread searches those (8 steps).  One wave
// per row made the largest hubs (~1.6e5 in-edges at R-MAT scale 22) a serial
// tail of ~2 ms; a binary search over all row offsets per block cost ~0.4 ms.
constexpr uint32_t kFillBlock = 256;
__global__ void k_block_rows(const uint32_t *__restrict__ roff, uint32_t nrows, uint32_t first,
                             uint32_t total, uint32_t *__restrict__ brow)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const uint32_t lo = max(roff[r], first), hi = min(roff[r + 1], total);
    if (lo >= hi) return;
    // blocks whose first slot lies in [lo, hi)
    for (uint32_t b = (lo - first + kFillBlock - 1) / kFillBlock; first + b * kFillBlock < hi; b++)
        brow[b] = r;
}
// The last k < nr with s_off[k] <= e (s_off[0] <= e < s_off[nr]).
__device__ __forceinline__ uint32_t staged_row(const uint32_t *s_off, uint32_t nr, uint32_t e)
{
    uint32_t lo = 0, hi = nr;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (s_off[mid] <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}
__global__ __launch_bounds__(kFillBlock) void k_fill_csr_edges(
    const uint32_t *__restrict__ order, const uint32_t *__restrict__ roff,
    const uint32_t *__restrict__ brow, const uint32_t *__restrict__ inv_idx,
    const uint32_t *__restrict__ inv_dst, const uint32_t *__restrict__ rank, uint32_t nrows,
    uint32_t first, uint32_t total, uint32_t *__restrict__ cols)
{
    __shared__ uint32_t s_off[kFillBlock + 1];
    const uint32_t e0 = first + blockIdx.x * kFillBlock;
    const uint32_t r0 = brow[blockIdx.x];
    // offsets of rows r0 .. r0 + 256 (clamped to nrows): the block's slots
    // end before roff[min(r0 + 256, nrows)]
    const uint32_t nr = min(kFillBlock, nrows - r0);
    for (uint32_t i = threadIdx.x; i <= nr; i += kFillBlock) s_off[i] = roff[r0 + i];
    __syncthreads();
    const uint32_t e = e0 + threadIdx.x;
    if (e >= total) return;
    const uint32_t lo = staged_row(s_off, nr, e);
    const uint32_t v = order[r0 + lo];
    cols[e] = rank[inv_dst[inv_idx[v] + (e - s_off[lo])]];
}

// SELL rows: one thread per row, four edges' loads in flight at a time.
// With ecol, the row's sources also go to ecol[roff[row] ..] in row order
// (the partitioned-table build reads them from there).
__global__ void k_fill_sell_rows(const uint32_t *__restrict__ order, const uint32_t *__restrict__ rdeg,
                                 const uint32_t *__restrict__ inv_idx,
                                 const uint32_t *__restrict__ inv_dst,
                                 const uint32_t *__restrict__ rank,
                                 const uint32_t *__restrict__ slice_ptr, uint32_t low_begin,
                                 uint32_t nlow, uint32_t *__restrict__ cols,
                                 const uint32_t *__restrict__ roff, uint32_t *__restrict__ ecol)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nlow) return;
    const uint32_t row = low_begin + t;
    const uint32_t v = order[row];
    const uint32_t n = rdeg[row];
    const uint32_t s = inv_idx[v];
    uint32_t *p = cols + slice_ptr[t / 64] + (t % 64);
    uint32_t *q = ecol ? ecol + roff[row] : nullptr;
    uint32_t k = 0;
    for (; k + 4 <= n; k += 4) {
        uint32_t d[4], r[4];
#pragma unroll
        for (int j = 0; j < 4; j++) d[j] = inv_dst[s + k + j];
#pragma unroll
        for (int j = 0; j < 4; j++) r[j] = rank[d[j]];
#pragma unroll
        for (int j = 0; j < 4; j++) p[(k + j) * 64] = r[j];
        if (q) {
#pragma unroll
            for (int j = 0; j < 4; j++) q[k + j] = r[j];
        }
    }
    for (; k < n; k++) {
        const uint32_t r = rank[inv_dst[s + k]];
        p[k * 64] = r;
        if (q) q[k] = r;
    }
}

// ---- partitioned-table layout construction ----

// Edges per table: a table holds the edges whose source label l has
// l / kTabEntries == t, so its count is the sum of those labels' out-degrees,
// for the candidate LDS tables 0 .. kTabHist-2 (hist[kTabHist-1], the rest,
// is never read and stays as the caller cleared it).  One workgroup per
// table, 8 loads in flight per thread, a plain store each.
constexpr int kTabHist = 64;
__global__ __launch_bounds__(1024) void k_tab_hist(const uint32_t *__restrict__ outdeg, uint32_t N,
                                                   uint32_t *__restrict__ hist)
{
    __shared__ uint32_t ws[16];
    const uint32_t t = blockIdx.x;
    if (t + 1 >= (uint32_t)kTabHist) return;
    const uint32_t l0 = (uint32_t)min<uint64_t>(N, (uint64_t)t * kTabEntries);
    const uint32_t l1 = (uint32_t)min<uint64_t>(N, (uint64_t)l0 + kTabEntries);
    uint32_t sum = 0;
    for (uint32_t l = l0 + threadIdx.x; l < l1; l += 8 * 1024) {
        uint32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = l + k * 1024u < l1 ? outdeg[l + k * 1024u] : 0u;
#pragma unroll
        for (int k = 0; k < 8; k++) sum += v[k];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int i = 0; i < 16; i++) tot += ws[i];
        hist[t] = tot;
    }
}

// Table of every label-order edge (the sort key of the table layout).
__global__ void k_tab_keys(const uint32_t *__restrict__ ecol, uint32_t E, uint32_t nt,
                           uint32_t *__restrict__ key)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    key[e] = min(ecol[e] / kTabEntries, nt);
}

// erow[e] = row of edge slot e < total (rows [0, nrows), each with an
// in-edge, total <= roff[nrows]): the row search of k_fill_csr_edges.
__global__ __launch_bounds__(kFillBlock) void k_edge_rows(const uint32_t *__restrict__ roff,
                                                          const uint32_t *__restrict__ brow,
                                                          uint32_t nrows, uint32_t total,
                                                          uint32_t *__restrict__ erow)
{
    __shared__ uint32_t s_off[kFillBlock + 1];
    const uint32_t e0 = blockIdx.x * kFillBlock;
    const uint32_t r0 = brow[blockIdx.x];
    const uint32_t nr = min(kFillBlock, nrows - r0);
    for (uint32_t i = threadIdx.x; i <= nr; i += kFillBlock) s_off[i] = roff[r0 + i];
    __syncthreads();
    const uint32_t e = e0 + threadIdx.x;
    if (e < total) erow[e] = r0 + staged_row(s_off, nr, e);
}

// 1 where a sub-row (table, row) starts in the table-sorted edge order.
__global__ void k_tab_flags(const uint32_t *__restrict__ skey, const uint32_t *__restrict__ srow,
                            uint32_t E, uint32_t *__restrict__ flag)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E) return;
    flag[i] = (i == 0 || skey[i] != skey[i - 1] || srow[i] != srow[i - 1]) ? 1u : 0u;
}

// Per sub-row: row, table, first table-sorted edge (sub_e0[S] = E).
__global__ void k_tab_subrows(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ fscan,
                              const uint32_t *__restrict__ skey, const uint32_t *__restrict__ srow,
                              uint32_t E, uint32_t *__restrict__ sub_row,
                              uint32_t *__restrict__ sub_tab, uint32_t *__restrict__ sub_e0)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E) return;
    if (flag[i]) {
        const uint32_t sid = fscan[i];
        sub_row[sid] = srow[i];
        sub_tab[sid] = skey[i];
        sub_e0[sid] = i;
    }
    if (i == E - 1) sub_e0[fscan[E]] = E;
}

// Padded length of each sub-row in units of 4 entries.
__global__ void k_tab_len4(const uint32_t *__restrict__ sub_e0, uint32_t S, uint32_t *__restrict__ len4)
{
    const uint32_t sid = blockIdx.x * blockDim.x + threadIdx.x;
    if (sid < S) len4[sid] = (sub_e0[sid + 1] - sub_e0[sid] + 3u) / 4u;
}

__global__ void k_fill_u16(uint16_t *__restrict__ p, uint64_t n, uint16_t v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u32(uint32_t *__restrict__ p, uint64_t n, uint32_t v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Source index of every table-sorted edge (scol: its source label) into its
// sub-row's slot.
__global__ void k_tab_fill_idx(const uint32_t *__restrict__ skey, const uint32_t *__restrict__ scol,
                               const uint32_t *__restrict__ fscan,
                               const uint32_t *__restrict__ sub_e0,
                               const uint32_t *__restrict__ sub_start, uint32_t E, uint32_t T,
                               uint32_t nt, uint32_t cold_base4, uint16_t *__restrict__ idx16,
                               uint32_t *__restrict__ idx32)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E) return;
    const uint32_t sid = fscan[i + 1] - 1u;
    const uint32_t t = skey[i];
    const uint32_t u = scol[i];
    const uint64_t pos = (uint64_t)sub_start[sid] * 4u + (i - sub_e0[sid]);
    if (t < nt)
        idx16[pos] = (uint16_t)(u - t * T);
    else
        idx32[pos - (uint64_t)cold_base4 * 4u] = u;
}

// First sub-row of every table that has one.
__global__ void k_tab_bounds(const uint32_t *__restrict__ sub_tab, uint32_t S, uint32_t *__restrict__ sb)
{
    const uint32_t sid = blockIdx.x * blockDim.x + threadIdx.x;
    if (sid >= S) return;
    if (sid == 0 || sub_tab[sid] != sub_tab[sid - 1]) sb[sub_tab[sid]] = sid;
}

// Class of a sub-row (its unit kind): hub row, else by padded length: one
// lane (up to lane_max entries), 4 lanes (up to 64), 16 lanes (up to 256),
// 64 lanes.  Key = table * 8 + class, value = sub-row.
__global__ void k_tab_classes(const uint32_t *__restrict__ sub_row, const uint32_t *__restrict__ sub_tab,
                              const uint32_t *__restrict__ len4, uint32_t S, uint32_t n_hub,
                              uint32_t lane_max4, uint32_t *__restrict__ key, uint32_t *__restrict__ val)
{
    const uint32_t sid = blockIdx.x * blockDim.x + threadIdx.x;
    if (sid >= S) return;
    const uint32_t l4 = len4[sid];
    uint32_t c;
    if (sub_row[sid] < n_hub) c = 4;
    else if (l4 <= lane_max4) c = kUnitLane;
    else if (l4 <= 16) c = kUnitG4;
    else if (l4 <= 64) c = kUnitG16;
    else c = kUnitG64;
    key[sid] = sub_tab[sid] * 8u + c;
    val[sid] = sid;
}

// Combine records, one thread per (64-row wave w, table t): the slot base
// (the first sub-row of table t whose row is >= the wave's first row; rows
// ascend inside a table) and the mask of the wave's rows with a sub-row in t
// (the run of sub-rows from the base whose rows lie in the wave).
__global__ void k_tab_bases(const uint32_t *__restrict__ sub_row, const uint32_t *__restrict__ sb,
                            uint32_t NT, uint32_t NTpad, uint32_t grp, uint32_t nwaves, uint32_t n_hub,
                            uint32_t *__restrict__ wtab)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nwaves * NT) return;
    const uint32_t w = i / NT, t = i % NT;
    const uint32_t row0 = n_hub + w * 64u;
    const uint32_t end = sb[t + 1];
    uint32_t lo = sb[t], hi = end; // first index in [lo, hi) with sub_row >= row0
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (sub_row[mid] < row0) lo = mid + 1;
        else hi = mid;
    }
    // the wave's rows with a sub-row in t: the run from lo of rows < row0 + 64
    uint64_t m = 0;
    for (uint32_t j = lo; j < end; j++) {
        const uint32_t r = sub_row[j];
        if (r >= row0 + 64u) break;
        m |= 1ull << (r - row0);
    }
    uint32_t *rec = wtab + (size_t)w * 3u * NTpad + 3u * (t / grp) * grp;
    rec[2u * (t % grp)] = (uint32_t)m;
    rec[2u * (t % grp) + 1u] = (uint32_t)(m >> 32);
    rec[2u * grp + t % grp] = lo;
}

// Cost of each work unit in padded index positions (lane units: per wave,
// 64 lanes times the largest per-lane total over its kLaneRows sub-rows; long
// and hub units: their positions), for the static split over workgroups.
__global__ __launch_bounds__(kSweepBlock) void k_tab_unit_cost(const uint4 *__restrict__ units, uint32_t nunits,
                                const uint32_t *__restrict__ sub_start,
                                const uint4 *__restrict__ litems,
                                const uint4 *__restrict__ hub_items, uint32_t lane_max4,
                                uint32_t *__restrict__ cost)
{
    // one workgroup per unit, thread (w, lane) as in the table kernel
    __shared__ unsigned long long s_c;
    const uint32_t ui = blockIdx.x;
    if (ui >= nunits) return;
    if (threadIdx.x == 0) s_c = 0;
    __syncthreads();
    const uint4 un = units[ui];
    const uint32_t kind = un.x & 0xffu;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    if (kind == kUnitLane) {
        uint32_t tot = 0;
        for (uint32_t j = 0; j < kLaneRows; j++) {
            const uint32_t k = (j * kSweepWaves + w) * 64 + lane;
            if (k >= un.z) continue;
            const uint32_t l4 = sub_start[un.y + k + 1] - sub_start[un.y + k];
            if (l4 <= lane_max4) tot += l4;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) tot = max(tot, (uint32_t)__shfl_xor((int)tot, off, 64));
        if (lane == 0) atomicAdd(&s_c, 64ull * tot);
    } else {
        unsigned long long c = 0;
        for (uint32_t k = threadIdx.x; k < un.z; k += blockDim.x)
            c += kind == kUnitHub ? hub_items[un.y + k].y - hub_items[un.y + k].x
                                  : litems[un.y + k].z - litems[un.y + k].y;
        if (c) atomicAdd(&s_c, c);
    }
    __syncthreads();
    if (threadIdx.x == 0) cost[ui] = (uint32_t)min<unsigned long long>(s_c, 0xffffffffull);
}

// Long-sub-row items (sub-row, first position, end position) of one
// (table, class) range of the class-sorted sub-rows.
// Long-sub-row items of every (table, class) segment in one launch: segment
// g = (first position in `sorted`, count, first item) for g < nseg, counts
// above zero and first items ascending, so item q belongs to the last
// segment whose first item is <= q.
__global__ void k_tab_litems(const uint32_t *__restrict__ sorted, const uint4 *__restrict__ seg,
                             uint32_t nseg, uint32_t total, const uint32_t *__restrict__ sub_start,
                             uint4 *__restrict__ out)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= total) return;
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (seg[mid].z <= q) lo = mid;
        else hi = mid;
    }
    const uint4 g = seg[lo];
    const uint32_t sid = sorted[g.x + (q - g.z)];
    out[q] = make_uint4(sid, sub_start[sid], sub_start[sid + 1], 0u);
}

// out[i] = a[idx[i]], out[n + i] = b[idx[i]], out[2n + i] = c[idx[i]]
__global__ void k_gather3(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
                          const uint32_t *__restrict__ c, const uint32_t *__restrict__ idx,
                          uint32_t n, uint32_t *__restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = idx[i];
    out[i] = a[k];
    out[n + i] = b[k];
    out[2 * (size_t)n + i] = c[k];
}

// out[i] = a[idx[i]] for i < n (one block)
__global__ void k_gather1(const uint32_t *__restrict__ a, const uint32_t *__restrict__ idx,
                          uint32_t n, uint32_t *__restrict__ out)
{
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) out[i] = a[idx[i]];
}

} // namespace

static void build_tables(ag_ctx *ctx, ag_graph *g, Plan *p, const Tuning &tn, bool sell_ecol);

static Plan *build_plan(ag_ctx *ctx, ag_graph *g)
{
    hipStream_t s = ctx->stream;
    const uint32_t N = g->N, E = g->E;
    const Tuning tn = g->has_tune ? g->tune : tuning();
    std::unique_ptr<Plan> p(new Plan());
    p->N = N;
    p->E = E;
    p->low_max = tn.low_max;
    p->hub_min = tn.hub_min;
    p->hub_chunk = tn.hub_chunk;
    p->hot_bytes = tn.hot_bytes;
    p->lane_max = tn.lane_max;
    p->max_tables = tn.max_tables;
    if (p->hot_bytes) {
        static bool attr_set = false;
        if (!attr_set) {
            const void *ks[4] = {(const void *)k_pr_sweep<float, false>,
                                 (const void *)k_pr_sweep<bf16_t, false>,
                                 (const void *)k_pr_sweep<__half, false>,
                                 (const void *)k_pr_sweep<float, true>};
            for (const void *k : ks)
                AG_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(kLdsBytes - kStaticLds)));
            const void *kt[4] = {(const void *)k_pr_tables<float, false>,
                                 (const void *)k_pr_tables<bf16_t, false>,
                                 (const void *)k_pr_tables<__half, false>,
                                 (const void *)k_pr_tables<float, true>};
            for (const void *k : kt)
                AG_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(kLdsBytes - kStaticLds)));
            attr_set = true;
        }
    }
    if (!g->inv_outdeg.p || !g->inv_idx.p || !g->inv_dst.p || !g->fwd_outdeg.p)
        fail(AG_ERR_INVALID, "graph has no inverse CSR (built with AG_BUILD_NO_INVERSE?)");

    p->order.alloc((size_t)N * 4);
    p->outdeg.alloc((size_t)N * 4);
    p->roff.alloc((size_t)(N + 1) * 4);
    if (N == 0) {
        AG_HIP(hipMemsetAsync(p->roff.p, 0, 4, s));
        AG_HIP(hipStreamSynchronize(s));
        return p.release();
    }
    const uint32_t maxdeg = device_max1(ctx, g->inv_outdeg.as<uint32_t>(), N);
    const uint32_t kbits = bits_for((uint64_t)maxdeg + 1);
    // temporaries in the context's scratch slots (build_tables, run at the
    // end, leaves Lay6 / Lay7 alone)
    std::lock_guard<std::recursive_mutex> lk(ctx->scratch_mu);
    // partitioned tables when forced, or (auto) when one LDS table cannot
    // hold every source; their build reads the SELL rows' sources in row
    // order, written here beside the SELL slices (scratch slot InSrc)
    const bool tables = tn.layout == AG_LAYOUT_TABLES || (tn.layout == AG_LAYOUT_AUTO && N > kTabEntries);
    // PairA / PairB also hold the table build's per-edge arrays: sized for
    // them now, a first build does not regrow them (a stream wait) there
    const size_t npair = tables ? std::max<size_t>(N, E) : N;
    uint32_t *keys = scratch_as<uint32_t>(ctx, kScrPairA, npair);
    uint32_t *ids = scratch_as<uint32_t>(ctx, kScrPairB, npair);
    uint32_t *skeys = scratch_as<uint32_t>(ctx, kScrLay0 + 7, N);
    uint32_t *rdeg = scratch_as<uint32_t>(ctx, kScrLay0 + 6, N);
    p->rank.alloc((size_t)N * 4);
    DevBuf counts(32);
    const uint32_t nb = ceil_div(N, 256);
    hipLaunchKernelGGL(k_degree_keys, dim3(nb), dim3(256), 0, s, g->inv_outdeg.as<uint32_t>(), N,
                       maxdeg, keys, ids);
    AG_HIP(hipGetLastError());
    radix_sort_pairs(ctx, keys, ids, skeys, p->order.as<uint32_t>(), N, kbits);
    AG_HIP(hipMemsetAsync(counts.p, 0, 32, s));
    hipLaunchKernelGGL(k_relabel_vertex, dim3(std::min<uint32_t>(nb, (uint32_t)ctx->num_cus * 8)),
                       dim3(256), 0, s, p->order.as<uint32_t>(), skeys, maxdeg,
                       g->fwd_outdeg.as<uint32_t>(), N, p->rank.as<uint32_t>(), rdeg,
                       p->outdeg.as<uint32_t>(), tn.low_max, tn.hub_min, counts.as<uint32_t>());
    AG_HIP(hipGetLastError());
    scan_exclusive_u32(ctx, rdeg, p->roff.as<uint32_t>(), N);
    uint32_t cnt[8];
    AG_HIP(hipMemcpyAsync(cnt, counts.p, 32, hipMemcpyDeviceToHost, s));
    AG_HIP(hipStreamSynchronize(s));
    p->n_hub = cnt[0];
    p->n_nonlow = cnt[1];
    p->n_nonzero = cnt[2];
    p->n_dangling = cnt[3];
    // lane-group bins inside the CSR rows: [n_hub, n_w) 64 lanes per row,
    // [n_w, n_h) 32, [n_h, n_nonlow) 16 (labels are sorted by degree)
    p->n_w = std::min(std::max(cnt[4], p->n_hub), p->n_nonlow);
    p->n_h = std::min(std::max(cnt[5], p->n_w), p->n_nonlow);
    uint32_t csr_edges = 0;
    AG_HIP(hipMemcpyAsync(&csr_edges, p->roff.as<uint32_t>() + p->n_nonlow, 4,
                          hipMemcpyDeviceToHost, s));
    AG_HIP(hipStreamSynchronize(s));
    p->csr_edges = csr_edges;
    const uint32_t nlow = p->n_nonzero - p->n_nonlow;
    p->low_slices = ceil_div(nlow, 64);
    p->slice_ptr.alloc((size_t)(p->low_slices + 1) * 4);
    uint32_t sell_slots = 0;
    if (p->low_slices) {
        hipLaunchKernelGGL(k_slice_width, dim3(ceil_div(p->low_slices, 256)), dim3(256), 0, s,
                           rdeg, p->n_nonlow, p->low_slices,
                           p->slice_ptr.as<uint32_t>());
        scan_exclusive_u32(ctx, p->slice_ptr.as<uint32_t>(), p->slice_ptr.as<uint32_t>(),
                           p->low_slices);
        AG_HIP(hipMemcpyAsync(&sell_slots, p->slice_ptr.as<uint32_t>() + p->low_slices, 4,
                              hipMemcpyDeviceToHost, s));
        AG_HIP(hipStreamSynchronize(s));
        if ((uint64_t)csr_edges + sell_slots > 0xffffffffull)
            fail(AG_ERR_INVALID, "gather layout exceeds 2^32 slots");
        hipLaunchKernelGGL(k_add_base, dim3(ceil_div(p->low_slices + 1, 256)), dim3(256), 0, s,
                           p->slice_ptr.as<uint32_t>(), p->low_slices + 1, csr_edges);
    } else {
        AG_HIP(hipMemsetAsync(p->slice_ptr.p, 0, 4, s));
    }
    p->sell_slots = sell_slots;
    const uint64_t ncols = (uint64_t)csr_edges + sell_slots;
    p->cols.alloc((size_t)ncols * 4);
    AG_HIP(hipMemsetAsync(p->cols.p, 0, (size_t)ncols * 4, s));
    if (p->n_nonlow && csr_edges) {
        uint32_t *brow = scratch_as<uint32_t>(ctx, kScrBlk, ceil_div(csr_edges, kFillBlock));
        hipLaunchKernelGGL(k_block_rows, dim3(ceil_div(p->n_nonlow, 256)), dim3(256), 0, s,
                           p->roff.as<uint32_t>(), p->n_nonlow, 0u, csr_edges, brow);
        hipLaunchKernelGGL(k_fill_csr_edges, dim3(ceil_div(csr_edges, kFillBlock)), dim3(kFillBlock),
                           0, s, p->order.as<uint32_t>(), p->roff.as<uint32_t>(), brow,
                           g->inv_idx.as<uint32_t>(), g->inv_dst.as<uint32_t>(),
                           p->rank.as<uint32_t>(), p->n_nonlow, 0u, csr_edges,
                           p->cols.as<uint32_t>());
    }
    uint32_t *ecol_sell = tables && nlow ? scratch_as<uint32_t>(ctx, kScrInSrc, E) : nullptr;
    if (nlow)
        hipLaunchKernelGGL(k_fill_sell_rows, dim3(ceil_div(nlow, 256)), dim3(256), 0, s,
                           p->order.as<uint32_t>(), rdeg,
                           g->inv_idx.as<uint32_t>(), g->inv_dst.as<uint32_t>(),
                           p->rank.as<uint32_t>(), p->slice_ptr.as<uint32_t>(), p->n_nonlow, nlow,
                           p->cols.as<uint32_t>(), p->roff.as<uint32_t>(), ecol_sell);
    AG_HIP(hipGetLastError());

    // Hub chunk table (host: hub rows are few).
    std::vector<uint32_t> hdeg(p->n_hub);
    if (p->n_hub)
        AG_HIP(hipMemcpyAsync(hdeg.data(), rdeg, (size_t)p->n_hub * 4, hipMemcpyDeviceToHost, s));
    AG_HIP(hipStreamSynchronize(s));
    if (tn.hub_chunk == AG_TUNE_AUTO) {
        // Hub chunks are the sweep's first work units; the persistent
        // workgroups (one per CU) balance when each gets several.  R-MAT s22
        // (too few hub edges for 8 chunks of 4,096 per CU) runs 239 us per
        // sweep with 2,048-edge chunks against 264 us; s24 runs 1.15 ms with
        // 4,096 against 1.18 ms.
        uint64_t n4096 = 0;
        for (uint32_t i = 0; i < p->n_hub; i++) n4096 += ceil_div(hdeg[i], 4096u);
        p->hub_chunk = n4096 >= 8ull * (uint64_t)ctx->num_cus ? 4096u : 2048u;
    }
    std::vector<uint32_t> first(p->n_hub + 1), crow;
    uint32_t tot = 0;
    for (uint32_t i = 0; i < p->n_hub; i++) {
        first[i] = tot;
        const uint32_t nch = ceil_div(hdeg[i], p->hub_chunk);
        for (uint32_t q = 0; q < nch; q++) crow.push_back(i);
        tot += nch;
    }
    first[p->n_hub] = tot;
    p->hub_chunks = tot;
    p->chunk_first.alloc((size_t)(p->n_hub + 1) * 4);
    p->chunk_row.alloc((size_t)std::max<uint32_t>(tot, 1) * 4);
    p->hub_partial.alloc((size_t)std::max<uint32_t>(tot, 1) * 4);
    AG_HIP(hipMemcpyAsync(p->chunk_first.p, first.data(), first.size() * 4, hipMemcpyHostToDevice, s));
    if (tot) AG_HIP(hipMemcpyAsync(p->chunk_row.p, crow.data(), crow.size() * 4, hipMemcpyHostToDevice, s));

    p->nb_hub = ceil_div(tot, kSweepWaves);
    p->nb_w = ceil_div(p->n_w - p->n_hub, kSweepWaves);
    p->nb_h = ceil_div(p->n_h - p->n_w, 2 * kSweepWaves);
    p->nb_q = ceil_div(p->n_nonlow - p->n_h, 4 * kSweepWaves);
    p->nb_low = ceil_div(p->low_slices, kSweepWaves);
    p->nb_zero = ceil_div(N - p->n_nonzero, kZeroRowsPerUnit);
    p->nvblocks = p->nb_hub + p->nb_w + p->nb_h + p->nb_q + p->nb_low + p->nb_zero;
    AG_HIP(hipStreamSynchronize(s));
    if (tables) build_tables(ctx, g, p.get(), tn, ecol_sell != nullptr);
    return p.release();
}

// Builds the partitioned-table layout on top of the row plan (labels, row
// offsets).  Per-edge work runs on the device; per-table and per-unit
// bookkeeping on the host from small downloads.
static constexpr uint32_t kMinTabEdges = 65536; // an LDS table needs this many edges
// Fixed cost of a unit, in index positions: each wave of a unit pays about two
// dependent memory round trips whatever its length.
static constexpr uint32_t kUnitOverhead = 4096;

static void build_tables(ag_ctx *ctx, ag_graph *g, Plan *p, const Tuning &tn, bool sell_ecol)
{
    hipStream_t s = ctx->stream;
    const uint32_t N = p->N, E = p->E, nrows = p->n_nonzero, n_hub = p->n_hub;
    if (E == 0 || nrows == 0) return; // no gather: the row layout serves
    std::unique_ptr<TabPlan> tp(new TabPlan());
    const uint32_t T = kTabEntries; // LDS holds T + 1 floats
    static_assert((kTabEntries + 1) * 4 <= kLdsBytes - kStaticLds, "LDS table too large");
    tp->T = T;
    auto dl = [&](void *dst, const void *src, size_t bytes) {
        if (bytes) AG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    };
    // Temporaries live in the context's scratch slots (kept across builds:
    // freeing a large buffer waits for the device).  Slot use by step:
    //   InSrc: ecol (1-3);  PairA/PairB: key/erow (3), -/flag (4),
    //   key2/val2 (8);  InDst: fscan (4-7), sk2 (8);  Lay0/Lay1: skey/scol
    //   (3-7);  Lay8: srow (3-4);  Lay2/Lay3: sub_row/sub_tab (4-12);  Lay4:
    //   sub_e0 (4-7), sorted_sub (8-10);  Lay5: len4 (5-9);  Lay6 is
    //   build_plan's.
    std::lock_guard<std::recursive_mutex> lk(ctx->scratch_mu);
    // 1. relabelled sources of every row with in-edges, label order: the row
    // layout's CSR part holds those of rows [0, n_nonlow) already, and with
    // sell_ecol the SELL fill has written the rest here
    uint32_t *ecol = scratch_as<uint32_t>(ctx, kScrInSrc, E);
    const uint32_t e_csr = (uint32_t)std::min<uint64_t>(p->csr_edges, E);
    if (e_csr)
        AG_HIP(hipMemcpyAsync(ecol, p->cols.p, (size_t)e_csr * 4, hipMemcpyDeviceToDevice, s));
    if (e_csr < E && !sell_ecol) {
        uint32_t *brow = scratch_as<uint32_t>(ctx, kScrBlk, ceil_div(E - e_csr, kFillBlock));
        hipLaunchKernelGGL(k_block_rows, dim3(ceil_div(nrows, 256)), dim3(256), 0, s,
                           p->roff.as<uint32_t>(), nrows, e_csr, E, brow);
        hipLaunchKernelGGL(k_fill_csr_edges, dim3(ceil_div(E - e_csr, kFillBlock)), dim3(kFillBlock),
                           0, s, p->order.as<uint32_t>(), p->roff.as<uint32_t>(), brow,
                           g->inv_idx.as<uint32_t>(), g->inv_dst.as<uint32_t>(),
                           p->rank.as<uint32_t>(), nrows, e_csr, E, ecol);
    }
    AG_HIP(hipGetLastError());
    // 2. edges per table; LDS tables while each holds enough edges
    {
        DevBuf hist(kTabHist * 4);
        AG_HIP(hipMemsetAsync(hist.p, 0, kTabHist * 4, s));
        hipLaunchKernelGGL(k_tab_hist, dim3(kTabHist - 1), dim3(1024), 0, s, p->outdeg.as<uint32_t>(), N,
                           hist.as<uint32_t>());
        AG_HIP(hipGetLastError());
        uint32_t h[kTabHist];
        dl(h, hist.p, sizeof(h));
        AG_HIP(hipStreamSynchronize(s));
        const uint32_t min_edges = std::max<uint32_t>(1, std::min<uint32_t>(kMinTabEdges, E / 64));
        uint32_t nt = 0;
        while (nt < tn.max_tables && nt < (uint32_t)kTabHist - 1 && (uint64_t)nt * T < N &&
               h[nt] >= min_edges)
            nt++;
        tp->nt = nt;
        tp->NT = nt + 1;
    }
    const uint32_t nt = tp->nt, NT = tp->NT;
    // 3. stable sort of the edges by table, each edge's row and source label
    // moving with it: each table's edges stay in (row, edge) order
    uint32_t *skey = scratch_as<uint32_t>(ctx, kScrLay0, E);
    uint32_t *scol = scratch_as<uint32_t>(ctx, kScrLay0 + 1, E);
    uint32_t *srow = scratch_as<uint32_t>(ctx, kScrLay8, E);
    {
        uint32_t *key = scratch_as<uint32_t>(ctx, kScrPairA, E);
        uint32_t *erow = scratch_as<uint32_t>(ctx, kScrPairB, E);
        hipLaunchKernelGGL(k_tab_keys, dim3(ceil_div(E, 256)), dim3(256), 0, s, ecol, E, nt, key);
        // row of every label-order edge slot
        uint32_t *brow = scratch_as<uint32_t>(ctx, kScrBlk, ceil_div(E, kFillBlock));
        hipLaunchKernelGGL(k_block_rows, dim3(ceil_div(nrows, 256)), dim3(256), 0, s,
                           p->roff.as<uint32_t>(), nrows, 0u, E, brow);
        hipLaunchKernelGGL(k_edge_rows, dim3(ceil_div(E, kFillBlock)), dim3(kFillBlock), 0, s,
                           p->roff.as<uint32_t>(), brow, nrows, E, erow);
        AG_HIP(hipGetLastError());
        radix_sort_triples(ctx, key, erow, ecol, skey, srow, scol, E, bits_for(NT));
    }
    // 4. sub-rows: starts where (table, row) changes
    uint32_t *fscan = scratch_as<uint32_t>(ctx, kScrInDst, (size_t)E + 1);
    uint32_t S = 0;
    uint32_t *sub_row, *sub_tab, *sub_e0;
    {
        uint32_t *flag = scratch_as<uint32_t>(ctx, kScrPairB, E);
        hipLaunchKernelGGL(k_tab_flags, dim3(ceil_div(E, 256)), dim3(256), 0, s, skey, srow, E, flag);
        AG_HIP(hipGetLastError());
        scan_exclusive_u32(ctx, flag, fscan, E);
        dl(&S, fscan + E, 4);
        AG_HIP(hipStreamSynchronize(s));
        sub_row = scratch_as<uint32_t>(ctx, kScrLay0 + 2, S);
        sub_tab = scratch_as<uint32_t>(ctx, kScrLay0 + 3, S);
        sub_e0 = scratch_as<uint32_t>(ctx, kScrLay0 + 4, (size_t)S + 1);
        hipLaunchKernelGGL(k_tab_subrows, dim3(ceil_div(E, 256)), dim3(256), 0, s, flag, fscan, skey,
                           srow, E, sub_row, sub_tab, sub_e0);
        AG_HIP(hipGetLastError());
    }
    tp->S = S;
    // 5. padded lengths and starts (units of 4 entries, all tables in one range)
    uint32_t *len4 = scratch_as<uint32_t>(ctx, kScrLay0 + 5, S);
    tp->sub_start.alloc(((size_t)S + 1) * 4);
    hipLaunchKernelGGL(k_tab_len4, dim3(ceil_div(S, 256)), dim3(256), 0, s, sub_e0, S, len4);
    AG_HIP(hipGetLastError());
    scan_exclusive_u32(ctx, len4, tp->sub_start.as<uint32_t>(), S);
    // 6. first sub-row of each table
    DevBuf sbd(((size_t)NT + 1) * 4);
    AG_HIP(hipMemsetAsync(sbd.p, 0xff, ((size_t)NT + 1) * 4, s));
    hipLaunchKernelGGL(k_tab_bounds, dim3(ceil_div(S, 256)), dim3(256), 0, s, sub_tab, S,
                       sbd.as<uint32_t>());
    AG_HIP(hipGetLastError());
    tp->SB.assign(NT + 1, 0);
    dl(tp->SB.data(), sbd.p, (size_t)NT * 4);
    uint32_t total4 = 0;
    dl(&total4, tp->sub_start.as<uint32_t>() + S, 4);
    AG_HIP(hipStreamSynchronize(s));
    tp->SB[NT] = S;
    for (int t = (int)NT - 1; t >= 0; t--)
        if (tp->SB[t] == 0xffffffffu) tp->SB[t] = tp->SB[t + 1]; // empty table
    AG_HIP(hipMemcpyAsync(sbd.p, tp->SB.data(), ((size_t)NT + 1) * 4, hipMemcpyHostToDevice, s));
    uint32_t cold_base4 = total4;
    if (tp->SB[nt] < S) dl(&cold_base4, tp->sub_start.as<uint32_t>() + tp->SB[nt], 4);
    AG_HIP(hipStreamSynchronize(s));
    // the table kernel addresses its arrays by 32-bit byte offsets (at32):
    // larger layouts are left to the row layout
    if ((uint64_t)total4 * 16 >= (1ull << 32) || (uint64_t)S * 16 >= (1ull << 32) ||
        ((uint64_t)N + 1) * 4 >= (1ull << 32))
        return;
    tp->cold_base4 = cold_base4;
    tp->idx16_len = (uint64_t)cold_base4 * 4;
    tp->idx32_len = (uint64_t)(total4 - cold_base4) * 4;
    // 7. source indices, padded with the sentinels
    tp->idx16.alloc(std::max<uint64_t>(tp->idx16_len, 4) * 2);
    tp->idx32.alloc(std::max<uint64_t>(tp->idx32_len, 4) * 4);
    const uint32_t fill_grid = ctx->num_cus * 8;
    hipLaunchKernelGGL(k_fill_u16, dim3(fill_grid), dim3(256), 0, s, tp->idx16.as<uint16_t>(),
                       (uint64_t)tp->idx16.bytes / 2, (uint16_t)T);
    hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid), dim3(256), 0, s, tp->idx32.as<uint32_t>(),
                       (uint64_t)tp->idx32.bytes / 4, N);
    hipLaunchKernelGGL(k_tab_fill_idx, dim3(ceil_div(E, 256)), dim3(256), 0, s, skey, scol,
                       fscan, sub_e0, tp->sub_start.as<uint32_t>(), E, T, nt, cold_base4,
                       tp->idx16.as<uint16_t>(), tp->idx32.as<uint32_t>());
    AG_HIP(hipGetLastError());
    // 8. sub-rows by (table, class): the long ones' lists and the hub sub-rows
    const uint32_t lane_max4 = tn.lane_max / 4;
    const uint32_t nkeys = NT * 8;
    std::vector<uint32_t> ccnt(nkeys), cfirst(nkeys);
    uint32_t *sorted_sub = scratch_as<uint32_t>(ctx, kScrLay0 + 4, S); // sub_e0 is dead
    {
        uint32_t *key2 = scratch_as<uint32_t>(ctx, kScrPairA, S);
        uint32_t *val2 = scratch_as<uint32_t>(ctx, kScrPairB, S);
        uint32_t *sk2 = scratch_as<uint32_t>(ctx, kScrInDst, S);
        DevBuf cnt(nkeys * 4), fst(nkeys * 4);
        hipLaunchKernelGGL(k_tab_classes, dim3(ceil_div(S, 256)), dim3(256), 0, s, sub_row, sub_tab,
                           len4, S, n_hub, lane_max4, key2, val2);
        AG_HIP(hipGetLastError());
        radix_sort_pairs(ctx, key2, val2, sk2, sorted_sub, S, bits_for(nkeys));
        vertex_map(ctx, sk2, S, nkeys, cnt.as<uint32_t>(), fst.as<uint32_t>(), nullptr);
        dl(ccnt.data(), cnt.p, (size_t)nkeys * 4);
        dl(cfirst.data(), fst.p, (size_t)nkeys * 4);
        AG_HIP(hipStreamSynchronize(s));
    }
    // 9. hub chunks: slots in (row, table, chunk) order, processed by table
    tp->hub_chunk = std::max<uint32_t>(4, p->hub_chunk / 4 * 4);
    const uint32_t chunk4 = tp->hub_chunk / 4;
    std::vector<uint32_t> H(NT);
    std::vector<std::vector<uint32_t>> hrow(NT), hlen(NT), hstart(NT);
    {
        // hub rows come first among a table's sub-rows (rows ascend): gather
        // (row, length, start) of sub-rows SB[t] .. SB[t] + H[t] - 1 of every
        // table on the device, one download
        std::vector<uint32_t> hidx;
        for (uint32_t t = 0; t < NT; t++) {
            H[t] = ccnt[t * 8 + 4];
            for (uint32_t k = 0; k < H[t]; k++) hidx.push_back(tp->SB[t] + k);
        }
        const uint32_t nh = (uint32_t)hidx.size();
        std::vector<uint32_t> hv((size_t)nh * 3);
        if (nh) {
            DevBuf didx((size_t)nh * 4), dout((size_t)nh * 12);
            AG_HIP(hipMemcpyAsync(didx.p, hidx.data(), (size_t)nh * 4, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_gather3, dim3(ceil_div(nh, 256)), dim3(256), 0, s, sub_row, len4,
                               tp->sub_start.as<uint32_t>(), didx.as<uint32_t>(), nh,
                               dout.as<uint32_t>());
            AG_HIP(hipGetLastError());
            dl(hv.data(), dout.p, (size_t)nh * 12);
            AG_HIP(hipStreamSynchronize(s));
        }
        for (uint32_t t = 0, q = 0; t < NT; t++) {
            hrow[t].resize(H[t]);
            hlen[t].resize(H[t]);
            hstart[t].resize(H[t]);
            for (uint32_t k = 0; k < H[t]; k++, q++) {
                hrow[t][k] = hv[q];
                hlen[t][k] = hv[nh + q];
                hstart[t][k] = hv[2 * (size_t)nh + q];
            }
        }
    }
    std::vector<uint32_t> first(n_hub + 1, 0);
    std::vector<std::vector<uint32_t>> slot_of(NT); // first slot of each hub sub-row
    {
        std::vector<std::vector<std::pair<uint32_t, uint32_t>>> per_row(n_hub); // (table, k)
        for (uint32_t t = 0; t < NT; t++) {
            slot_of[t].resize(H[t]);
            for (uint32_t k = 0; k < H[t]; k++) {
                if (hrow[t][k] >= n_hub) fail(AG_ERR_INVALID, "table layout: hub sub-row order");
                per_row[hrow[t][k]].push_back({t, k});
            }
        }
        uint32_t slot = 0;
        for (uint32_t h = 0; h < n_hub; h++) {
            first[h] = slot;
            for (auto &tk : per_row[h]) { // tables ascend
                slot_of[tk.first][tk.second] = slot;
                slot += ceil_div(hlen[tk.first][tk.second], chunk4);
            }
        }
        first[n_hub] = slot;
        tp->hub_chunks = slot;
    }
    std::vector<uint4> hub_items;
    hub_items.reserve(tp->hub_chunks);
    // 10. work units, table by table; long-sub-row items in unit order
    std::vector<uint4> units;
    uint32_t nlong = 0;
    for (uint32_t t = 0; t < NT; t++)
        for (uint32_t c = 1; c <= 3; c++) nlong += ccnt[t * 8 + c];
    tp->items.alloc((size_t)std::max<uint32_t>(nlong, 1) * sizeof(uint4));
    std::vector<uint4> lsegs; // (first in sorted_sub, count, first item) per (table, class)
    uint32_t lpos = 0;
    for (uint32_t t = 0; t < NT; t++) {
        const uint32_t tt = t << 8;
        const uint32_t h0 = (uint32_t)hub_items.size();
        for (uint32_t k = 0; k < H[t]; k++) {
            const uint32_t nch = ceil_div(hlen[t][k], chunk4);
            const uint32_t end = hstart[t][k] + hlen[t][k];
            for (uint32_t j = 0; j < nch; j++) {
                const uint32_t b = hstart[t][k] + j * chunk4;
                hub_items.push_back(make_uint4(b, std::min(end, b + chunk4), slot_of[t][k] + j,
                                               tp->SB[t] + k));
            }
        }
        const uint32_t h1 = (uint32_t)hub_items.size();
        for (uint32_t c = h0; c < h1; c += kSweepWaves)
            units.push_back(make_uint4(tt | kUnitHub, c, std::min<uint32_t>(kSweepWaves, h1 - c), 0u));
        // sub-rows per unit: kLongItems per lane group (16 / 4 / 1 groups per wave)
        const uint32_t per_unit[4] = {0, 16 * kSweepWaves * kLongItems, 4 * kSweepWaves * kLongItems,
                                      kSweepWaves * kLongItems};
        for (uint32_t c = 3; c >= 1; c--) {
            const uint32_t n = ccnt[t * 8 + c], f = cfirst[t * 8 + c];
            if (n) lsegs.push_back(make_uint4(f, n, lpos, 0u));
            for (uint32_t k = 0; k < n; k += per_unit[c])
                units.push_back(make_uint4(tt | c, lpos + k, std::min<uint32_t>(per_unit[c], n - k),
                                           lpos + n));
            lpos += n;
        }
        // cold lane units are smaller: the kernel takes them one at a time
        const uint32_t lrows = t < nt ? kLaneUnitRows : kSweepBlock;
        for (uint32_t k = tp->SB[t] + H[t]; k < tp->SB[t + 1]; k += lrows)
            units.push_back(make_uint4(tt | kUnitLane, k,
                                       std::min<uint32_t>(lrows, tp->SB[t + 1] - k), tp->SB[t + 1]));
    }
    if (lpos != nlong) fail(AG_ERR_INVALID, "table layout: long-item count");
    DevBuf dsegs(std::max<size_t>(lsegs.size(), 1) * sizeof(uint4));
    if (nlong) {
        AG_HIP(hipMemcpyAsync(dsegs.p, lsegs.data(), lsegs.size() * sizeof(uint4),
                              hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_tab_litems, dim3(ceil_div(nlong, 256)), dim3(256), 0, s, sorted_sub,
                           dsegs.as<uint4>(), (uint32_t)lsegs.size(), nlong,
                           tp->sub_start.as<uint32_t>(), tp->items.as<uint4>());
        AG_HIP(hipGetLastError());
    }
    tp->nunits = (uint32_t)units.size();
    // the cold table's lane units close the list: consecutive kSweepBlock
    // ranges of sub-rows (the kernel computes them from the unit index)
    tp->cold_lane_u0 = tp->nunits;
    while (tp->cold_lane_u0 > 0 && units[tp->cold_lane_u0 - 1].x == ((nt << 8) | kUnitLane))
        tp->cold_lane_u0--;
    tp->cold_lane_s0 = tp->cold_lane_u0 < tp->nunits ? units[tp->cold_lane_u0].y : 0u;
    tp->cold_lane_s1 = tp->SB[NT];
    for (uint32_t u = tp->cold_lane_u0; u < tp->nunits; u++)
        if (units[u].y != tp->cold_lane_s0 + (u - tp->cold_lane_u0) * kSweepBlock ||
            units[u].z != std::min<uint32_t>(kSweepBlock, tp->cold_lane_s1 - units[u].y))
            fail(AG_ERR_INVALID, "table layout: cold lane units are not consecutive ranges");
    tp->hub_items.alloc(std::max<size_t>(hub_items.size(), 1) * sizeof(uint4));
    if (!hub_items.empty())
        AG_HIP(hipMemcpyAsync(tp->hub_items.p, hub_items.data(), hub_items.size() * sizeof(uint4),
                              hipMemcpyHostToDevice, s));
    tp->units.alloc(std::max<size_t>(units.size(), 1) * sizeof(uint4));
    if (!units.empty())
        AG_HIP(hipMemcpyAsync(tp->units.p, units.data(), units.size() * sizeof(uint4),
                              hipMemcpyHostToDevice, s));
    tp->chunk_first.alloc(((size_t)n_hub + 1) * 4);
    AG_HIP(hipMemcpyAsync(tp->chunk_first.p, first.data(), first.size() * 4, hipMemcpyHostToDevice, s));
    tp->hub_partial.alloc(std::max<size_t>(tp->hub_chunks, 1) * 4);
    // one slot of slack: the combine kernel's lanes without a sub-row in a
    // table read (and discard) the slot after the wave's last one
    tp->part.alloc(((size_t)S + 1) * 4);
    // 11. split the units over one workgroup per CU by their cost
    std::vector<uint32_t> cost(units.size());
    if (!units.empty()) {
        DevBuf dcost(units.size() * 4);
        hipLaunchKernelGGL(k_tab_unit_cost, dim3((uint32_t)units.size()), dim3(kSweepBlock), 0, s,
                           tp->units.as<uint4>(), (uint32_t)units.size(),
                           tp->sub_start.as<uint32_t>(), tp->items.as<uint4>(),
                           tp->hub_items.as<uint4>(), lane_max4, dcost.as<uint32_t>());
        AG_HIP(hipGetLastError());
        dl(cost.data(), dcost.p, units.size() * 4);
        AG_HIP(hipStreamSynchronize(s));
    }
    // the cold table's units come last; the kernel deals them round-robin,
    // and the LDS units [0, cold_unit0) are split into contiguous ranges
    uint32_t cold_unit0 = (uint32_t)units.size();
    while (cold_unit0 > 0 && (units[cold_unit0 - 1].x >> 8) == nt) cold_unit0--;
    tp->cold_unit0 = cold_unit0;
    const uint32_t grid = std::max<uint32_t>(1, (uint32_t)ctx->num_cus);
    std::vector<uint64_t> pre(cold_unit0 + 1, 0);
    for (size_t u = 0; u < cold_unit0; u++) pre[u + 1] = pre[u] + cost[u] + kUnitOverhead;
    std::vector<uint32_t> wgb(grid + 1);
    for (uint32_t b = 0; b <= grid; b++) {
        const uint64_t target = pre.back() * b / grid;
        wgb[b] = (uint32_t)(std::lower_bound(pre.begin(), pre.end(), target) - pre.begin());
        wgb[b] = std::min<uint32_t>(wgb[b], cold_unit0);
    }
    wgb[0] = 0;
    wgb[grid] = cold_unit0;
    tp->grid = grid;
    tp->wg_begin.alloc(((size_t)grid + 1) * 4);
    AG_HIP(hipMemcpyAsync(tp->wg_begin.p, wgb.data(), wgb.size() * 4, hipMemcpyHostToDevice, s));
    // 12. row masks and per-wave slot bases of the combine kernel
    const uint32_t nnh = nrows - std::min(n_hub, nrows);
    tp->nwaves = ceil_div(nnh, 64);
    tp->row_blocks = ceil_div(tp->nwaves, kCombBlock / 64);
    tp->hub_blocks = ceil_div(std::min(n_hub, nrows), kCombBlock / 64);
    // tables per combine record (a multiple of the group the kernel loads at
    // once): 4, 8, 12, 17, 2 x 12, or 2, 3 or 4 x 17
    tp->NTpad = NT <= 4 ? 4u : NT <= 8 ? 8u : NT <= 12 ? 12u : NT <= 17 ? 17u : NT <= 24 ? 24u
              : NT <= 34 ? 34u : NT <= 51 ? 51u : 68u;
    tp->grp = tp->NTpad == 24 ? 12u : tp->NTpad > 17 ? 17u : tp->NTpad;
    static_assert(kMaxTabNT <= 68, "combine records hold at most 68 tables");
    const size_t wrec = ((size_t)tp->nwaves * 3 * tp->NTpad + 1) * 4;
    tp->wtab.alloc(wrec);
    AG_HIP(hipMemsetAsync(tp->wtab.p, 0, wrec, s));
    if (tp->nwaves)
        hipLaunchKernelGGL(k_tab_bases, dim3(ceil_div((uint64_t)tp->nwaves * NT, 256)), dim3(256), 0,
                           s, sub_row, sbd.as<uint32_t>(), NT, tp->NTpad, tp->grp, tp->nwaves, n_hub,
                           tp->wtab.as<uint32_t>());
    AG_HIP(hipGetLastError());
    tp->tab_edges.assign(NT, 0);
    {
        // index positions per table: sub_start at each table's first sub-row
        // (sbd holds SB[0 .. NT] on the device)
        std::vector<uint32_t> st(NT + 1);
        DevBuf dst_((size_t)(NT + 1) * 4);
        hipLaunchKernelGGL(k_gather1, dim3(1), dim3(256), 0, s, tp->sub_start.as<uint32_t>(),
                           sbd.as<uint32_t>(), NT + 1, dst_.as<uint32_t>());
        AG_HIP(hipGetLastError());
        dl(st.data(), dst_.p, (size_t)(NT + 1) * 4);
        AG_HIP(hipStreamSynchronize(s));
        for (uint32_t t = 0; t < NT; t++) tp->tab_edges[t] = (uint64_t)(st[t + 1] - st[t]) * 4;
    }
    AG_HIP(hipStreamSynchronize(s));
    p->tab = std::move(tp);
}

Plan *ensure_plan(ag_ctx *ctx, ag_graph *g, double *ms)
{
    if (ms) *ms = 0.0;
    if (g->plan) return g->plan;
    PhaseRange phase_("pagerank.layout_build");
    auto t0 = std::chrono::steady_clock::now();
    g->plan = build_plan(ctx, g);
    auto t1 = std::chrono::steady_clock::now();
    if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return g->plan;
}

// Persistent sweep workgroups: as many per CU as the LDS hot table allows (<= 2).
static uint32_t sweep_grid(const ag_ctx *ctx, const Plan *p)
{
    const uint32_t per_cu =
        std::min<uint32_t>(2, std::max<uint32_t>(1, kLdsBytes / (p->hot_bytes + kStaticLds)));
    return std::max<uint32_t>(1, std::min<uint32_t>(p->nvblocks, (uint32_t)ctx->num_cus * per_cu));
}

// Combine-kernel blocks of one sweep: the row blocks, plus the zero-in-degree
// rows' blocks unless those rows are skipped (at least one block).
static uint32_t tab_comb_blocks(const Plan *p, bool skip_zero)
{
    const uint32_t z = skip_zero ? 0u : ceil_div(p->N - p->n_nonzero, kCombBlock);
    return std::max<uint32_t>(1, p->tab->row_blocks + p->tab->hub_blocks + z);
}

// Residual-partial slots per sweep: the row-layout sweep grid, or the
// combine kernel's blocks with the partitioned tables.
static uint32_t part_slots(const ag_ctx *ctx, const Plan *p)
{
    const uint32_t n = sweep_grid(ctx, p);
    return p->tab ? std::max(n, tab_comb_blocks(p, false)) : n;
}

// Per-call device buffers of the sweep loop, as raw pointers: either the
// plan's cached Work buffers or a carve-up of the allocation-free scratch.
struct WorkPtrs {
    uint32_t grid = 0;
    float *pr = nullptr;
    void *c0 = nullptr, *c1 = nullptr;
    double *blk_err = nullptr, *blk_dang = nullptr;
    uint32_t *blk_act = nullptr;
    SweepState *st = nullptr;
    double *err_iter = nullptr;
    uint32_t *act_iter = nullptr, *fmt = nullptr;
    double *sum_part = nullptr; // ceil(N / kBlock) + 1 doubles
};

static void ensure_work(ag_ctx *ctx, Plan *p, uint32_t max_iter)
{
    Work &w = p->work;
    const uint32_t N = p->N;
    if (!w.pr.p) {
        const uint32_t grid = sweep_grid(ctx, p);
        w.grid = grid;
        // one more slot than vertices: the zero the cold table's padding reads
        w.pr.alloc((size_t)N * 4);
        w.c0.alloc(((size_t)N + 1) * 4);
        w.c1.alloc(((size_t)N + 1) * 4);
        const uint32_t parts = part_slots(ctx, p);
        w.blk_err.alloc((size_t)parts * 8);
        w.blk_dang.alloc((size_t)parts * 8);
        w.blk_act.alloc((size_t)parts * 4);
        w.state.alloc(sizeof(SweepState));
        w.sum_part.alloc((size_t)(ceil_div(N, kBlock) + 1) * 8 + 8);
        w.out.alloc((size_t)N * 4);
    }
    if (w.max_iter < max_iter || !w.err_iter.p) {
        w.err_iter.alloc((size_t)(max_iter + 1) * 8);
        w.act_iter.alloc((size_t)(max_iter + 1) * 4);
        w.fmt.alloc((size_t)(max_iter + 3) * 4);
        w.max_iter = max_iter;
    }
}

static WorkPtrs work_ptrs(const Work &w)
{
    WorkPtrs r;
    r.grid = w.grid;
    r.pr = w.pr.as<float>();
    r.c0 = w.c0.p;
    r.c1 = w.c1.p;
    r.blk_err = w.blk_err.as<double>();
    r.blk_dang = w.blk_dang.as<double>();
    r.blk_act = w.blk_act.as<uint32_t>();
    r.st = w.state.as<SweepState>();
    r.err_iter = w.err_iter.as<double>();
    r.act_iter = w.act_iter.as<uint32_t>();
    r.fmt = w.fmt.as<uint32_t>();
    r.sum_part = w.sum_part.as<double>();
    return r;
}

static SweepArgs make_args(const Plan *p, const WorkPtrs &w, const ag_pr_options &o,
                           uint32_t max_iter)
{
    SweepArgs a;
    std::memset(&a, 0, sizeof(a));
    a.N = p->N;
    a.n_hub = p->n_hub;
    a.n_nonlow = p->n_nonlow;
    a.n_nonzero = p->n_nonzero;
    a.nb_hub = p->nb_hub;
    a.nb_w = p->nb_w;
    a.nb_h = p->nb_h;
    a.nb_q = p->nb_q;
    a.n_w = p->n_w;
    a.n_h = p->n_h;
    a.hub_chunks = p->hub_chunks;
    a.nb_low = p->nb_low;
    a.nb_zero = p->nb_zero;
    a.nvblocks = p->nvblocks;
    a.hub_chunk = p->hub_chunk;
    a.mode = o.mode;
    a.max_iter = max_iter;
    a.finish_fp32 = (o.flags & AG_PR_FLAG_FINISH_FP32) ? 1u : 0u;
    a.extrap = (o.flags & AG_PR_FLAG_EXTRAPOLATE) ? 1u : 0u;
    // DATA_DRIVEN zero rows never change rank; with epsilon > 0 they never take
    // the "updated" mark either, so both buffers can hold them from init.
    // A warm start's zero rows do change (from the caller's value to base)
    // in the first sweep, so they are swept.
    a.skip_zero = !(o.flags & AG_PR_FLAG_WARM_START) &&
                          ((o.mode == AG_PR_REFERENCE && o.precision == AG_PREC_FP32) ||
                           (o.mode == AG_PR_DATA_DRIVEN && o.epsilon > 0.0))
                      ? 1u
                      : 0u;
    // STANDARD ranks are ~1/N: scale FP16 words by 2^ceil(log2 N)
    a.f16_scale = o.mode == AG_PR_STANDARD ? std::ldexp(1.0f, (int)bits_for(p->N)) : 1.0f;
    a.f16_inv = 1.0f / a.f16_scale;
    a.damp = o.damp;
    {
#pragma clang fp contract(off)
        a.base = 1.0f - o.damp;
    }
    a.epsilon = o.epsilon;
    a.inv_n = p->N ? 1.0 / (double)p->N : 0.0;
    a.roff = p->roff.as<uint32_t>();
    a.cols = p->cols.as<uint32_t>();
    a.slice_ptr = p->slice_ptr.as<uint32_t>();
    a.outdeg = p->outdeg.as<uint32_t>();
    a.chunk_row = p->chunk_row.as<uint32_t>();
    a.chunk_first = p->chunk_first.as<uint32_t>();
    a.hub_partial = p->hub_partial.as<float>();
    a.nparts = w.grid;
    a.fin_hubs = 1;
    if (p->tab) {
        // partitioned tables: hub rows are summed from their table chunks by
        // the combine kernel
        const TabPlan &tp = *p->tab;
        a.fin_hubs = 0;
        a.tab_hub_blocks = tp.hub_blocks;
        a.tab_T = tp.T;
        a.tab_nt = tp.nt;
        a.tab_NT = tp.NT;
        a.tab_cold_base4 = tp.cold_base4;
        a.tab_nwaves = tp.nwaves;
        a.tab_lane_max4 = p->lane_max / 4;
        a.tab_cold_unit0 = tp.cold_unit0;
        a.tab_cold_lane_u0 = tp.cold_lane_u0;
        a.tab_cold_lane_s0 = tp.cold_lane_s0;
        a.tab_cold_lane_s1 = tp.cold_lane_s1;
        a.tab_nunits = tp.nunits;
        a.tab_idx16 = tp.idx16.as<uint16_t>();
        a.tab_idx32 = tp.idx32.as<uint32_t>();
        a.tab_sub_start = tp.sub_start.as<uint32_t>();
        a.tab_part = tp.part.as<float>();
        a.tab_units = tp.units.as<uint4>();
        a.tab_wg_begin = tp.wg_begin.as<uint32_t>();
        a.tab_items = tp.items.as<uint4>();
        a.tab_hub_items = tp.hub_items.as<uint4>();
        a.tab_wtab = tp.wtab.as<uint32_t>();
        a.chunk_first = tp.chunk_first.as<uint32_t>();
        a.hub_partial = tp.hub_partial.as<float>();
        a.hub_chunk = tp.hub_chunk;
        a.hub_chunks = tp.hub_chunks;
        a.tab_row_blocks = tp.row_blocks;
        a.nparts = tab_comb_blocks(p, a.skip_zero != 0);
    }
    a.hot_bytes = p->hot_bytes;
    a.pr = w.pr;
    a.c0 = w.c0;
    a.c1 = w.c1;
    a.blk_err = w.blk_err;
    a.blk_dang = w.blk_dang;
    a.blk_act = w.blk_act;
    a.st = w.st;
    a.err_iter = w.err_iter;
    a.act_iter = w.act_iter;
    a.fmt = w.fmt;
    return a;
}

// Enqueues one sweep's gather kernels: with the row layout the sweep kernel
// of each storage format the run can read (a mixed-precision run launches the
// low-precision and the FP32 kernel; the one whose format does not match the
// sweep exits at once); with the partitioned tables the table kernel(s), then
// the combine kernel over a.nparts blocks (its residual partials).  `grid` is
// the row-layout sweep grid.
// Enqueues sweep `it`'s gather.  With the partitioned tables and fin_prev set,
// the table kernel also runs sweep it - 1's finalize (so that sweep's
// k_pr_finalize launch is left out); the row layout ignores fin_prev.
static void enqueue_gather(hipStream_t s, const Plan *p, const SweepArgs &a, uint32_t mode,
                           uint32_t precision, uint32_t flags, uint32_t grid, uint32_t it,
                           uint32_t fin_prev)
{
    const bool dd = mode == AG_PR_DATA_DRIVEN;
    const bool fp32 = !dd && (precision == AG_PREC_FP32 || (flags & AG_PR_FLAG_FINISH_FP32));
    if (p->tab) {
        const TabPlan &tp = *p->tab;
        const uint32_t lds = (tp.T + 1) * 4;
        // of two table kernels (low-precision run with the FP32 finish) only
        // the one whose format the sweep reads passes its entry check and
        // takes tickets, so the finalize runs once
        if (dd)
            hipLaunchKernelGGL((k_pr_tables<float, true>), dim3(tp.grid), dim3(kSweepBlock), lds, s,
                               a, it, fin_prev);
        else if (precision == AG_PREC_BF16)
            hipLaunchKernelGGL((k_pr_tables<bf16_t, false>), dim3(tp.grid), dim3(kSweepBlock), lds,
                               s, a, it, fin_prev);
        else if (precision == AG_PREC_FP16)
            hipLaunchKernelGGL((k_pr_tables<__half, false>), dim3(tp.grid), dim3(kSweepBlock), lds,
                               s, a, it, fin_prev);
        if (fp32)
            hipLaunchKernelGGL((k_pr_tables<float, false>), dim3(tp.grid), dim3(kSweepBlock), lds,
                               s, a, it, fin_prev);
        // the combine kernel's table loops are unrolled to the records' NTpad
        const void *kc;
        switch (tp.NTpad) {
        case 4: kc = dd ? (const void *)k_pr_combine<true, 4, 4> : (const void *)k_pr_combine<false, 4, 4>; break;
        case 8: kc = dd ? (const void *)k_pr_combine<true, 8, 8> : (const void *)k_pr_combine<false, 8, 8>; break;
        case 12: kc = dd ? (const void *)k_pr_combine<true, 12, 12> : (const void *)k_pr_combine<false, 12, 12>; break;
        case 17: kc = dd ? (const void *)k_pr_combine<true, 17, 17> : (const void *)k_pr_combine<false, 17, 17>; break;
        case 24: kc = dd ? (const void *)k_pr_combine<true, 24, 12> : (const void *)k_pr_combine<false, 24, 12>; break;
        case 34: kc = dd ? (const void *)k_pr_combine<true, 34, 17> : (const void *)k_pr_combine<false, 34, 17>; break;
        case 51: kc = dd ? (const void *)k_pr_combine<true, 51, 17> : (const void *)k_pr_combine<false, 51, 17>; break;
        default: kc = dd ? (const void *)k_pr_combine<true, 68, 17> : (const void *)k_pr_combine<false, 68, 17>;
        }
        void *kargs[] = {(void *)&a, (void *)&it};
        AG_HIP(hipLaunchKernel(kc, dim3(a.nparts), dim3(kCombBlock), kargs, 0, s));
        return;
    }
    if (dd)
        hipLaunchKernelGGL((k_pr_sweep<float, true>), dim3(grid), dim3(kSweepBlock), p->hot_bytes,
                           s, a, it);
    else if (precision == AG_PREC_BF16)
        hipLaunchKernelGGL((k_pr_sweep<bf16_t, false>), dim3(grid), dim3(kSweepBlock),
                           p->hot_bytes, s, a, it);
    else if (precision == AG_PREC_FP16)
        hipLaunchKernelGGL((k_pr_sweep<__half, false>), dim3(grid), dim3(kSweepBlock),
                           p->hot_bytes, s, a, it);
    if (fp32)
        hipLaunchKernelGGL((k_pr_sweep<float, false>), dim3(grid), dim3(kSweepBlock), p->hot_bytes,
                           s, a, it);
}

// Enqueues sweeps [it0, it1): the gather kernels, then the finalize kernel.
// `sev` (two events per sweep, or null) brackets the gather kernels.
static void enqueue_sweeps(hipStream_t s, const Plan *p, const SweepArgs &a,
                           const ag_pr_options &o, uint32_t grid, uint32_t it0, uint32_t it1,
                           hipEvent_t *sev)
{
    // partitioned tables: sweep it's finalize runs in sweep it + 1's table
    // kernel, the batch's last one as its own kernel
    const bool fused = p->tab != nullptr;
    for (uint32_t it = it0; it < it1; it++) {
        if (sev) AG_HIP(hipEventRecord(sev[2 * it], s));
        enqueue_gather(s, p, a, o.mode, o.precision, o.flags, grid, it, fused && it > it0 ? 1u : 0u);
        if (sev) AG_HIP(hipEventRecord(sev[2 * it + 1], s));
        if (!fused || it + 1 == it1)
            hipLaunchKernelGGL(k_pr_finalize, dim3(1), dim3(kFinBlock), 0, s, a, it);
    }
    AG_HIP(hipGetLastError());
}

// Sweep loops are captured into hipGraphs and replayed unless AG_NO_GRAPH=1
// is set in the environment (read once).
static bool loop_graphs_enabled()
{
    static const bool on = [] {
        const char *e = std::getenv("AG_NO_GRAPH");
        return !(e && e[0] == '1');
    }();
    return on;
}

// Runs `enqueue` on stream s through a captured graph cached on the plan
// under `key` (captured on first use, replayed after).  The captured work
// holds kernel launches only: no memset nodes.
template <typename F>
static void run_captured(Plan *p, hipStream_t s, const std::vector<unsigned char> &key, F &&enqueue)
{
    for (size_t i = 0; i < p->loop_graphs.size(); i++) {
        if (p->loop_graphs[i].key == key) {
            AG_HIP(hipGraphLaunch(p->loop_graphs[i].exec, s));
            return;
        }
    }
    LoopGraph lg;
    lg.key = key;
    PhaseRange phase_("pagerank.graph_capture");
    AG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        enqueue();
    } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(s, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
    }
    AG_HIP(hipStreamEndCapture(s, &lg.graph));
    const hipError_t e = hipGraphInstantiate(&lg.exec, lg.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(lg.graph);
        fail(AG_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    if (p->loop_graphs.size() >= Plan::kLoopGraphs) {
        LoopGraph &old = p->loop_graphs.front();
        (void)hipGraphExecDestroy(old.exec);
        (void)hipGraphDestroy(old.graph);
        p->loop_graphs.erase(p->loop_graphs.begin());
    }
    p->loop_graphs.push_back(std::move(lg));
    AG_HIP(hipGraphLaunch(p->loop_graphs.back().exec, s));
}

template <typename T> static void key_add(std::vector<unsigned char> &k, const T &v)
{
    const unsigned char *b = reinterpret_cast<const unsigned char *>(&v);
    k.insert(k.end(), b, b + sizeof(T));
}

// Runs the sweep loop; ranks land in d_out (device, caller order).
//
// Three ways to run it:
//   * eager, with HIP events (AG_PR_FLAG_TIME_SWEEPS, or AG_NO_GRAPH=1, or a
//     personalized query, whose seed table changes per call);
//   * replayed from hipGraphs cached on the plan: one graph for init + all
//     sweeps + finish when the sweep count is fixed (epsilon <= 0), else one
//     per group of 16 sweeps with the convergence check on the host between;
//   * on caller-provided per-call buffers (`fixed_work`, from the
//     allocation-free entry point): the graph's plan must exist already and no
//     event, graph or buffer is created.  Timing then comes from the host clock.
static void run_pagerank_impl(ag_ctx *ctx, ag_graph *g, const ag_pr_options &o,
                              const SeedTable *seeds, float *d_out, uint32_t *active_iter,
                              double *err_iter, ag_pr_result *res,
                              const WorkPtrs *fixed_work = nullptr)
{
    const bool no_alloc = fixed_work != nullptr;
    hipStream_t s = ctx->stream;
    ag_pr_result r;
    std::memset(&r, 0, sizeof(r));
    Plan *p = nullptr;
    if (no_alloc) {
        p = g->plan;
        if (!p) fail(AG_ERR_INVALID, "no gather layout: call ag_pagerank_prepare first");
    } else {
        p = ensure_plan(ctx, g, &r.plan_ms);
    }
    std::lock_guard<std::mutex> lk(p->mu);
    const uint32_t N = p->N;
    const uint32_t max_iter = o.iterations;
    if (o.mode != AG_PR_REFERENCE && o.mode != AG_PR_STANDARD && o.mode != AG_PR_DATA_DRIVEN &&
        o.mode != kModePPR)
        fail(AG_ERR_INVALID, "unknown PageRank mode %u", o.mode);
    const bool dd = o.mode == AG_PR_DATA_DRIVEN;
    if (dd && (o.precision != AG_PREC_FP32 || (o.flags & AG_PR_FLAG_FINISH_FP32)))
        fail(AG_ERR_INVALID, "data-driven pull stores FP32 contributions only");
    if ((o.mode == kModePPR) != (seeds != nullptr))
        fail(AG_ERR_INVALID, "personalized mode needs a seed table");
    if ((o.flags & AG_PR_FLAG_EXTRAPOLATE) && o.mode != kModePPR && o.mode != AG_PR_STANDARD)
        fail(AG_ERR_INVALID, "extrapolated steps are offered for STANDARD and personalized PageRank only");
    const bool warm = (o.flags & AG_PR_FLAG_WARM_START) != 0;
    if (warm && o.mode == kModePPR)
        fail(AG_ERR_INVALID, "personalized PageRank starts at its seed vector: no warm start");
    if (o.precision > AG_PREC_FP16) fail(AG_ERR_INVALID, "unknown precision %u", o.precision);
    if (!(o.damp > 0.0f && o.damp < 1.0f)) fail(AG_ERR_INVALID, "damping must be in (0, 1)");
    if (N == 0) {
        if (res) *res = r;
        return;
    }
    if (warm && !d_out) fail(AG_ERR_INVALID, "a warm start reads the start from the output buffer");
    WorkPtrs w;
    if (no_alloc) {
        w = *fixed_work;
    } else {
        ensure_work(ctx, p, max_iter);
        w = work_ptrs(p->work);
    }
    SweepArgs a = make_args(p, w, o, max_iter);
    if (seeds) {
        a.seed_bits = seeds->bits;
        a.seed_v = seeds->v;
        a.seed_w = seeds->w;
        a.nseeds = seeds->n;
    }
    const bool ref_scale = o.mode == AG_PR_REFERENCE || dd; // pr0 = base, ranks pr/N
    const float x0 = ref_scale ? a.base : (float)(1.0 / (double)N);
    const double dang0 = (double)p->n_dangling * (1.0 / (double)N);
    const float tele0 = (float)((1.0 - (double)o.damp) / N + (double)o.damp * dang0 / N);
    const bool timed = !no_alloc && (o.flags & AG_PR_FLAG_TIME_SWEEPS) != 0;
    const bool use_graph = !no_alloc && !timed && !seeds && loop_graphs_enabled();
    const bool can_stop = o.epsilon > 0.0;
    const uint32_t group = can_stop ? 16u : max_iter;
    const uint32_t grid = w.grid;
    const uint32_t nbf = ceil_div(N, kBlock);
    double *parts = w.sum_part;

    auto enqueue_init = [&] {
        hipLaunchKernelGGL(k_pr_init, dim3(ceil_div((uint64_t)std::max(N + 1, max_iter + 2), 256)),
                           dim3(256), 0, s, a, x0, tele0, o.precision, warm ? 1u : 0u);
        if (warm) {
            hipLaunchKernelGGL(k_pr_init_from, dim3(nbf), dim3(kBlock), 0, s, a,
                               (const float *)d_out, p->rank.as<uint32_t>(),
                               ref_scale ? (float)N : 1.0f, o.precision, parts);
            hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kSumBlock), 0, s, parts, nbf, parts + nbf);
            hipLaunchKernelGGL(k_pr_warm_tele, dim3(1), dim3(64), 0, s, a, (const double *)parts + nbf);
        }
        AG_HIP(hipGetLastError());
    };
    auto enqueue_finish = [&] {
        hipLaunchKernelGGL(k_pr_finish, dim3(nbf), dim3(kBlock), 0, s, w.pr, p->rank.as<uint32_t>(),
                           N, ref_scale ? 1u : 0u, d_out, parts);
        hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kSumBlock), 0, s, parts, nbf, parts + nbf);
        AG_HIP(hipGetLastError());
    };
    std::vector<unsigned char> key0;
    if (use_graph) {
        key_add(key0, a);
        key_add(key0, x0);
        key_add(key0, tele0);
        key_add(key0, d_out);
        key_add(key0, grid);
        key_add(key0, p->hot_bytes);
        key_add(key0, o.precision);
        key_add(key0, o.flags);
    }
    auto seg_key = [&](uint32_t seg) {
        std::vector<unsigned char> k = key0;
        key_add(k, seg);
        return k;
    };

    // Host-side events and timers (none in the allocation-free mode).
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    if (!no_alloc)
        for (auto &e : ev) AG_HIP(hipEventCreate(&e));
    struct EvGuard { hipEvent_t *e; ~EvGuard() { for (int i = 0; i < 4; i++) if (e[i]) (void)hipEventDestroy(e[i]); } } eg{ev};
    std::vector<hipEvent_t> sev(timed ? 2 * (size_t)max_iter : 0);
    for (auto &e : sev) AG_HIP(hipEventCreate(&e));
    struct SevGuard { std::vector<hipEvent_t> &v; ~SevGuard() { for (auto e : v) (void)hipEventDestroy(e); } } sg{sev};
    const auto h0 = std::chrono::steady_clock::now();

    if (ev[0]) AG_HIP(hipEventRecord(ev[0], s));
    if (use_graph && !can_stop) {
        // one replayable graph: init, every sweep, finish
        run_captured(p, s, seg_key(0), [&] {
            enqueue_init();
            enqueue_sweeps(s, p, a, o, grid, 0, max_iter, nullptr);
            enqueue_finish();
        });
        if (ev[1]) AG_HIP(hipEventRecord(ev[1], s));
        if (ev[2]) AG_HIP(hipEventRecord(ev[2], s));
    } else {
        if (!use_graph) enqueue_init();
        if (ev[1]) AG_HIP(hipEventRecord(ev[1], s));
        for (uint32_t it0 = 0; it0 < max_iter; it0 += group) {
            const uint32_t it1 = std::min(max_iter, it0 + group);
            if (use_graph) {
                run_captured(p, s, seg_key(1 + it0 / group), [&] {
                    if (it0 == 0) enqueue_init();
                    enqueue_sweeps(s, p, a, o, grid, it0, it1, nullptr);
                });
            } else {
                enqueue_sweeps(s, p, a, o, grid, it0, it1, timed ? sev.data() : nullptr);
            }
            if (can_stop && it1 < max_iter) {
                PhaseRange phase_("pagerank.stop_check");
                uint32_t done = 0;
                AG_HIP(hipMemcpyAsync(&done, &a.st->done, 4, hipMemcpyDeviceToHost, s));
                AG_HIP(hipStreamSynchronize(s));
                if (done) break;
            }
        }
        if (ev[2]) AG_HIP(hipEventRecord(ev[2], s));
        if (use_graph) run_captured(p, s, seg_key(0xffffffffu), enqueue_finish);
        else enqueue_finish();
    }
    if (ev[3]) AG_HIP(hipEventRecord(ev[3], s));
    SweepState st;
    AG_HIP(hipMemcpyAsync(&st, a.st, sizeof(st), hipMemcpyDeviceToHost, s));
    AG_HIP(hipMemcpyAsync(&r.pr_sum, parts + nbf, 8, hipMemcpyDeviceToHost, s));
    AG_HIP(hipStreamSynchronize(s));
    if (active_iter && st.sweeps)
        AG_HIP(hipMemcpyAsync(active_iter, a.act_iter, (size_t)st.sweeps * 4,
                              hipMemcpyDeviceToHost, s));
    if (err_iter && st.sweeps)
        AG_HIP(hipMemcpyAsync(err_iter, a.err_iter, (size_t)st.sweeps * 8,
                              hipMemcpyDeviceToHost, s));
    AG_HIP(hipStreamSynchronize(s));
    if (no_alloc) {
        const double ms = std::chrono::duration<double, std::milli>(
                              std::chrono::steady_clock::now() - h0).count();
        r.loop_ms = ms;
        r.total_ms = ms;
    } else {
        float ms = 0.f;
        // a single replayed graph holds init and finish too: loop_ms then
        // spans the whole replay
        AG_HIP(hipEventElapsedTime(&ms, (use_graph && !can_stop) ? ev[0] : ev[1],
                                   (use_graph && !can_stop) ? ev[3] : ev[2]));
        r.loop_ms = ms;
        AG_HIP(hipEventElapsedTime(&ms, ev[0], ev[3]));
        r.total_ms = ms;
    }
    r.sweep_kernel_ms = 0.0;
    for (uint32_t it = 0; timed && it < st.sweeps && it < max_iter; it++) {
        float ms = 0.f;
        AG_HIP(hipEventElapsedTime(&ms, sev[2 * it], sev[2 * it + 1]));
        r.sweep_kernel_ms += ms;
    }
    r.iterations = st.iterations;
    r.sweeps = st.sweeps;
    r.active_vertices = st.last_active;
    r.error_total = st.last_err;
    r.fp32_switch_sweep = st.fp32_switch;
    if (res) *res = r;
}

static void run_pagerank(ag_ctx *ctx, ag_graph *g, const ag_pr_options &o, float *d_out,
                         uint32_t *active_iter, double *err_iter, ag_pr_result *res)
{
    run_pagerank_impl(ctx, g, o, nullptr, d_out, active_iter, err_iter, res);
}

void run_pagerank_seeded(ag_ctx *ctx, ag_graph *g, const ag_pr_options &o, const SeedTable &st,
                         float *d_out, ag_pr_result *res)
{
    run_pagerank_impl(ctx, g, o, &st, d_out, nullptr, nullptr, res);
}

// SPMV pull (SPMV.c:622-750): `iterations` sweeps of y = A^T x over the
// gather layout, as the reference restarts its compute unit once per
// iteration on the same input.  d_x / d_y are device buffers in caller order.
static void run_spmv(ag_ctx *ctx, ag_graph *g, const float *d_x, float *d_y,
                     uint32_t iterations, ag_spmv_result *res)
{
    hipStream_t s = ctx->stream;
    ag_spmv_result r;
    std::memset(&r, 0, sizeof(r));
    Plan *p = ensure_plan(ctx, g, &r.plan_ms);
    std::lock_guard<std::mutex> lk(p->mu);
    const uint32_t N = p->N;
    if (N == 0) {
        r.iterations = iterations;
        if (res) *res = r;
        return;
    }
    if (iterations == 0) { // the compute unit never runs: the output stays zero
        AG_HIP(hipMemsetAsync(d_y, 0, (size_t)N * 4, s));
        AG_HIP(hipStreamSynchronize(s));
        if (res) *res = r;
        return;
    }
    ensure_work(ctx, p, iterations);
    const WorkPtrs w = work_ptrs(p->work);
    ag_pr_options o;
    ag_pr_options_default(&o);
    o.mode = kModeSpmv;
    o.epsilon = 0.0;
    SweepArgs a = make_args(p, w, o, iterations);
    a.c1 = a.c0; // every sweep reads the same input
    a.skip_zero = 0;
    if (p->tab) a.nparts = tab_comb_blocks(p, false);
    hipEvent_t ev[4];
    for (auto &e : ev) AG_HIP(hipEventCreate(&e));
    struct EvGuard { hipEvent_t *e; ~EvGuard() { for (int i = 0; i < 4; i++) (void)hipEventDestroy(e[i]); } } eg{ev};
    AG_HIP(hipEventRecord(ev[0], s));
    hipLaunchKernelGGL(k_spmv_init, dim3(ceil_div((uint64_t)std::max(N + 1, iterations + 2), 256)),
                       dim3(256), 0, s, a, d_x, p->order.as<uint32_t>());
    AG_HIP(hipEventRecord(ev[1], s));
    const bool fused = p->tab != nullptr; // as in enqueue_sweeps
    for (uint32_t it = 0; it < iterations; it++) {
        enqueue_gather(s, p, a, kModeSpmv, AG_PREC_FP32, 0u, w.grid, it, fused && it > 0 ? 1u : 0u);
        if (!fused || it + 1 == iterations)
            hipLaunchKernelGGL(k_pr_finalize, dim3(1), dim3(kFinBlock), 0, s, a, it);
    }
    AG_HIP(hipGetLastError());
    AG_HIP(hipEventRecord(ev[2], s));
    const uint32_t nbf = ceil_div(N, kBlock);
    double *parts = w.sum_part;
    hipLaunchKernelGGL(k_pr_finish, dim3(nbf), dim3(kBlock), 0, s, w.pr,
                       p->rank.as<uint32_t>(), N, 0u, d_y, parts);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kSumBlock), 0, s, parts, nbf, parts + nbf);
    AG_HIP(hipGetLastError());
    AG_HIP(hipEventRecord(ev[3], s));
    AG_HIP(hipMemcpyAsync(&r.y_sum, parts + nbf, 8, hipMemcpyDeviceToHost, s));
    AG_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    AG_HIP(hipEventElapsedTime(&ms, ev[1], ev[2]));
    r.loop_ms = ms;
    AG_HIP(hipEventElapsedTime(&ms, ev[0], ev[3]));
    r.total_ms = ms;
    r.iterations = iterations;
    if (res) *res = r;
}

// ---- allocation-free entry point: static device scratch ----------------------
// ag_pagerank_pull_device_noalloc takes its per-call buffers (ranks in label
// order, both contribution buffers, per-workgroup partials, sweep state,
// per-sweep arrays, reduction partials) from `scratch`, a static device array
// of kNoAllocBytes that every device the library runs on holds until the
// process ends.  In host code the array's name is only a host-side stand-in:
// its device address, per device, comes from hipGetSymbolAddress at context
// creation (noalloc_scratch_init) and only that address is used afterwards.
constexpr size_t kNoAllocBytes = (size_t)16 << 20;
constexpr int kNoAllocMaxDevices = 64;

} // namespace ag

static __device__ __align__(256) unsigned char scratch[ag::kNoAllocBytes];

namespace ag {

namespace {
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kNoAllocMaxDevices] = {};
} // namespace

// Called by context creation with `device` current, once per device.
void noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kNoAllocMaxDevices) return; // the entry point then refuses
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return;
    void *ptr = nullptr;
    AG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(::scratch)));
    if (!ptr) fail(AG_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    g_noalloc_ptr[device] = ptr;
}

static unsigned char *noalloc_scratch_ptr(int device)
{
    if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return static_cast<unsigned char *>(g_noalloc_ptr[device]);
}

// Carves the per-call buffers of the sweep loop out of the device's scratch
// (256-byte aligned pieces); refuses sizes beyond kNoAllocBytes.
static WorkPtrs noalloc_work(ag_ctx *ctx, const Plan *p, uint32_t max_iter)
{
    const uint32_t grid = sweep_grid(ctx, p);
    const size_t parts = part_slots(ctx, p);
    const size_t N = p->N;
    const size_t sizes[] = {N * 4, (N + 1) * 4, (N + 1) * 4, parts * 8, parts * 8,
                            parts * 4, sizeof(SweepState), ((size_t)max_iter + 1) * 8,
                            ((size_t)max_iter + 1) * 4, ((size_t)max_iter + 3) * 4,
                            ((size_t)ceil_div(N, kBlock) + 1) * 8 + 8};
    size_t need = 0;
    for (size_t b : sizes) need += (b + 255) & ~(size_t)255;
    if (need > kNoAllocBytes)
        fail(AG_ERR_INVALID,
             "allocation-free PageRank needs %zu bytes of scratch for N=%zu and %u sweeps; the "
             "static scratch holds %zu (about 1.3 M vertices at 20 sweeps)",
             need, N, max_iter, kNoAllocBytes);
    unsigned char *base = noalloc_scratch_ptr(ctx->device);
    if (!base) fail(AG_ERR_INVALID, "no allocation-free scratch on device %d", ctx->device);
    size_t off = 0;
    void *q[11];
    for (int i = 0; i < 11; i++) {
        q[i] = base + off;
        off += (sizes[i] + 255) & ~(size_t)255;
    }
    WorkPtrs w;
    w.grid = grid;
    w.pr = static_cast<float *>(q[0]);
    w.c0 = q[1];
    w.c1 = q[2];
    w.blk_err = static_cast<double *>(q[3]);
    w.blk_dang = static_cast<double *>(q[4]);
    w.blk_act = static_cast<uint32_t *>(q[5]);
    w.st = static_cast<SweepState *>(q[6]);
    w.err_iter = static_cast<double *>(q[7]);
    w.act_iter = static_cast<uint32_t *>(q[8]);
    w.fmt = static_cast<uint32_t *>(q[9]);
    w.sum_part = static_cast<double *>(q[10]);
    return w;
}

} // namespace ag

using namespace ag;

extern "C" {

void ag_pr_options_default(ag_pr_options *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->iterations = 20;
    o->mode = AG_PR_REFERENCE;
    o->precision = AG_PREC_FP32;
    o->epsilon = 1e-4;
    o->damp = 0.85f;
}

int ag_pagerank_prepare(ag_ctx *ctx, ag_graph *g, double *ms)
{
    return guard([&] {
        PhaseRange phase_("ag_pagerank_prepare");
        check_ctx(ctx);
        if (!g) fail(AG_ERR_INVALID, "null graph");
        ensure_plan(ctx, g, ms);
    });
}

int ag_pagerank_autotune(ag_ctx *ctx, ag_graph *g, uint32_t sweeps, ag_tuning *best, double *ms)
{
    return guard([&] {
        PhaseRange phase_("ag_pagerank_autotune");
        check_ctx(ctx);
        if (!g) fail(AG_ERR_INVALID, "null graph");
        if (sweeps == 0 || sweeps > 1000) fail(AG_ERR_INVALID, "sweeps must be in 1..1000");
        const auto t0 = std::chrono::steady_clock::now();
        Tuning cur = g->has_tune ? g->tune : tuning();
        // a layout built from `tn` becomes g's plan (the previous one is kept
        // by the caller or destroyed)
        auto build = [&](const Tuning &tn) {
            g->tune = tn;
            g->has_tune = true;
            g->plan = nullptr;
            return ensure_plan(ctx, g, nullptr);
        };
        ag_pr_options o;
        ag_pr_options_default(&o);
        o.iterations = sweeps;
        o.epsilon = 0.0;
        DevBuf out((size_t)std::max<uint32_t>(g->N, 1) * 4);
        // FP32 pull, `sweeps` sweeps: the first call captures the loop's
        // graph, the faster of two replays is the candidate's time
        auto measure = [&] {
            ag_pr_result r;
            double t = 1e300;
            for (int k = 0; k < 3; k++) {
                run_pagerank(ctx, g, o, out.as<float>(), nullptr, nullptr, &r);
                if (k > 0) t = std::min(t, r.total_ms);
            }
            return t;
        };
        plan_destroy(g->plan);
        Plan *best_plan = build(cur);
        double best_ms = g->N && g->E ? measure() : 0.0;
        Tuning best_t = cur;
        if (g->N && g->E) {
            struct Knob { uint32_t Tuning::*f; std::vector<uint32_t> vals; };
            std::vector<Knob> knobs;
            if (best_plan->tab) {
                knobs.push_back({&Tuning::max_tables, {8, 12, 16, 20, 24, 31, 40, 48, 62}});
                knobs.push_back({&Tuning::lane_max, {8, 12, 16, 24}});
            } else {
                knobs.push_back({&Tuning::low_max, {16, 32, 64, 128}});
            }
            knobs.push_back({&Tuning::hub_min, {4096, 6144, 8192, 12288}});
            knobs.push_back({&Tuning::hub_chunk, {AG_TUNE_AUTO, 1024, 2048, 4096}});
            // one pass of coordinate search; a candidate replaces the best
            // only when at least 1% faster (timing noise is below that)
            for (const Knob &k : knobs)
                for (uint32_t v : k.vals) {
                    if (best_t.*k.f == v) continue;
                    Tuning t = best_t;
                    t.*k.f = v;
                    Plan *cand = build(t);
                    const double m = measure();
                    if (m < 0.99 * best_ms) {
                        plan_destroy(best_plan);
                        best_plan = cand;
                        best_ms = m;
                        best_t = t;
                    } else {
                        plan_destroy(cand);
                    }
                    g->plan = nullptr;
                }
        }
        g->plan = best_plan;
        g->tune = best_t;
        g->has_tune = true;
        g->plan->autotuned = true;
        if (best) {
            std::memset(best, 0, sizeof(*best));
            best->low_max = best_t.low_max;
            best->hub_min = best_t.hub_min;
            best->hub_chunk = best_t.hub_chunk;
            best->hot_bytes = best_t.hot_bytes;
            best->layout = best_t.layout;
            best->max_tables = best_t.max_tables;
            best->lane_max = best_t.lane_max;
        }
        if (ms)
            *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
}

int ag_pagerank_plan_info(const ag_graph *g, ag_plan_info *info)
{
    return guard([&] {
        if (!g || !info) fail(AG_ERR_INVALID, "null argument");
        std::memset(info, 0, sizeof(*info));
        const Plan *p = g->plan;
        if (!p) fail(AG_ERR_INVALID, "no plan built yet (call ag_pagerank_prepare)");
        info->n_hub = p->n_hub;
        info->n_mid = p->n_nonlow - p->n_hub;
        info->n_low = p->n_nonzero - p->n_nonlow;
        info->n_zero = p->N - p->n_nonzero;
        info->hub_chunks = p->hub_chunks;
        info->low_max = p->low_max;
        info->hub_min = p->hub_min;
        info->sell_slots = p->sell_slots;
        info->plan_bytes = p->bytes();
        info->hub_chunk = p->tab ? p->tab->hub_chunk : p->hub_chunk;
        info->layout = p->tab ? AG_LAYOUT_TABLES : AG_LAYOUT_ROWS;
        info->lane_max = p->lane_max;
        info->max_tables = p->max_tables;
        info->autotuned = p->autotuned ? 1u : 0u;
        if (p->tab) {
            info->hub_chunks = p->tab->hub_chunks;
            info->tables = p->tab->nt;
            info->table_workgroups = p->tab->grid;
            info->subrows = p->tab->S;
            info->cold_edges = p->tab->tab_edges.empty() ? 0 : p->tab->tab_edges.back();
        }
    });
}

int ag_pagerank_pull(ag_ctx *ctx, ag_graph *g, const ag_pr_options *opt, float *ranks,
                     uint32_t *active_iter, double *err_iter, ag_pr_result *res)
{
    return guard([&] {
        PhaseRange phase_("ag_pagerank_pull");
        check_ctx(ctx);
        if (!g || !opt) fail(AG_ERR_INVALID, "null argument");
        Plan *p = ensure_plan(ctx, g, nullptr);
        if (g->N == 0) {
            run_pagerank(ctx, g, *opt, nullptr, active_iter, err_iter, res);
            return;
        }
        ensure_work(ctx, p, opt->iterations);
        if (opt->flags & AG_PR_FLAG_WARM_START) {
            if (!ranks) fail(AG_ERR_INVALID, "a warm start reads the start from `ranks`");
            AG_HIP(hipMemcpyAsync(p->work.out.p, ranks, (size_t)g->N * 4, hipMemcpyHostToDevice,
                                  ctx->stream));
        }
        run_pagerank(ctx, g, *opt, p->work.out.as<float>(), active_iter, err_iter, res);
        if (ranks) {
            AG_HIP(hipMemcpyAsync(ranks, p->work.out.p, (size_t)g->N * 4, hipMemcpyDeviceToHost,
                                  ctx->stream));
            AG_HIP(hipStreamSynchronize(ctx->stream));
        }
    });
}

int ag_spmv_pull(ag_ctx *ctx, ag_graph *g, const float *x, float *y, uint32_t iterations,
                 ag_spmv_result *res)
{
    return guard([&] {
        PhaseRange phase_("ag_spmv_pull");
        check_ctx(ctx);
        if (!g) fail(AG_ERR_INVALID, "null graph");
        if (g->N && (!x || !y)) fail(AG_ERR_INVALID, "null vector");
        if (g->N == 0) {
            run_spmv(ctx, g, nullptr, nullptr, iterations, res);
            return;
        }
        DevBuf dx((size_t)g->N * 4), dy((size_t)g->N * 4);
        AG_HIP(hipMemcpyAsync(dx.p, x, (size_t)g->N * 4, hipMemcpyHostToDevice, ctx->stream));
        run_spmv(ctx, g, dx.as<float>(), dy.as<float>(), iterations, res);
        AG_HIP(hipMemcpyAsync(y, dy.p, (size_t)g->N * 4, hipMemcpyDeviceToHost, ctx->stream));
        AG_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int ag_spmv_pull_device(ag_ctx *ctx, ag_graph *g, const float *d_x, float *d_y,
                        uint32_t iterations, ag_spmv_result *res)
{
    return guard([&] {
        PhaseRange phase_("ag_spmv_pull_device");
        check_ctx(ctx);
        if (!g) fail(AG_ERR_INVALID, "null graph");
        if (g->N && (!d_x || !d_y)) fail(AG_ERR_INVALID, "null vector");
        if (d_x == d_y && g->N) fail(AG_ERR_INVALID, "input and output must not alias");
        run_spmv(ctx, g, d_x, d_y, iterations, res);
    });
}

int ag_pagerank_pull_device(ag_ctx *ctx, ag_graph *g, const ag_pr_options *opt, float *d_ranks,
                            ag_pr_result *res)
{
    return guard([&] {
        PhaseRange phase_("ag_pagerank_pull_device");
        check_ctx(ctx);
        if (!g || !opt) fail(AG_ERR_INVALID, "null argument");
        if (g->N && !d_ranks) fail(AG_ERR_INVALID, "null output");
        run_pagerank(ctx, g, *opt, d_ranks, nullptr, nullptr, res);
    });
}

int ag_pagerank_pull_device_noalloc(ag_ctx *ctx, ag_graph *g, const ag_pr_options *opt,
                                    float *d_ranks, ag_pr_result *res)
{
    return guard([&] {
        PhaseRange phase_("ag_pagerank_pull_device_noalloc");
        check_ctx(ctx);
        if (!g || !opt) fail(AG_ERR_INVALID, "null argument");
        if (g->N && !d_ranks) fail(AG_ERR_INVALID, "null output");
        if (opt->flags & AG_PR_FLAG_TIME_SWEEPS)
            fail(AG_ERR_INVALID, "the allocation-free entry point does not time sweeps "
                                 "(HIP events are created per call)");
        if (!g->plan)
            fail(AG_ERR_INVALID, "the allocation-free entry point needs the graph's gather "
                                 "layout: call ag_pagerank_prepare first");
        if (g->N == 0) {
            run_pagerank_impl(ctx, g, *opt, nullptr, d_ranks, nullptr, nullptr, res);
            return;
        }
        const WorkPtrs w = noalloc_work(ctx, g->plan, opt->iterations);
        run_pagerank_impl(ctx, g, *opt, nullptr, d_ranks, nullptr, nullptr, res, &w);
    });
}

} // extern "C"
