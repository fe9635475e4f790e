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
// Layout: vertex-major [N][capacity] (FP32 iterate; contributions stored in
// FP32, BF16 or FP16), in the degree-sorted labels of the graph's plan.  A
// wave is split into 64/CPW lane groups; each group covers CPW consecutive
// columns.  Hub rows are split in fixed edge chunks (one workgroup each),
// mid rows take one wave each (groups stride over the edges, shuffle
// reduction), low rows one group each in the plan's SELL-64 slice order.  A
// sweep is three launches: the sweep kernel, a hub-row kernel that sums the
// chunk partials in order, and a one-workgroup column reduction; kernel
// boundaries order them, with no device-scope fence (an L2 write-back +
// invalidate on gfx950).  Columns retire individually: a done column is
// frozen (no writes) and the host compacts slots between sweeps.  Residual and dangling sums are FP64, reduced in a fixed order, so a
// column's result does not depend on which other columns share the batch.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>

#include "ag_ppr.h"

namespace ag {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMaxCap = 64;

struct PprArgs {
    uint32_t N, cap;
    uint32_t n_hub, n_nonlow, n_nonzero, hub_chunk, nb_hub, nmid_units, nlow_units;
    float damp;
    uint32_t fmt;
    const uint32_t *__restrict__ roff;
    const uint32_t *__restrict__ cols;
    const uint32_t *__restrict__ slice_ptr;
    const uint32_t *__restrict__ outdeg;
    const uint32_t *__restrict__ chunk_row;
    const uint32_t *__restrict__ chunk_first;
    float *X;
    const void *Cin;
    void *Cout;
    const uint32_t *seed_bits, *seed_v, *seed_c;
    const float *seed_w;
    PprCol *cs;
    PprBatchState *bs;
    double *blk_res, *blk_dang; // [nparts][64]
    double *hub_res, *hub_dang; // [n_hub][64]
    float *hub_part;            // [hub chunks][cap]
    uint32_t nparts;
};

// Per-slot descriptor of one admission (or migration: e0 = destination slot).
struct PprAdmit {
    uint32_t slot, e0, e1, max_iter;
    double tol;
};

__device__ __forceinline__ float ld(const float *p, uint64_t i) { return p[i]; }
__device__ __forceinline__ float ld(const bf16_t *p, uint64_t i)
{
    return __uint_as_float((uint32_t)p[i].x << 16);
}
__device__ __forceinline__ float ld(const __half *p, uint64_t i) { return __half2float(p[i]); }

__device__ __forceinline__ void st(void *buf, uint32_t fmt, uint64_t i, float c)
{
    if (fmt == AG_PREC_FP32) {
        static_cast<float *>(buf)[i] = c;
    } else if (fmt == AG_PREC_BF16) {
        const uint32_t b = __float_as_uint(c);
        static_cast<uint16_t *>(buf)[i] = (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
    } else {
        static_cast<__half *>(buf)[i] = __float2half_rn(c);
    }
}

__device__ __forceinline__ uint32_t ldnt(const uint32_t *p) { return __builtin_nontemporal_load(p); }

template <int CPW>
__device__ __forceinline__ float group_sum(float x)
{
#pragma unroll
    for (int off = CPW; off < 64; off <<= 1) x += __shfl_xor(x, off, 64);
    return x;
}
template <int CPW>
__device__ __forceinline__ double group_sum(double x)
{
#pragma unroll
    for (int off = CPW; off < 64; off <<= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// Column update of label `row` (only called for active columns).
__device__ __forceinline__ void ppr_update(const PprArgs &a, uint32_t row, uint32_t c, float acc,
                                          float t, double &res, double &dang)
{
#pragma clang fp contract(off)
    float sv = 0.0f;
    if ((a.seed_bits[row >> 5] >> (row & 31)) & 1u) {
        const uint32_t ns = a.bs->nseeds;
        for (uint32_t q = 0; q < ns; q++)
            if (a.seed_v[q] == row && a.seed_c[q] == c) sv += a.seed_w[q];
    }
    const uint64_t i = (uint64_t)row * a.cap + c;
    const float old = a.X[i];
    const float gathered = a.damp * acc;
    const float nw = gathered + sv * t;
    a.X[i] = nw;
    res += fabs((double)nw - (double)old);
    const uint32_t od = a.outdeg[row];
    if (od == 0) dang += (double)nw;
    st(a.Cout, a.fmt, i, od ? nw / (float)od : 0.0f);
}

template <int CPW, typename CT>
__global__ __launch_bounds__(kBlock) void k_ppr_sweep(PprArgs a)
{
    constexpr int G = 64 / CPW;
    __shared__ double s_res[kMaxCap], s_dang[kMaxCap];
    __shared__ double w_res[4][CPW], w_dang[4][CPW];
    __shared__ float hub_red[4][CPW];
    const uint32_t K = a.bs->K;
    if (K == 0) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = lane / CPW, cl = lane % CPW;
    if (threadIdx.x < kMaxCap) { s_res[threadIdx.x] = 0.0; s_dang[threadIdx.x] = 0.0; }
    __syncthreads();
    const CT *__restrict__ cin = static_cast<const CT *>(a.Cin);
    const uint32_t ncg = (K + CPW - 1) / CPW;
    const uint32_t U = a.nb_hub + a.nmid_units + a.nlow_units + 1;
    const uint32_t total = U * ncg;
    double res = 0.0, dang = 0.0;
    int cur = -1;

    auto flush = [&](int cg) {
        double r = group_sum<CPW>(res), d = group_sum<CPW>(dang);
        if (g == 0) { w_res[w][cl] = r; w_dang[w][cl] = d; }
        __syncthreads();
        if (threadIdx.x < CPW) {
            const int c = cg * CPW + threadIdx.x;
            if (c < (int)kMaxCap) {
                s_res[c] += (w_res[0][threadIdx.x] + w_res[1][threadIdx.x]) +
                            (w_res[2][threadIdx.x] + w_res[3][threadIdx.x]);
                s_dang[c] += (w_dang[0][threadIdx.x] + w_dang[1][threadIdx.x]) +
                             (w_dang[2][threadIdx.x] + w_dang[3][threadIdx.x]);
            }
        }
        __syncthreads();
        res = 0.0;
        dang = 0.0;
    };

    for (uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const uint32_t u = vb / ncg, cg = vb % ncg;
        if ((int)cg != cur) {
            if (cur >= 0) flush(cur);
            cur = (int)cg;
        }
        const uint32_t c = cg * CPW + cl;
        const bool act = c < K && !a.cs[c].done && !a.cs[c].stalled;
        const float t = act ? a.cs[c].t : 0.0f;
        if (u < a.nb_hub) {
            // hub chunk: the workgroup's 4*G lane groups stride over the chunk
            const uint32_t row = a.chunk_row[u];
            const uint32_t k = u - a.chunk_first[row];
            const uint32_t rb = a.roff[row], re = a.roff[row + 1];
            const uint32_t b0 = rb + k * a.hub_chunk;
            const uint32_t b1 = min(re, b0 + a.hub_chunk);
            float acc = 0.0f;
            if (act) {
                float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
                constexpr uint32_t S = 4 * G; // lane groups of the workgroup
                uint32_t j = b0 + w * G + g;
                for (; j + 3 * S < b1; j += 4 * S) {
                    const uint32_t u0 = ldnt(a.cols + j), u1 = ldnt(a.cols + j + S);
                    const uint32_t u2 = ldnt(a.cols + j + 2 * S), u3 = ldnt(a.cols + j + 3 * S);
                    a0 += ld(cin, (uint64_t)u0 * a.cap + c);
                    a1 += ld(cin, (uint64_t)u1 * a.cap + c);
                    a2 += ld(cin, (uint64_t)u2 * a.cap + c);
                    a3 += ld(cin, (uint64_t)u3 * a.cap + c);
                }
                for (; j < b1; j += S) a0 += ld(cin, (uint64_t)ldnt(a.cols + j) * a.cap + c);
                acc = (a0 + a1) + (a2 + a3);
            }
            acc = group_sum<CPW>(acc);
            if (g == 0) hub_red[w][cl] = acc;
            __syncthreads();
            if (threadIdx.x < CPW && cg * CPW + threadIdx.x < a.cap) {
                const float tot = (hub_red[0][threadIdx.x] + hub_red[1][threadIdx.x]) +
                                  (hub_red[2][threadIdx.x] + hub_red[3][threadIdx.x]);
                a.hub_part[(uint64_t)u * a.cap + cg * CPW + threadIdx.x] = tot;
            }
            __syncthreads();
        } else if (u < a.nb_hub + a.nmid_units) {
            // mid row: one wave, lane groups stride over the row's edges
            const uint32_t row = a.n_hub + (u - a.nb_hub) * 4 + w;
            if (row < a.n_nonlow) {
                const uint32_t rb = a.roff[row], re = a.roff[row + 1];
                float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
                if (act) {
                    uint32_t j = rb + g;
                    for (; j + 3 * G < re; j += 4 * G) {
                        const uint32_t u0 = ldnt(a.cols + j), u1 = ldnt(a.cols + j + G);
                        const uint32_t u2 = ldnt(a.cols + j + 2 * G), u3 = ldnt(a.cols + j + 3 * G);
                        a0 += ld(cin, (uint64_t)u0 * a.cap + c);
                        a1 += ld(cin, (uint64_t)u1 * a.cap + c);
                        a2 += ld(cin, (uint64_t)u2 * a.cap + c);
                        a3 += ld(cin, (uint64_t)u3 * a.cap + c);
                    }
                    for (; j < re; j += G) a0 += ld(cin, (uint64_t)ldnt(a.cols + j) * a.cap + c);
                }
                const float acc = group_sum<CPW>((a0 + a1) + (a2 + a3));
                if (g == 0 && act) ppr_update(a, row, c, acc, t, res, dang);
            }
        } else if (u < a.nb_hub + a.nmid_units + a.nlow_units) {
            // low rows: one lane group per row, edges in the plan's order
            const uint32_t rl = (u - a.nb_hub - a.nmid_units) * (4 * G) + w * G + g;
            const uint32_t row = a.n_nonlow + rl;
            if (row < a.n_nonzero && act) {
                const uint32_t deg = a.roff[row + 1] - a.roff[row];
                const uint32_t *p = a.cols + a.slice_ptr[rl >> 6] + (rl & 63);
                float acc = 0.0f;
                uint32_t k = 0;
                for (; k + 4 <= deg; k += 4) {
                    const uint32_t u0 = ldnt(p + k * 64), u1 = ldnt(p + (k + 1) * 64);
                    const uint32_t u2 = ldnt(p + (k + 2) * 64), u3 = ldnt(p + (k + 3) * 64);
                    const float x0 = ld(cin, (uint64_t)u0 * a.cap + c);
                    const float x1 = ld(cin, (uint64_t)u1 * a.cap + c);
                    const float x2 = ld(cin, (uint64_t)u2 * a.cap + c);
                    const float x3 = ld(cin, (uint64_t)u3 * a.cap + c);
                    acc += x0; // edge order
                    acc += x1;
                    acc += x2;
                    acc += x3;
                }
                for (; k < deg; k++) acc += ld(cin, (uint64_t)ldnt(p + k * 64) * a.cap + c);
                ppr_update(a, row, c, acc, t, res, dang);
            }
        } else {
            // seeds without in-edges: x' = s_v * t
            if (threadIdx.x < CPW && act) {
                const uint32_t ns = a.bs->nseeds;
                for (uint32_t q = 0; q < ns; q++)
                    if (a.seed_c[q] == c && a.seed_v[q] >= a.n_nonzero)
                        ppr_update(a, a.seed_v[q], c, 0.0f, t, res, dang);
            }
        }
    }
    if (cur >= 0) flush(cur);

    // per-workgroup column partials; k_ppr_reduce sums them in a fixed order
    if (threadIdx.x < kMaxCap) {
        a.blk_res[(uint64_t)blockIdx.x * kMaxCap + threadIdx.x] = s_res[threadIdx.x];
        a.blk_dang[(uint64_t)blockIdx.x * kMaxCap + threadIdx.x] = s_dang[threadIdx.x];
    }
}

// Hub rows, one workgroup per row and one lane per column, after the sweep
// (the kernel boundary orders the chunk partials): chunk partials summed in
// chunk order, then the column update.
__global__ __launch_bounds__(kMaxCap) void k_ppr_hub(PprArgs a)
{
    const uint32_t K = a.bs->K;
    if (K == 0) return;
    const uint32_t row = blockIdx.x, c = threadIdx.x;
    double res = 0.0, dang = 0.0;
    if (c < K && !a.cs[c].done && !a.cs[c].stalled) {
        const uint32_t f0 = a.chunk_first[row], f1 = a.chunk_first[row + 1];
        float sum = 0.0f;
        for (uint32_t q = f0; q < f1; q++) sum += a.hub_part[(uint64_t)q * a.cap + c];
        ppr_update(a, row, c, sum, a.cs[c].t, res, dang);
    }
    a.hub_res[(uint64_t)row * kMaxCap + c] = res;
    a.hub_dang[(uint64_t)row * kMaxCap + c] = dang;
}

// One workgroup: per column, residual and dangling sums over the sweep and hub
// partials in a fixed order, then convergence.
constexpr int kRedBlock = 1024;
__global__ __launch_bounds__(kRedBlock) void k_ppr_reduce(PprArgs a)
{
    __shared__ double q_res[kRedBlock / kMaxCap][kMaxCap], q_dang[kRedBlock / kMaxCap][kMaxCap];
    constexpr uint32_t P = kRedBlock / kMaxCap;
    const uint32_t K = a.bs->K;
    if (K == 0) return;
    const uint32_t c = threadIdx.x % kMaxCap, part = threadIdx.x / kMaxCap;
    double r = 0.0, d = 0.0;
    for (uint32_t b = part * a.nparts / P; b < (part + 1) * a.nparts / P; b++) {
        r += a.blk_res[(uint64_t)b * kMaxCap + c];
        d += a.blk_dang[(uint64_t)b * kMaxCap + c];
    }
    for (uint32_t h = part * a.n_hub / P; h < (part + 1) * a.n_hub / P; h++) {
        r += a.hub_res[(uint64_t)h * kMaxCap + c];
        d += a.hub_dang[(uint64_t)h * kMaxCap + c];
    }
    q_res[part][c] = r;
    q_dang[part][c] = d;
    __syncthreads();
    if (threadIdx.x < K) {
        PprCol &col = a.cs[c];
        if (!col.done && !col.stalled) {
            double R = 0.0, D = 0.0;
            for (uint32_t i = 0; i < P; i++) { R += q_res[i][c]; D += q_dang[i][c]; }
            const double prev = col.residual;
            col.iters += 1;
            col.residual = R;
            col.dangling = D;
            col.t = (float)((1.0 - (double)a.damp) + (double)a.damp * D);
            if (col.iters >= col.max_iter)
                col.done = 1;
            else if (col.finish && (R <= col.tol || (col.iters >= 3 && R > 0.97 * prev)))
                col.stalled = 1; // low-precision floor reached: finish in FP32
            else if (R <= col.tol)
                col.done = 1;
        }
    }
    if (threadIdx.x == 0) a.bs->sweeps += 1;
}

// ---- slot management kernels ----
// Column (re)initialisation of the slots admitted in one step: one pass over
// the rows clears X and both contribution buffers for every listed slot.
__global__ void k_ppr_clear_cols(float *X, void *C0, void *C1, uint32_t N, uint32_t cap,
                                 const PprAdmit *adm, uint32_t na, uint32_t fmt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (uint32_t q = 0; q < na; q++) {
        const uint64_t k = (uint64_t)i * cap + adm[q].slot;
        X[k] = 0.0f;
        st(C0, fmt, k, 0.0f);
        st(C1, fmt, k, 0.0f);
    }
}

// One workgroup per admitted slot: x0 = s on the seed rows, contributions
// s / outdeg in the buffer the next sweep reads, column state.
__global__ void k_ppr_seed_cols(float *X, void *Cin, const uint32_t *outdeg, uint32_t cap,
                                uint32_t fmt, const PprAdmit *adm, const uint32_t *sv,
                                const float *sw, PprCol *cs, float damp)
{
    const PprAdmit d = adm[blockIdx.x];
    for (uint32_t q = d.e0 + threadIdx.x; q < d.e1; q += blockDim.x) {
        const uint32_t v = sv[q];
        const uint64_t k = (uint64_t)v * cap + d.slot;
        X[k] = sw[q];
        const uint32_t od = outdeg[v];
        st(Cin, fmt, k, od ? sw[q] / (float)od : 0.0f);
    }
    if (threadIdx.x == 0) {
        double d0 = 0.0;
        for (uint32_t q = d.e0; q < d.e1; q++)
            if (outdeg[sv[q]] == 0) d0 += (double)sw[q];
        PprCol col;
        col.t = (float)((1.0 - (double)damp) + (double)damp * d0);
        col.done = 0;
        col.iters = 0;
        col.max_iter = d.max_iter;
        col.tol = d.tol;
        col.residual = INFINITY;
        col.dangling = d0;
        col.stalled = 0;
        col.finish = fmt != AG_PREC_FP32 ? 1u : 0u;
        cs[d.slot] = col;
    }
}

// Stops columns released while running (a cancelled query).
__global__ void k_ppr_stop_cols(PprCol *cs, const PprAdmit *adm, uint32_t n)
{
    const uint32_t q = threadIdx.x;
    if (q < n) cs[adm[q].slot].done = 1;
}

// Moves stalled columns into an FP32 batch: iterate copied, both FP32
// contribution buffers recomputed from it (x / outdeg), state carried over.
// adm[q].slot = source slot, adm[q].e0 = destination slot.
__global__ void k_ppr_migrate_cols(const float *Xs, uint32_t scap, float *Xd, float *C0, float *C1,
                                   uint32_t dcap, const PprAdmit *adm, uint32_t na,
                                   const uint32_t *outdeg, uint32_t N, const PprCol *cs_src,
                                   PprCol *cs_dst)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) {
        PprCol col = cs_src[adm[i].slot];
        col.stalled = 0;
        col.done = 0;
        col.finish = 0;
        cs_dst[adm[i].e0] = col;
    }
    if (i >= N) return;
    const uint32_t od = outdeg[i];
    for (uint32_t q = 0; q < na; q++) {
        const float x = Xs[(uint64_t)i * scap + adm[q].slot];
        const float c = od ? x / (float)od : 0.0f;
        const uint64_t k = (uint64_t)i * dcap + adm[q].e0;
        Xd[k] = x;
        C0[k] = c;
        C1[k] = c;
    }
}

__global__ void k_ppr_extract(const float *X, const uint32_t *order, uint32_t N, uint32_t cap,
                              uint32_t c, float *out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    out[order[i]] = X[(uint64_t)i * cap + c];
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

template <int CPW, typename CT>
void launch(const PprArgs &a, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL((k_ppr_sweep<CPW, CT>), dim3(grid), dim3(kBlock), 0, s, a);
}

template <typename CT>
void launch_cpw(uint32_t cpw, const PprArgs &a, uint32_t grid, hipStream_t s)
{
    switch (cpw) {
    case 4: launch<4, CT>(a, grid, s); break;
    case 8: launch<8, CT>(a, grid, s); break;
    case 16: launch<16, CT>(a, grid, s); break;
    case 32: launch<32, CT>(a, grid, s); break;
    default: launch<64, CT>(a, grid, s); break;
    }
}

uint32_t cbytes(uint32_t fmt) { return fmt == AG_PREC_FP32 ? 4u : 2u; }

} // namespace

uint64_t PprBatch::bytes_for(uint32_t N, uint32_t cap, uint32_t fmt)
{
    return (uint64_t)N * cap * (4ull + 2ull * cbytes(fmt)) + (uint64_t)N * 4 + (1u << 20);
}

PprBatch::PprBatch(ag_ctx *ctx, ag_graph *g, uint32_t cap, uint32_t cpw, uint32_t fmt, float damp)
    : ctx_(ctx), cap_(cap), cpw_(cpw), fmt_(fmt), damp_(damp)
{
    if (cap == 0 || cap > kMaxCap || cap % 4) fail(AG_ERR_INVALID, "batch capacity must be 4..64, multiple of 4");
    if (cpw != 4 && cpw != 8 && cpw != 16 && cpw != 32 && cpw != 64)
        fail(AG_ERR_INVALID, "cols_per_wave must be 4, 8, 16, 32 or 64");
    if (fmt > AG_PREC_FP16) fail(AG_ERR_INVALID, "unknown precision %u", fmt);
    if (!(damp > 0.0f && damp < 1.0f)) fail(AG_ERR_INVALID, "damping must be in (0, 1)");
    p_ = ensure_plan(ctx, g, nullptr);
    N_ = p_->N;
    if (N_ == 0) fail(AG_ERR_INVALID, "graph has no vertices");
    hipStream_t s = ctx->stream;
    grid_ = std::max<uint32_t>(64, (uint32_t)ctx->num_cus * 8);
    X_.alloc((size_t)N_ * cap * 4);
    C_[0].alloc((size_t)N_ * cap * cbytes(fmt));
    C_[1].alloc((size_t)N_ * cap * cbytes(fmt));
    cols_.alloc(sizeof(PprCol) * kMaxCap);
    bstate_.alloc(sizeof(PprBatchState));
    blk_res_.alloc((size_t)grid_ * kMaxCap * 8);
    blk_dang_.alloc((size_t)grid_ * kMaxCap * 8);
    hub_part_.alloc((size_t)std::max<uint32_t>(p_->hub_chunks, 1) * cap * 4);
    hub_res_.alloc((size_t)std::max<uint32_t>(p_->n_hub, 1) * kMaxCap * 8);
    hub_dang_.alloc((size_t)std::max<uint32_t>(p_->n_hub, 1) * kMaxCap * 8);
    seed_bits_.alloc((size_t)((N_ + 31) / 32) * 4);
    out_.alloc((size_t)N_ * 4);
    desc_.alloc(sizeof(PprAdmit) * kMaxCap);
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
    (void)hipStreamSynchronize(ctx_->stream);
}

uint64_t PprBatch::bytes() const
{
    return X_.bytes + C_[0].bytes + C_[1].bytes + blk_res_.bytes + blk_dang_.bytes +
           hub_part_.bytes + hub_res_.bytes + hub_dang_.bytes + seed_bits_.bytes + out_.bytes +
           seed_v_.bytes + seed_c_.bytes + seed_w_.bytes + desc_.bytes;
}

void PprBatch::set_cpw(uint32_t cpw)
{
    if (cpw != 4 && cpw != 8 && cpw != 16 && cpw != 32 && cpw != 64)
        fail(AG_ERR_INVALID, "bad cols_per_wave");
    cpw_ = cpw;
    if (gexec_) { (void)hipGraphExecDestroy(gexec_); gexec_ = nullptr; }
    if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
}

PprSeedSet PprBatch::make_seeds(const uint32_t *seeds, uint32_t n, double tol, uint32_t max_iter)
{
    if (n == 0 || !seeds) fail(AG_ERR_INVALID, "a query needs at least one seed vertex");
    std::map<uint32_t, uint32_t> cnt;
    for (uint32_t i = 0; i < n; i++) {
        if (seeds[i] >= N_) fail(AG_ERR_INVALID, "seed vertex %u >= num_vertices %u", seeds[i], N_);
        cnt[seeds[i]]++;
    }
    std::vector<uint32_t> ids;
    PprSeedSet q;
    for (auto &kv : cnt) {
        ids.push_back(kv.first);
        q.w.push_back((float)((double)kv.second / (double)n));
    }
    // vertex ids -> labels (device map through the plan's rank array)
    DevBuf din(ids.size() * 4), dout(ids.size() * 4);
    q.labels.resize(ids.size());
    AG_HIP(hipMemcpyAsync(din.p, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, ctx_->stream));
    hipLaunchKernelGGL(k_map_labels, dim3(ceil_div(ids.size(), 256)), dim3(256), 0, ctx_->stream,
                       p_->rank.as<uint32_t>(), din.as<uint32_t>(), (uint32_t)ids.size(),
                       dout.as<uint32_t>());
    AG_HIP(hipMemcpyAsync(q.labels.data(), dout.p, ids.size() * 4, hipMemcpyDeviceToHost,
                          ctx_->stream));
    AG_HIP(hipStreamSynchronize(ctx_->stream));
    q.tol = tol;
    q.max_iter = max_iter ? max_iter : 1000;
    return q;
}

void PprBatch::upload_seed_table()
{
    hipStream_t s = ctx_->stream;
    std::vector<uint32_t> v, c;
    std::vector<float> w;
    for (uint32_t k = 0; k < cap_; k++) {
        if (!used_[k]) continue;
        for (size_t q = 0; q < q_[k].labels.size(); q++) {
            v.push_back(q_[k].labels[q]);
            c.push_back(k);
            w.push_back(q_[k].w[q]);
        }
    }
    const size_t n = v.size();
    const size_t need = std::max<size_t>(n, 1) * 4;
    if (seed_v_.bytes < need) {
        const size_t nb = std::max<size_t>(need * 2, (size_t)cap_ * 64 * 4);
        seed_v_.alloc(nb);
        seed_c_.alloc(nb);
        seed_w_.alloc(nb);
        if (gexec_) { (void)hipGraphExecDestroy(gexec_); gexec_ = nullptr; }
        if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
    }
    if (n) {
        AG_HIP(hipMemcpyAsync(seed_v_.p, v.data(), n * 4, hipMemcpyHostToDevice, s));
        AG_HIP(hipMemcpyAsync(seed_c_.p, c.data(), n * 4, hipMemcpyHostToDevice, s));
        AG_HIP(hipMemcpyAsync(seed_w_.p, w.data(), n * 4, hipMemcpyHostToDevice, s));
    }
    const uint64_t words = (N_ + 31) / 32;
    hipLaunchKernelGGL(k_fill_u32, dim3(ceil_div(words, 256)), dim3(256), 0, s,
                       seed_bits_.as<uint32_t>(), words, 0u);
    if (n)
        hipLaunchKernelGGL(k_set_bits, dim3(ceil_div(n, 256)), dim3(256), 0, s,
                           seed_bits_.as<uint32_t>(), seed_v_.as<uint32_t>(), (uint32_t)n);
    // K (slot high-water mark) and the entry count live on the device so that
    // captured sweeps stay valid.
    PprBatchState bs{hi_, 0u, 0u, (uint32_t)n};
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

// Applies the step's slot changes: stops released columns, rebuilds the seed
// table, and initialises all admitted columns in one pass over the rows.
void PprBatch::flush()
{
    if (!dirty_) return;
    hipStream_t s = ctx_->stream;
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
        // entry ranges of each admitted slot in the table just uploaded
        adm.clear();
        uint32_t off = 0;
        for (uint32_t k = 0; k < cap_; k++) {
            if (!used_[k]) continue;
            const uint32_t n = (uint32_t)q_[k].labels.size();
            if (std::find(pending_.begin(), pending_.end(), k) != pending_.end())
                adm.push_back({k, off, off + n, q_[k].max_iter, q_[k].tol});
            off += n;
        }
        AG_HIP(hipMemcpyAsync(desc_.p, adm.data(), adm.size() * sizeof(PprAdmit),
                              hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_ppr_clear_cols, dim3(ceil_div(N_, 256)), dim3(256), 0, s,
                           X_.as<float>(), C_[0].p, C_[1].p, N_, cap_, desc_.as<PprAdmit>(),
                           (uint32_t)adm.size(), fmt_);
        hipLaunchKernelGGL(k_ppr_seed_cols, dim3((uint32_t)adm.size()), dim3(256), 0, s,
                           X_.as<float>(), C_[cur_].p, p_->outdeg.as<uint32_t>(), cap_, fmt_,
                           desc_.as<PprAdmit>(), seed_v_.as<uint32_t>(), seed_w_.as<float>(),
                           cols_.as<PprCol>(), damp_);
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
    const uint32_t G = 64 / cpw_;
    a.N = N_;
    a.cap = cap_;
    a.n_hub = p_->n_hub;
    a.n_nonlow = p_->n_nonlow;
    a.n_nonzero = p_->n_nonzero;
    a.hub_chunk = p_->hub_chunk;
    a.nb_hub = p_->hub_chunks;
    a.nmid_units = ceil_div(p_->n_nonlow - p_->n_hub, 4);
    a.nlow_units = ceil_div(p_->n_nonzero - p_->n_nonlow, 4 * G);
    a.damp = damp_;
    a.fmt = fmt_;
    a.roff = p_->roff.as<uint32_t>();
    a.cols = p_->cols.as<uint32_t>();
    a.slice_ptr = p_->slice_ptr.as<uint32_t>();
    a.outdeg = p_->outdeg.as<uint32_t>();
    a.chunk_row = p_->chunk_row.as<uint32_t>();
    a.chunk_first = p_->chunk_first.as<uint32_t>();
    a.X = X_.as<float>();
    a.Cin = C_[from0 ? 0 : 1].p;
    a.Cout = C_[from0 ? 1 : 0].p;
    a.seed_bits = seed_bits_.as<uint32_t>();
    a.seed_v = seed_v_.as<uint32_t>();
    a.seed_c = seed_c_.as<uint32_t>();
    a.seed_w = seed_w_.as<float>();
    a.cs = cols_.as<PprCol>();
    a.bs = bstate_.as<PprBatchState>();
    a.blk_res = blk_res_.as<double>();
    a.blk_dang = blk_dang_.as<double>();
    a.hub_part = hub_part_.as<float>();
    a.hub_res = hub_res_.as<double>();
    a.hub_dang = hub_dang_.as<double>();
    a.nparts = grid_;
    hipStream_t s = ctx_->stream;
    if (fmt_ == AG_PREC_FP32) launch_cpw<float>(cpw_, a, grid_, s);
    else if (fmt_ == AG_PREC_BF16) launch_cpw<bf16_t>(cpw_, a, grid_, s);
    else launch_cpw<__half>(cpw_, a, grid_, s);
    if (p_->n_hub) hipLaunchKernelGGL(k_ppr_hub, dim3(p_->n_hub), dim3(kMaxCap), 0, s, a);
    hipLaunchKernelGGL(k_ppr_reduce, dim3(1), dim3(kRedBlock), 0, s, a);
    AG_HIP(hipGetLastError());
}

void PprBatch::sweep(uint32_t n, bool replay)
{
    flush();
    if (used_count_ == 0 || n == 0) return;
    hipStream_t s = ctx_->stream;
    hipEvent_t e0, e1;
    AG_HIP(hipEventCreate(&e0));
    AG_HIP(hipEventCreate(&e1));
    AG_HIP(hipEventRecord(e0, s));
    uint32_t done = 0;
    if (replay) {
        // A captured pair of sweeps (C0 -> C1 -> C0), replayed while the
        // current buffer is C0.
        if (cur_ == 1) { launch_sweep(false); cur_ = 0; done++; }
        if (n - done >= 2 && !gexec_) {
            AG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
            launch_sweep(true);
            launch_sweep(false);
            AG_HIP(hipStreamEndCapture(s, &graph_));
            AG_HIP(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
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
                          ctx_->stream));
    AG_HIP(hipStreamSynchronize(ctx_->stream));
    for (uint32_t k = 0; k < hi_; k++)
        if (used_[k] && (cs[k].done || cs[k].stalled))
            out.push_back({k, tags_[k], cs[k].iters, cs[k].residual, cs[k].stalled != 0});
    return out;
}

void PprBatch::migrate(const std::vector<uint32_t> &slots, PprBatch &dst)
{
    if (slots.empty()) return;
    if (dst.fmt_ != AG_PREC_FP32) fail(AG_ERR_INVALID, "migration target must store FP32");
    if (dst.free_slots() < slots.size()) fail(AG_ERR_ADMISSION, "no free FP32 slots");
    dst.flush();
    hipStream_t s = ctx_->stream;
    std::vector<PprAdmit> mv;
    for (uint32_t c : slots) {
        if (!used(c)) fail(AG_ERR_INVALID, "slot %u not in use", c);
        uint32_t d = 0;
        while (dst.used_[d]) d++;
        dst.used_[d] = 1;
        dst.used_count_++;
        dst.hi_ = std::max(dst.hi_, d + 1);
        dst.q_[d] = q_[c];
        dst.tags_[d] = tags_[c];
        mv.push_back({c, d, 0, 0, 0.0});
    }
    AG_HIP(hipMemcpyAsync(desc_.p, mv.data(), mv.size() * sizeof(PprAdmit), hipMemcpyHostToDevice,
                          s));
    hipLaunchKernelGGL(k_ppr_migrate_cols, dim3(ceil_div(std::max<uint32_t>(N_, kMaxCap), 256)),
                       dim3(256), 0, s, X_.as<float>(), cap_, dst.X_.as<float>(),
                       dst.C_[0].as<float>(), dst.C_[1].as<float>(), dst.cap_,
                       desc_.as<PprAdmit>(), (uint32_t)mv.size(), p_->outdeg.as<uint32_t>(), N_,
                       cols_.as<PprCol>(), dst.cols_.as<PprCol>());
    AG_HIP(hipGetLastError());
    AG_HIP(hipStreamSynchronize(s)); // mv is a temporary
    dst.dirty_ = true;
    for (uint32_t c : slots) {
        used_[c] = 0;
        used_count_--;
        q_[c] = PprSeedSet();
        tags_[c] = ~0ull;
    }
    while (hi_ && !used_[hi_ - 1]) hi_--;
    dirty_ = true; // stalled columns are already inactive on the device
}

void PprBatch::extract(uint32_t slot, float *host_out)
{
    if (!used(slot)) fail(AG_ERR_INVALID, "slot %u not in use", slot);
    if (!host_out) return;
    hipStream_t s = ctx_->stream;
    hipLaunchKernelGGL(k_ppr_extract, dim3(ceil_div(N_, 256)), dim3(256), 0, s, X_.as<float>(),
                       p_->order.as<uint32_t>(), N_, cap_, slot, out_.as<float>());
    AG_HIP(hipGetLastError());
    AG_HIP(hipMemcpyAsync(host_out, out_.p, (size_t)N_ * 4, hipMemcpyDeviceToHost, s));
    AG_HIP(hipStreamSynchronize(s));
}

PprPipeline::PprPipeline(ag_ctx *ctx, ag_graph *g, uint32_t cap, uint32_t cpw, uint32_t fmt,
                         float damp)
{
    first_.reset(new PprBatch(ctx, g, cap, cpw, fmt, damp));
    if (fmt != AG_PREC_FP32) fp32_.reset(new PprBatch(ctx, g, cap, cpw, AG_PREC_FP32, damp));
}

uint64_t PprPipeline::bytes_for(uint32_t N, uint32_t cap, uint32_t fmt)
{
    uint64_t b = PprBatch::bytes_for(N, cap, fmt);
    if (fmt != AG_PREC_FP32) b += PprBatch::bytes_for(N, cap, AG_PREC_FP32);
    return b;
}

void PprPipeline::sweep(uint32_t n, bool replay)
{
    if (first_->active()) first_->sweep(n, replay);
    if (fp32_ && fp32_->active()) fp32_->sweep(n, replay);
}

bool PprPipeline::cancel(uint64_t tag)
{
    for (PprBatch *b : {first_.get(), fp32_.get()}) {
        if (!b) continue;
        for (uint32_t k = 0; k < b->capacity(); k++)
            if (b->used(k) && b->tag(k) == tag) {
                b->release(k);
                return true;
            }
    }
    return false;
}

uint32_t ppr_autotune(ag_ctx *ctx, ag_graph *g, uint32_t fmt)
{
    if (fmt > AG_PREC_FP16) fail(AG_ERR_INVALID, "unknown precision %u", fmt);
    if (g->ppr_cpw[fmt]) return g->ppr_cpw[fmt];
    Plan *p = ensure_plan(ctx, g, nullptr);
    const uint32_t N = p->N;
    PprBatch b(ctx, g, kMaxCap, 32, fmt, 0.85f);
    for (uint32_t q = 0; q < kMaxCap; q++) {
        const uint32_t seed = (uint32_t)(((uint64_t)q * 2654435761u) % N);
        b.admit(b.make_seeds(&seed, 1, 0.0, 1u << 30), q);
    }
    uint32_t best = 32;
    double best_ms = 1e30;
    for (uint32_t cpw : {4u, 8u, 16u, 32u, 64u}) {
        b.set_cpw(cpw);
        b.sweep(1, false);
        b.sweep(2, false);
        if (b.last_sweep_ms() < best_ms) {
            best_ms = b.last_sweep_ms();
            best = cpw;
        }
    }
    g->ppr_cpw[fmt] = best;
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

int ag_ppr_autotune(ag_ctx *ctx, ag_graph *g, uint32_t precision, uint32_t *cpw)
{
    return guard([&] {
        check_ctx(ctx);
        if (!g || !cpw) fail(AG_ERR_INVALID, "null argument");
        *cpw = ppr_autotune(ctx, g, precision);
    });
}

int ag_ppr_batch(ag_ctx *ctx, ag_graph *g, const ag_ppr_query *q, uint32_t nq,
                 const ag_ppr_options *opt, float *x, ag_ppr_stat *stats, double *elapsed_ms)
{
    return guard([&] {
        check_ctx(ctx);
        if (!g || (nq && !q)) fail(AG_ERR_INVALID, "null argument");
        ag_ppr_options o;
        if (opt) o = *opt; else ag_ppr_options_default(&o);
        if (nq == 0) return;
        const uint32_t cpw = o.cols_per_wave ? o.cols_per_wave : ppr_autotune(ctx, g, o.precision);
        const uint32_t cap = std::min<uint32_t>(kMaxCap, ((nq + 3) / 4) * 4);
        PprPipeline pl(ctx, g, cap, cpw, o.precision, o.damp);
        const uint32_t N = g->N;
        std::vector<PprSeedSet> sets(nq);
        for (uint32_t i = 0; i < nq; i++)
            sets[i] = pl.make_seeds(q[i].seeds, q[i].nseeds, q[i].tol, q[i].max_iter);
        AG_HIP(hipStreamSynchronize(ctx->stream));
        const auto t0 = std::chrono::steady_clock::now();
        uint32_t next = 0, finished = 0;
        while (next < nq && pl.free_slots()) { pl.admit(sets[next], next); next++; }
        while (finished < nq) {
            pl.sweep(4, o.use_graph_replay != 0);
            auto fin = pl.retire([&](uint64_t tag) { return x ? x + (size_t)tag * N : nullptr; });
            for (auto &f : fin) {
                if (stats) {
                    stats[f.tag].iterations = f.iters;
                    stats[f.tag].residual = f.residual;
                    stats[f.tag].status = AG_OK;
                    stats[f.tag].latency_ms = std::chrono::duration<double, std::milli>(
                        std::chrono::steady_clock::now() - t0).count();
                }
                finished++;
            }
            while (next < nq && pl.free_slots()) { pl.admit(sets[next], next); next++; }
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
    return guard([&] {
        check_ctx(ctx);
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
        o.flags = precision != AG_PREC_FP32 ? AG_PR_FLAG_FINISH_FP32 : 0u;
        ag_pr_result r;
        const auto t0 = std::chrono::steady_clock::now();
        run_pagerank_seeded(ctx, g, o, st, out.as<float>(), &r);
        if (x) AG_HIP(hipMemcpyAsync(x, out.p, (size_t)N * 4, hipMemcpyDeviceToHost, s));
        AG_HIP(hipStreamSynchronize(s));
        if (stat) {
            stat->iterations = r.iterations;
            stat->status = AG_OK;
            stat->residual = r.error_total;
            stat->latency_ms = std::chrono::duration<double, std::milli>(
                std::chrono::steady_clock::now() - t0).count();
        }
    });
}

} // extern "C"
