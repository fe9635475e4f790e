/*
 * CPU oracle (serial restatement + OpenMP baselines) for the GraphCSR builder
 * and pageRankPullGraphCSR.  TEST INFRASTRUCTURE ONLY: see oracle.h.
 *
 * Compiled with -ffp-contract=off so that `base + damp * next` is a separate
 * multiply and add, as written in pageRank.c:1058.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static void *xcalloc(size_t n, size_t sz)
{
    void *p = calloc(n ? n : 1, sz ? sz : 1);
    if (!p) abort();
    return p;
}

uint32_t oracle_num_vertices(uint32_t E, const uint32_t *src, const uint32_t *dst)
{
    uint32_t m = 0;
    if (E == 0) return 0;
    for (uint32_t e = 0; e < E; e++) {
        if (src[e] > m) m = src[e];
        if (dst[e] > m) m = dst[e];
    }
    return m + 1;
}

/* Textbook stable counting sort: histogram, exclusive scan, scatter in input
 * order.  This is sortRunAlgorithms(el, 2) (count sort by src; modes listed in
 * README.md:434-439, call site SSSP.c:548-553). */
void oracle_count_sort(uint32_t N, uint32_t E, const uint32_t *key, const uint32_t *val,
                       uint32_t *key_out, uint32_t *val_out)
{
    uint64_t *pos = (uint64_t *)xcalloc((size_t)N + 1, sizeof(uint64_t));
    for (uint32_t e = 0; e < E; e++) pos[key[e] + 1]++;
    for (uint32_t v = 0; v < N; v++) pos[v + 1] += pos[v];
    for (uint32_t e = 0; e < E; e++) {
        uint64_t p = pos[key[e]]++;
        key_out[p] = key[e];
        val_out[p] = val[e];
    }
    free(pos);
}

/* graphCSRAssignEdgeList's vertex mapping (OpenGraph, absent; fields read at
 * capienv.c:409-414): degree = segment length, edges_idx = segment start, and
 * 0 for an empty segment (convention inferred, inert for every consumer since
 * they read [edges_idx, edges_idx + degree)). */
void oracle_vertex_map(uint32_t N, uint32_t E, const uint32_t *sorted_key,
                       uint32_t *degree, uint32_t *edges_idx)
{
    memset(degree, 0, (size_t)N * sizeof(uint32_t));
    memset(edges_idx, 0, (size_t)N * sizeof(uint32_t));
    for (uint32_t e = 0; e < E; e++) {
        uint32_t v = sorted_key[e];
        if (degree[v] == 0) edges_idx[v] = e;
        degree[v]++;
    }
}

void oracle_build_csr(uint32_t N, uint32_t E, const uint32_t *src, const uint32_t *dst,
                      uint32_t sort_mode,
                      uint32_t *fwd_src, uint32_t *fwd_dst, uint32_t *fwd_outdeg,
                      uint32_t *fwd_indeg, uint32_t *fwd_idx,
                      uint32_t *inv_src, uint32_t *inv_dst, uint32_t *inv_outdeg,
                      uint32_t *inv_indeg, uint32_t *inv_idx)
{
    uint32_t mode = sort_mode & 0xff;
    int inverse_from_input = (sort_mode & 0x100) != 0;
    int by_src_dst = (mode == 1 || mode == 3);
    uint32_t *ta = (uint32_t *)xcalloc(E, sizeof(uint32_t));
    uint32_t *tb = (uint32_t *)xcalloc(E, sizeof(uint32_t));

    /* forward: stable by src, or lexicographic (src, dst) via two stable passes */
    if (by_src_dst) {
        oracle_count_sort(N, E, dst, src, tb, ta);      /* tb = dst sorted, ta = src */
        oracle_count_sort(N, E, ta, tb, fwd_src, fwd_dst);
    } else {
        oracle_count_sort(N, E, src, dst, fwd_src, fwd_dst);
    }

    /* inverse: swap endpoints (SSSP.c:518-520) of the forward-sorted list (or of
     * the input list), then sort by the new src with the same mode. */
    const uint32_t *isrc = inverse_from_input ? dst : fwd_dst;
    const uint32_t *idst = inverse_from_input ? src : fwd_src;
    if (by_src_dst) {
        oracle_count_sort(N, E, idst, isrc, tb, ta);
        oracle_count_sort(N, E, ta, tb, inv_src, inv_dst);
    } else {
        oracle_count_sort(N, E, isrc, idst, inv_src, inv_dst);
    }

    oracle_vertex_map(N, E, fwd_src, fwd_outdeg, fwd_idx);
    oracle_vertex_map(N, E, inv_src, inv_outdeg, inv_idx);
    memcpy(fwd_indeg, inv_outdeg, (size_t)N * sizeof(uint32_t));
    memcpy(inv_indeg, fwd_outdeg, (size_t)N * sizeof(uint32_t));
    free(ta);
    free(tb);
}

/* pageRankPullGraphCSR, FP32 (pageRank.c:942-1108), with the OpenMP gather
 * form of pageRank.c:2890-2907 run serially:
 *   init        pr[v] = base_pr (pageRank.c:57-58,71), next[v] = 0 (:1014-1019)
 *   contrib     outdeg ? pr/outdeg : 0 with the FORWARD out-degree (:1024-1033)
 *   gather      sequential FP32 += over the inverse edge range (:2900-2907)
 *   update      new = base + damp*next; err = fabs(new-old) (float diff,
 *               promoted); error_total += err/N; active += err >= epsilon
 *               (:1055-1070); break if active == 0 without ++iterations (:1073)
 *   normalise   pr /= N; sum += pr in double (:1079-1085)                   */
uint32_t oracle_pagerank_pull_f32(uint32_t N, const uint32_t *fwd_outdeg,
                                  const uint32_t *inv_deg, const uint32_t *inv_idx,
                                  const uint32_t *inv_dest, uint32_t iterations,
                                  double epsilon, float *pr, double *error_total,
                                  double *pr_sum, uint32_t *active_iter, double *err_iter)
{
    return oracle_pagerank_pull_f32_from(N, fwd_outdeg, inv_deg, inv_idx, inv_dest, iterations,
                                         epsilon, NULL, pr, error_total, pr_sum, active_iter,
                                         err_iter);
}

uint32_t oracle_pagerank_pull_f32_from(uint32_t N, const uint32_t *fwd_outdeg,
                                       const uint32_t *inv_deg, const uint32_t *inv_idx,
                                       const uint32_t *inv_dest, uint32_t iterations,
                                       double epsilon, const float *init, float *pr,
                                       double *error_total, double *pr_sum,
                                       uint32_t *active_iter, double *err_iter)
{
    const float damp = 0.85f;
    const float base = 1.0f - damp;
    float *contrib = (float *)xcalloc(N, sizeof(float));
    float *next = (float *)xcalloc(N, sizeof(float));
    double err_tot = 0.0;
    uint32_t it;

    /* pageRank.c starts at base; a warm start is the caller's pr / N, times N */
    for (uint32_t v = 0; v < N; v++) { pr[v] = init ? init[v] * (float)N : base; next[v] = 0.0f; }

    for (it = 0; it < iterations; it++) {
        uint32_t active = 0;
        err_tot = 0.0;
        for (uint32_t v = 0; v < N; v++)
            contrib[v] = fwd_outdeg[v] ? pr[v] / (float)fwd_outdeg[v] : 0.0f;
        for (uint32_t v = 0; v < N; v++) {
            float s = 0.0f;
            uint32_t j0 = inv_idx[v], j1 = inv_idx[v] + inv_deg[v];
            for (uint32_t j = j0; j < j1; j++) s += contrib[inv_dest[j]];
            next[v] = s;
        }
        for (uint32_t v = 0; v < N; v++) {
            float old = pr[v];
            float nw = base + damp * next[v];
            pr[v] = nw;
            next[v] = 0.0f;
            double e = fabs(nw - old);
            err_tot += e / N;
            if (e >= epsilon) active++;
        }
        if (active_iter) active_iter[it] = active;
        if (err_iter) err_iter[it] = err_tot;
        if (active == 0) break;
    }

    double sum = 0.0;
    for (uint32_t v = 0; v < N; v++) {
        pr[v] = pr[v] / N;
        sum += pr[v];
    }
    if (error_total) *error_total = err_tot;
    if (pr_sum) *pr_sum = sum;
    free(contrib);
    free(next);
    return it;
}

uint32_t oracle_pagerank_pull_f64(uint32_t N, const uint32_t *fwd_outdeg,
                                  const uint32_t *inv_deg, const uint32_t *inv_idx,
                                  const uint32_t *inv_dest, uint32_t iterations,
                                  double epsilon, double *pr, double *error_total)
{
    const double damp = (double)0.85f;
    const double base = (double)(1.0f - 0.85f);
    double *contrib = (double *)xcalloc(N, sizeof(double));
    double err_tot = 0.0;
    uint32_t it;

    for (uint32_t v = 0; v < N; v++) pr[v] = base;
    for (it = 0; it < iterations; it++) {
        uint32_t active = 0;
        err_tot = 0.0;
        for (uint32_t v = 0; v < N; v++)
            contrib[v] = fwd_outdeg[v] ? pr[v] / (double)fwd_outdeg[v] : 0.0;
        for (uint32_t v = 0; v < N; v++) {
            double s = 0.0;
            uint32_t j0 = inv_idx[v], j1 = inv_idx[v] + inv_deg[v];
            for (uint32_t j = j0; j < j1; j++) s += contrib[inv_dest[j]];
            double nw = base + damp * s;
            double e = fabs(nw - pr[v]);
            pr[v] = nw;
            err_tot += e / N;
            if (e >= epsilon) active++;
        }
        if (active == 0) break;
    }
    for (uint32_t v = 0; v < N; v++) pr[v] /= N;
    if (error_total) *error_total = err_tot;
    free(contrib);
    return it;
}

uint32_t oracle_pagerank_standard_f64(uint32_t N, const uint32_t *fwd_outdeg,
                                      const uint32_t *inv_deg, const uint32_t *inv_idx,
                                      const uint32_t *inv_dest, double damp,
                                      uint32_t iterations, double tol, double *x,
                                      double *residual)
{
    return oracle_pagerank_standard_f64_from(N, fwd_outdeg, inv_deg, inv_idx, inv_dest, damp,
                                             iterations, tol, NULL, x, residual);
}

uint32_t oracle_pagerank_standard_f64_from(uint32_t N, const uint32_t *fwd_outdeg,
                                           const uint32_t *inv_deg, const uint32_t *inv_idx,
                                           const uint32_t *inv_dest, double damp,
                                           uint32_t iterations, double tol, const double *init,
                                           double *x, double *residual)
{
    double *contrib = (double *)xcalloc(N, sizeof(double));
    double *nx = (double *)xcalloc(N, sizeof(double));
    double res = 0.0;
    uint32_t it;
    if (N == 0) { if (residual) *residual = 0.0; free(contrib); free(nx); return 0; }
    for (uint32_t v = 0; v < N; v++) x[v] = init ? init[v] : 1.0 / N;
    for (it = 0; it < iterations; it++) {
        double dangling = 0.0;
        for (uint32_t v = 0; v < N; v++) {
            if (fwd_outdeg[v]) contrib[v] = x[v] / (double)fwd_outdeg[v];
            else { contrib[v] = 0.0; dangling += x[v]; }
        }
        double tele = (1.0 - damp) / N + damp * dangling / N;
        res = 0.0;
        for (uint32_t v = 0; v < N; v++) {
            double s = 0.0;
            uint32_t j0 = inv_idx[v], j1 = inv_idx[v] + inv_deg[v];
            for (uint32_t j = j0; j < j1; j++) s += contrib[inv_dest[j]];
            nx[v] = tele + damp * s;
            res += fabs(nx[v] - x[v]);
        }
        memcpy(x, nx, (size_t)N * sizeof(double));
        if (tol > 0.0 && res <= tol) { it++; break; }
    }
    if (residual) *residual = res;
    free(contrib);
    free(nx);
    return it;
}

uint32_t oracle_ppr_f64(uint32_t N, const uint32_t *fwd_outdeg, const uint32_t *inv_deg,
                        const uint32_t *inv_idx, const uint32_t *inv_dest, double damp,
                        const uint32_t *seeds, uint32_t nseeds, double tol,
                        uint32_t max_iter, double *x, double *residual)
{
    double *s = (double *)xcalloc(N, sizeof(double));
    double *contrib = (double *)xcalloc(N, sizeof(double));
    double *nx = (double *)xcalloc(N, sizeof(double));
    double res = 0.0;
    uint32_t it;
    for (uint32_t k = 0; k < nseeds; k++) s[seeds[k]] += 1.0 / nseeds;
    memcpy(x, s, (size_t)N * sizeof(double));
    for (it = 0; it < max_iter; it++) {
        double dangling = 0.0;
        for (uint32_t v = 0; v < N; v++) {
            if (fwd_outdeg[v]) contrib[v] = x[v] / (double)fwd_outdeg[v];
            else { contrib[v] = 0.0; dangling += x[v]; }
        }
        res = 0.0;
        for (uint32_t v = 0; v < N; v++) {
            double acc = 0.0;
            uint32_t j0 = inv_idx[v], j1 = inv_idx[v] + inv_deg[v];
            for (uint32_t j = j0; j < j1; j++) acc += contrib[inv_dest[j]];
            nx[v] = (1.0 - damp) * s[v] + damp * (acc + dangling * s[v]);
            res += fabs(nx[v] - x[v]);
        }
        memcpy(x, nx, (size_t)N * sizeof(double));
        if (res <= tol) { it++; break; }
    }
    if (residual) *residual = res;
    free(s);
    free(contrib);
    free(nx);
    return it;
}

/* Data-driven pull, after pageRank.c:2810-2968 (DIRECTED: gather over the
 * inverse CSR, marks over the forward CSR).  The contribution vector is taken
 * from the ranks as they stand at the start of the sweep (:2876-2883), so the
 * in-place rank writes of one sweep are seen only by the next (Jacobi). */
uint32_t oracle_pagerank_dd_pull_f32(uint32_t N, const uint32_t *fwd_outdeg,
                                     const uint32_t *fwd_idx, const uint32_t *fwd_dest,
                                     const uint32_t *inv_deg, const uint32_t *inv_idx,
                                     const uint32_t *inv_dest, uint32_t iterations,
                                     double epsilon, float *pr, double *error_total,
                                     double *pr_sum, uint32_t *active_iter, double *err_iter,
                                     uint32_t *work_iter)
{
    return oracle_pagerank_dd_pull_f32_from(N, fwd_outdeg, fwd_idx, fwd_dest, inv_deg, inv_idx,
                                            inv_dest, iterations, epsilon, NULL, pr, error_total,
                                            pr_sum, active_iter, err_iter, work_iter);
}

uint32_t oracle_pagerank_dd_pull_f32_from(uint32_t N, const uint32_t *fwd_outdeg,
                                          const uint32_t *fwd_idx, const uint32_t *fwd_dest,
                                          const uint32_t *inv_deg, const uint32_t *inv_idx,
                                          const uint32_t *inv_dest, uint32_t iterations,
                                          double epsilon, const float *init, float *pr,
                                          double *error_total, double *pr_sum,
                                          uint32_t *active_iter, double *err_iter,
                                          uint32_t *work_iter)
{
    const float damp = 0.85f;
    const float base = 1.0f - damp;
    float *contrib = (float *)xcalloc(N ? N : 1, sizeof(float));
    uint8_t *cur = (uint8_t *)xcalloc(N ? N : 1, 1);
    uint8_t *nxt = (uint8_t *)xcalloc(N ? N : 1, 1);
    double err_tot = 0.0;
    uint32_t it;

    for (uint32_t v = 0; v < N; v++) { pr[v] = init ? init[v] * (float)N : base; cur[v] = 1; } /* :2862-2870 */

    for (it = 0; it < iterations; it++) {
        uint32_t active = 0, work = 0;
        err_tot = 0.0;
        for (uint32_t v = 0; v < N; v++)
            contrib[v] = fwd_outdeg[v] ? pr[v] / (float)fwd_outdeg[v] : 0.0f;
        for (uint32_t v = 0; v < N; v++) {
            if (!cur[v]) continue;
            work++;
            float s = 0.0f;
            uint32_t j0 = inv_idx[v], j1 = inv_idx[v] + inv_deg[v];
            for (uint32_t j = j0; j < j1; j++) s += contrib[inv_dest[j]];
            float old = pr[v];
            float nw = base + damp * s;
            double e = fabs(nw - old);
            err_tot += e / N;
            if (e >= epsilon) {                                     /* :2910-2926 */
                pr[v] = nw;
                uint32_t k0 = fwd_idx[v], k1 = fwd_idx[v] + fwd_outdeg[v];
                for (uint32_t k = k0; k < k1; k++) nxt[fwd_dest[k]] = 1;
                active++;
            }
        }
        uint8_t *t = cur; cur = nxt; nxt = t;                        /* :2930-2931 */
        memset(nxt, 0, N);
        if (active_iter) active_iter[it] = active;
        if (err_iter) err_iter[it] = err_tot;
        if (work_iter) work_iter[it] = work;
        if (active == 0) break;
    }

    double sum = 0.0;
    for (uint32_t v = 0; v < N; v++) {
        pr[v] = pr[v] / N;
        sum += pr[v];
    }
    if (error_total) *error_total = err_tot;
    if (pr_sum) *pr_sum = sum;
    free(contrib);
    free(cur);
    free(nxt);
    return it;
}

void oracle_spmv_input(uint32_t N, const uint32_t *fwd_outdeg, float *x)
{
    for (uint32_t v = 0; v < N; v++)
        x[v] = fwd_outdeg[v] ? (1.0f / fwd_outdeg[v]) : 0.001f;
}

void oracle_spmv_pull_f32(uint32_t N, const uint32_t *inv_deg, const uint32_t *inv_idx,
                          const uint32_t *inv_dest, const float *x, float *y)
{
    for (uint32_t v = 0; v < N; v++) {
        float s = 0.0f;
        uint32_t j0 = inv_idx[v], j1 = inv_idx[v] + inv_deg[v];
        for (uint32_t j = j0; j < j1; j++) s += x[inv_dest[j]];
        y[v] = s;
    }
}

double oracle_spmv_sum(uint32_t N, const float *y)
{
    double sum = 0.0;
    for (uint32_t v = 0; v < N; v++) sum += ((int)(y[v] * 100 + .5) / 100.0);
    return sum;
}

/* Relabelling: a key per vertex, then ids ordered by (key, id).  The keys
 * are encoded so that ascending order is the mode's order. */
typedef struct { uint64_t key; uint32_t id; } reorder_item;

static int reorder_cmp(const void *a, const void *b)
{
    const reorder_item *x = (const reorder_item *)a, *y = (const reorder_item *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id ? 1 : 0);
}

static uint32_t lowbias32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

int oracle_reorder_labels(uint32_t N, uint32_t E, const uint32_t *src, const uint32_t *dst,
                          uint32_t mode, uint32_t seed, uint32_t *labels)
{
    if (mode > 10) return -1;
    uint32_t *deg = (uint32_t *)xcalloc(N ? N : 1, sizeof(uint32_t));
    const int use_out = mode == 1 || mode == 3 || mode == 4 || mode == 6 || mode == 8;
    const int use_in = mode == 2 || mode == 3 || mode == 5 || mode == 7 || mode == 9;
    for (uint32_t e = 0; e < E; e++) {
        if (use_out) deg[src[e]]++;
        if (use_in) deg[dst[e]]++;
    }
    const uint64_t avg = N ? E / N : 0;
    reorder_item *it = (reorder_item *)xcalloc(N ? N : 1, sizeof(reorder_item));
    for (uint32_t v = 0; v < N; v++) {
        const uint64_t d = deg[v];
        uint64_t k = 0;
        switch (mode) {
        case 0: k = 0; break;
        case 1: case 2: case 3: k = UINT64_MAX - d; break;
        case 4: case 5:
            k = d >= 32 * avg ? 0 : d >= 16 * avg ? 1 : d >= 8 * avg ? 2 : d >= 4 * avg ? 3
              : d >= 2 * avg ? 4 : d >= avg ? 5 : 2 * d >= avg ? 6 : 7;
            break;
        case 6: case 7: k = d > avg ? UINT64_MAX - 1 - d : UINT64_MAX; break;
        case 8: case 9: k = d > avg ? 0 : 1; break;
        case 10: k = lowbias32(v * 0x9E3779B9u ^ seed); break;
        }
        it[v].key = k;
        it[v].id = v;
    }
    qsort(it, N, sizeof(reorder_item), reorder_cmp);
    for (uint32_t i = 0; i < N; i++) labels[it[i].id] = i;
    free(it);
    free(deg);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* OpenMP baselines                                                          */
/* ------------------------------------------------------------------------- */

int oracle_omp_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static double now_s(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* Parallel stable counting sort: per-thread histograms over contiguous input
 * chunks, offsets ordered (vertex, thread), scatter in chunk order. */
static void omp_count_sort(uint32_t N, uint32_t E, const uint32_t *key, const uint32_t *val,
                           uint32_t *kout, uint32_t *vout, int T)
{
    uint32_t *cnt = (uint32_t *)xcalloc((size_t)T * N, sizeof(uint32_t));
    uint64_t *off = (uint64_t *)xcalloc((size_t)N + 1, sizeof(uint64_t));
    #pragma omp parallel num_threads(T)
    {
        int t = 0, nt = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        nt = omp_get_num_threads();
#endif
        uint64_t lo = (uint64_t)E * t / nt, hi = (uint64_t)E * (t + 1) / nt;
        uint32_t *c = cnt + (size_t)t * N;
        for (uint64_t e = lo; e < hi; e++) c[key[e]]++;
        #pragma omp barrier
        #pragma omp for schedule(static)
        for (uint32_t v = 0; v < N; v++) {
            uint32_t run = 0;
            for (int u = 0; u < nt; u++) {
                uint32_t x = cnt[(size_t)u * N + v];
                cnt[(size_t)u * N + v] = run;
                run += x;
            }
            off[v + 1] = run;
        }
        #pragma omp single
        for (uint32_t v = 0; v < N; v++) off[v + 1] += off[v];
        for (uint64_t e = lo; e < hi; e++) {
            uint32_t k = key[e];
            uint64_t p = off[k] + c[k]++;
            kout[p] = k;
            vout[p] = val[e];
        }
    }
    free(cnt);
    free(off);
}

static void omp_vertex_map(uint32_t N, uint32_t E, const uint32_t *sk, uint32_t *deg,
                           uint32_t *idx, int T)
{
    #pragma omp parallel for num_threads(T) schedule(static)
    for (uint32_t v = 0; v < N; v++) { deg[v] = 0; idx[v] = 0; }
    #pragma omp parallel for num_threads(T) schedule(static)
    for (uint32_t e = 0; e < E; e++) {
        if (e == 0 || sk[e] != sk[e - 1]) {
            uint32_t v = sk[e];
            uint32_t end = e + 1;
            while (end < E && sk[end] == v) end++;
            idx[v] = e;
            deg[v] = end - e;
        }
    }
}

double oracle_csr_build_omp(uint32_t N, uint32_t E, const uint32_t *src, const uint32_t *dst,
                            uint32_t *fwd_src, uint32_t *fwd_dst, uint32_t *fwd_outdeg,
                            uint32_t *fwd_idx, uint32_t *inv_src, uint32_t *inv_dst,
                            uint32_t *inv_outdeg, uint32_t *inv_idx, int threads)
{
    int T = threads > 0 ? threads : oracle_omp_max_threads();
    double t0 = now_s();
    omp_count_sort(N, E, src, dst, fwd_src, fwd_dst, T);
    omp_count_sort(N, E, fwd_dst, fwd_src, inv_src, inv_dst, T);
    omp_vertex_map(N, E, fwd_src, fwd_outdeg, fwd_idx, T);
    omp_vertex_map(N, E, inv_src, inv_outdeg, inv_idx, T);
    return now_s() - t0;
}

/* pageRankPullGraphCSR with the reference's OpenMP loop structure:
 * contrib `parallel for` (pageRank.c:1024), gather `schedule(dynamic,1024)`
 * (pageRank.c:2890), update `reduction(+:error_total,activeVertices)`
 * (pageRank.c:1055), normalise `reduction(+:sum)` (pageRank.c:1079). */
double oracle_pagerank_pull_omp(uint32_t N, const uint32_t *fwd_outdeg,
                                const uint32_t *inv_deg, const uint32_t *inv_idx,
                                const uint32_t *inv_dest, uint32_t iterations,
                                double epsilon, float *pr, double *error_total,
                                uint32_t *iters_done, int threads)
{
    int T = threads > 0 ? threads : oracle_omp_max_threads();
    const float damp = 0.85f;
    const float base = 1.0f - damp;
    float *contrib = (float *)xcalloc(N, sizeof(float));
    float *next = (float *)xcalloc(N, sizeof(float));
    double err_tot = 0.0;
    uint32_t it;
    #pragma omp parallel for num_threads(T)
    for (uint32_t v = 0; v < N; v++) { pr[v] = base; next[v] = 0.0f; }

    double t0 = now_s();
    for (it = 0; it < iterations; it++) {
        uint32_t active = 0;
        double e_sum = 0.0;
        #pragma omp parallel for num_threads(T)
        for (uint32_t v = 0; v < N; v++)
            contrib[v] = fwd_outdeg[v] ? pr[v] / (float)fwd_outdeg[v] : 0.0f;
        #pragma omp parallel for num_threads(T) schedule(dynamic, 1024)
        for (uint32_t v = 0; v < N; v++) {
            float s = 0.0f;
            uint32_t j0 = inv_idx[v], j1 = inv_idx[v] + inv_deg[v];
            for (uint32_t j = j0; j < j1; j++) s += contrib[inv_dest[j]];
            next[v] = s;
        }
        #pragma omp parallel for num_threads(T) reduction(+ : e_sum, active)
        for (uint32_t v = 0; v < N; v++) {
            float old = pr[v];
            float nw = base + damp * next[v];
            pr[v] = nw;
            next[v] = 0.0f;
            double e = fabs(nw - old);
            e_sum += e / N;
            if (e >= epsilon) active++;
        }
        err_tot = e_sum;
        if (active == 0) break;
    }
    double secs = now_s() - t0;
    #pragma omp parallel for num_threads(T)
    for (uint32_t v = 0; v < N; v++) pr[v] = pr[v] / N;
    if (error_total) *error_total = err_tot;
    if (iters_done) *iters_done = it;
    free(contrib);
    free(next);
    return secs;
}
