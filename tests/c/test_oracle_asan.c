/* The oracle built with AddressSanitizer + UndefinedBehaviorSanitizer (CPU
 * only): every oracle function runs on small seeded graphs, edge cases
 * included (no edges, one vertex, self-loops, a vertex holding every edge),
 * and a few invariants are checked so that the run exercises real paths.
 * Exits non-zero on the first failed invariant; the sanitizers abort on the
 * first memory or UB error. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int failures = 0;
#define EXPECT(c, what) do { if (!(c)) { printf("FAIL %s\n", what); failures++; } } while (0)

static uint32_t lcg(uint32_t *s) { *s = *s * 1664525u + 1013904223u; return *s >> 8; }

static void run_graph(uint32_t N, uint32_t E, const uint32_t *src, const uint32_t *dst)
{
    uint32_t *fs = calloc(E + 1, 4), *fd = calloc(E + 1, 4), *fo = calloc(N + 1, 4),
             *fi = calloc(N + 1, 4), *fx = calloc(N + 1, 4), *is = calloc(E + 1, 4),
             *id = calloc(E + 1, 4), *io = calloc(N + 1, 4), *ii = calloc(N + 1, 4),
             *ix = calloc(N + 1, 4), *lab = calloc(N + 1, 4), *act = calloc(64, 4),
             *work = calloc(64, 4);
    float *r32 = calloc(N + 1, 4), *x = calloc(N + 1, 4), *y = calloc(N + 1, 4);
    double *r64 = calloc(N + 1, 8), *err = calloc(64, 8), et = 0, ps = 0, res = 0;
    for (uint32_t mode = 0; mode < 4; mode++) {
        oracle_build_csr(N, E, src, dst, mode, fs, fd, fo, fi, fx, is, id, io, ii, ix);
        uint64_t so = 0, si = 0;
        for (uint32_t v = 0; v < N; v++) { so += fo[v]; si += io[v]; }
        EXPECT(so == E && si == E, "degrees sum to E");
    }
    oracle_pagerank_pull_f32(N, fo, io, ix, id, 20, 0.0, r32, &et, &ps, act, err);
    oracle_pagerank_pull_f32(N, fo, io, ix, id, 20, 1e-4, r32, &et, &ps, act, err);
    oracle_pagerank_pull_f64(N, fo, io, ix, id, 20, 0.0, r64, &et);
    if (N) {
        oracle_pagerank_standard_f64(N, fo, io, ix, id, 0.85, 50, 1e-9, r64, &res);
        double s = 0;
        for (uint32_t v = 0; v < N; v++) s += r64[v];
        EXPECT(fabs(s - 1.0) < 1e-9, "STANDARD mass is 1");
        uint32_t seeds[3] = {0, N / 2, N - 1};
        oracle_ppr_f64(N, fo, io, ix, id, 0.85, seeds, 3, 1e-8, 200, r64, &res);
        s = 0;
        for (uint32_t v = 0; v < N; v++) s += r64[v];
        EXPECT(fabs(s - 1.0) < 1e-6, "PPR mass is 1");
    }
    oracle_pagerank_dd_pull_f32(N, fo, fx, fd, io, ix, id, 20, 1e-4, r32, &et, &ps, act, err,
                                work);
    oracle_spmv_input(N, fo, x);
    oracle_spmv_pull_f32(N, io, ix, id, x, y);
    (void)oracle_spmv_sum(N, y);
    for (uint32_t mode = 0; mode <= 10; mode++) {
        EXPECT(oracle_reorder_labels(N, E, src, dst, mode, 7, lab) == 0, "reorder mode");
        for (uint32_t v = 0; v < N; v++) EXPECT(lab[v] < N, "label in range");
    }
    EXPECT(oracle_reorder_labels(N, E, src, dst, 99, 0, lab) == -1, "unknown mode refused");
    oracle_csr_build_omp(N, E, src, dst, fs, fd, fo, fx, is, id, io, ix, 2);
    uint32_t done = 0;
    oracle_pagerank_pull_omp(N, fo, io, ix, id, 20, 0.0, r32, &et, &done, 2);
    free(fs); free(fd); free(fo); free(fi); free(fx); free(is); free(id); free(io); free(ii);
    free(ix); free(lab); free(act); free(work); free(r32); free(x); free(y); free(r64); free(err);
}

int main(void)
{
    uint32_t seed = 12345;
    const uint32_t sizes[][2] = {{1, 0}, {1, 5}, {2, 1}, {7, 40}, {64, 1000}, {1000, 20000}};
    for (size_t k = 0; k < sizeof(sizes) / sizeof(sizes[0]); k++) {
        uint32_t N = sizes[k][0], E = sizes[k][1];
        uint32_t *s = calloc(E + 1, 4), *d = calloc(E + 1, 4);
        for (uint32_t e = 0; e < E; e++) { s[e] = lcg(&seed) % N; d[e] = lcg(&seed) % N; }
        EXPECT(E == 0 || oracle_num_vertices(E, s, d) <= N, "num_vertices");
        run_graph(N, E, s, d);
        /* one vertex holding every edge */
        for (uint32_t e = 0; e < E; e++) d[e] = 0;
        run_graph(N, E, s, d);
        free(s); free(d);
    }
    printf("%s oracle under ASan/UBSan\n", failures ? "FAIL" : "PASS");
    return failures ? 1 : 0;
}
