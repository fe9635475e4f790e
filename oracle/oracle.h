/*
 * CPU oracle for the CSR builder and pull PageRank hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * never by the product library (accelgraph_express_amd), which must run on the
 * GPU or fail.
 *
 * Every function restates the reference's algorithm serially (or, for the
 * *_omp baselines, with the reference's OpenMP loop structure).  The reference
 * C files are cited per function.  The OpenGraph submodule that holds the
 * reference's CSR builder (graphCSR.c, edgeList.c, sortRun.c) is absent from
 * the reference tree, so the builder restatement follows the textbook stable
 * counting sort and the in-tree call sites (SSSP.c:443-565).
 *
 * Pinning (which check fixes which function) is listed in tests/test_oracle_*.py.
 */
#ifndef AG_ORACLE_H
#define AG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* num_vertices = max(src, dst) + 1 (0 for an empty list). */
uint32_t oracle_num_vertices(uint32_t E, const uint32_t *src, const uint32_t *dst);

/* Stable counting sort of (key, val) pairs by key (keys < N). */
void oracle_count_sort(uint32_t N, uint32_t E, const uint32_t *key, const uint32_t *val,
                       uint32_t *key_out, uint32_t *val_out);

/* Vertex mapping of a sorted edge list: out_degree[v] = #edges with key v,
 * edges_idx[v] = first index of v's segment, 0 when the degree is 0. */
void oracle_vertex_map(uint32_t N, uint32_t E, const uint32_t *sorted_key,
                       uint32_t *degree, uint32_t *edges_idx);

/* Full GraphCSR build: forward (stable by src; sort_mode 1/3 = by (src,dst))
 * and inverse (from the forward-sorted list with endpoints swapped, stable by
 * the new src).  in_degree of each side is the other side's out_degree. */
void oracle_build_csr(uint32_t N, uint32_t E, const uint32_t *src, const uint32_t *dst,
                      uint32_t sort_mode,
                      uint32_t *fwd_src, uint32_t *fwd_dst, uint32_t *fwd_outdeg,
                      uint32_t *fwd_indeg, uint32_t *fwd_idx,
                      uint32_t *inv_src, uint32_t *inv_dst, uint32_t *inv_outdeg,
                      uint32_t *inv_indeg, uint32_t *inv_idx);

/* Reference pull PageRank (pageRank.c:942-1108), FP32, serial.  Returns the
 * reported iteration count.  ranks[] receives the normalised ranks (pr/N).
 * active_iter/err_iter (may be NULL) receive per-iteration values, length
 * `iterations`. */
uint32_t oracle_pagerank_pull_f32(uint32_t N, const uint32_t *fwd_outdeg,
                                  const uint32_t *inv_deg, const uint32_t *inv_idx,
                                  const uint32_t *inv_dest, uint32_t iterations,
                                  double epsilon, float *ranks, double *error_total,
                                  double *pr_sum, uint32_t *active_iter, double *err_iter);

/* Same, from a caller's start: init[v] (the scale of the returned ranks,
 * pr / N) times N in FP32 replaces base.  init NULL: the reference's start.
 * Not in the reference (a warm start of the library's pull). */
uint32_t oracle_pagerank_pull_f32_from(uint32_t N, const uint32_t *fwd_outdeg,
                                       const uint32_t *inv_deg, const uint32_t *inv_idx,
                                       const uint32_t *inv_dest, uint32_t iterations,
                                       double epsilon, const float *init, float *ranks,
                                       double *error_total, double *pr_sum,
                                       uint32_t *active_iter, double *err_iter);

/* Same operator in FP64 (the damping constants are the FP32 ones, promoted). */
uint32_t oracle_pagerank_pull_f64(uint32_t N, const uint32_t *fwd_outdeg,
                                  const uint32_t *inv_deg, const uint32_t *inv_idx,
                                  const uint32_t *inv_dest, uint32_t iterations,
                                  double epsilon, double *ranks, double *error_total);

/* STANDARD formulation, FP64: x0 = 1/N; dangling mass redistributed uniformly;
 * runs `iterations` sweeps or until the L1 residual sum|x'-x| <= tol (tol<=0:
 * fixed count).  Returns sweeps done; *residual = last L1 residual. */
uint32_t oracle_pagerank_standard_f64(uint32_t N, const uint32_t *fwd_outdeg,
                                      const uint32_t *inv_deg, const uint32_t *inv_idx,
                                      const uint32_t *inv_dest, double damp,
                                      uint32_t iterations, double tol, double *x,
                                      double *residual);

/* Same from x0 = init (N doubles; NULL: 1/N). */
uint32_t oracle_pagerank_standard_f64_from(uint32_t N, const uint32_t *fwd_outdeg,
                                           const uint32_t *inv_deg, const uint32_t *inv_idx,
                                           const uint32_t *inv_dest, double damp,
                                           uint32_t iterations, double tol, const double *init,
                                           double *x, double *residual);

/* Personalized PageRank, FP64, one query: teleport vector s uniform over the
 * seed list (duplicates add mass); dangling mass returned to s.
 * x' = (1-d) s + d (A^T D^-1 x + dangling(x) s); x0 = s.  Stops when the L1
 * residual <= tol or after max_iter sweeps. */
uint32_t oracle_ppr_f64(uint32_t N, const uint32_t *fwd_outdeg, const uint32_t *inv_deg,
                        const uint32_t *inv_idx, const uint32_t *inv_dest, double damp,
                        const uint32_t *seeds, uint32_t nseeds, double tol,
                        uint32_t max_iter, double *x, double *residual);

/* Reference data-driven pull PageRank (pageRank.c:2810-2968), FP32, serial.
 * Sweep t gathers only the rows on the worklist; a row whose change is
 * >= epsilon takes its new rank and puts its forward out-neighbours on the
 * next worklist, a row below epsilon keeps its old rank.  The first worklist
 * holds every vertex.  Returns the reported iteration count; ranks[] receives
 * pr/N.  active_iter/err_iter/work_iter (may be NULL, length `iterations`)
 * receive per sweep the updated count, the error total (summed over the
 * worklist) and the worklist size. */
uint32_t oracle_pagerank_dd_pull_f32(uint32_t N, const uint32_t *fwd_outdeg,
                                     const uint32_t *fwd_idx, const uint32_t *fwd_dest,
                                     const uint32_t *inv_deg, const uint32_t *inv_idx,
                                     const uint32_t *inv_dest, uint32_t iterations,
                                     double epsilon, float *ranks, double *error_total,
                                     double *pr_sum, uint32_t *active_iter, double *err_iter,
                                     uint32_t *work_iter);

/* Same from the start init[v] * N (FP32; NULL: base), every vertex on the
 * first worklist.  Not in the reference (a warm start of the library's pull). */
uint32_t oracle_pagerank_dd_pull_f32_from(uint32_t N, const uint32_t *fwd_outdeg,
                                          const uint32_t *fwd_idx, const uint32_t *fwd_dest,
                                          const uint32_t *inv_deg, const uint32_t *inv_idx,
                                          const uint32_t *inv_dest, uint32_t iterations,
                                          double epsilon, const float *init, float *ranks,
                                          double *error_total, double *pr_sum,
                                          uint32_t *active_iter, double *err_iter,
                                          uint32_t *work_iter);

/* SPMVPullGraphCSR (SPMV.c:622-750), unweighted (graphConfig.h:7
 * WEIGHTED 0): the input vector of :699-706, x[v] = 1/out_degree(v) (FP32
 * divide) or 0.001f without out-edges ... */
void oracle_spmv_input(uint32_t N, const uint32_t *fwd_outdeg, float *x);
/* ... y[v] = sum of x over v's inverse edge range, FP32, in edge order (the
 * gather the SPMV compute unit runs with auxiliary1 = x, auxiliary2 = y). */
void oracle_spmv_pull_f32(uint32_t N, const uint32_t *inv_deg, const uint32_t *inv_idx,
                          const uint32_t *inv_dest, const float *x, float *y);
/* The printed checksum (:735-739): sum of (int)(y*100 + .5) / 100.0. */
double oracle_spmv_sum(uint32_t N, const float *y);

/* Locality relabelling, the reference's -l/-L/-O modes 0-10 (README.md:362-411;
 * the implementation is in the absent OpenGraph reorder module, so this
 * restates the published algorithms with the conventions below).  labels[v]
 * = new label of vertex v (the EdgeList label_array of BFS.c:301).  Degrees
 * count the edge list (out: src, in: dst, in+out: both); avg = E / N in
 * uint32 (the structs' avg_degree); a hub has degree > avg.
 *   1-3   degree sort: descending degree, ties by ascending id
 *   4-5   DBG (Faldu et al., IISWC 2019): 8 groups, d >= 32avg, >= 16avg,
 *         >= 8avg, >= 4avg, >= 2avg, >= avg, 2d >= avg, rest; groups in that
 *         order, ascending id inside a group
 *   6-7   HUBSort (Zhang et al., 2017): hubs by descending degree (ties by
 *         id), then the other vertices in id order
 *   8-9   HUBCluster (Balaji and Lucia, IISWC 2018): hubs in id order, then
 *         the other vertices in id order
 *   10    random: ascending 32-bit key lowbias32(v * 0x9E3779B9 ^ seed),
 *         ties by id
 * Returns 0, or -1 for an unknown mode. */
int oracle_reorder_labels(uint32_t N, uint32_t E, const uint32_t *src, const uint32_t *dst,
                          uint32_t mode, uint32_t seed, uint32_t *labels);

/* OpenMP CPU baselines (reported beside the GPU, never the product path). */
double oracle_csr_build_omp(uint32_t N, uint32_t E, const uint32_t *src, const uint32_t *dst,
                            uint32_t *fwd_src, uint32_t *fwd_dst, uint32_t *fwd_outdeg,
                            uint32_t *fwd_idx, uint32_t *inv_src, uint32_t *inv_dst,
                            uint32_t *inv_outdeg, uint32_t *inv_idx, int threads);

/* Returns seconds spent in the iteration loop; semantics as the f32 oracle. */
double oracle_pagerank_pull_omp(uint32_t N, const uint32_t *fwd_outdeg,
                                const uint32_t *inv_deg, const uint32_t *inv_idx,
                                const uint32_t *inv_dest, uint32_t iterations,
                                double epsilon, float *ranks, double *error_total,
                                uint32_t *iters_done, int threads);

int oracle_omp_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
