/*
 * accelgraph_amd.h - C ABI of the MI355X (gfx950) CSR-builder / pull-PageRank
 * engine.  Plain C types and pointers only; every call returns an AG_* status
 * and leaves a message for ag_last_error() (thread-local) on failure.
 *
 * This layer replaces the reference's CAPI runtime (capienv.h:146-166:
 * mapGraphCSRToWED / setupAFUGraphCSR / startAFU / startCU / waitAFU /
 * releaseAFU), which packs CSR pointers into a WED descriptor and drives the
 * FPGA gather (pageRank.c:993-1046).  Here a context owns one HIP stream on
 * one device, a graph owns its device-resident CSR, and one call runs the
 * whole PageRank loop on the device.  The reference-named functions in
 * graphCSR.h / sortRun.h / pageRank.h are built on top of this API.
 */
#ifndef ACCELGRAPH_AMD_H
#define ACCELGRAPH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define AG_OK              0
#define AG_ERR_INVALID     1  /* bad argument, shape or state */
#define AG_ERR_HIP         2  /* HIP runtime error (message names the call) */
#define AG_ERR_OOM         3  /* device or host allocation failed */
#define AG_ERR_ADMISSION   4  /* serving: HBM budget or slot limit refused */
#define AG_ERR_NOT_FOUND   5  /* unknown ticket / graph id */
#define AG_ERR_CANCELLED   6  /* serving: query cancelled */
#define AG_ERR_IO          7  /* file I/O */
#define AG_ERR_INJECTED    8  /* serving: injected fault (test hook) */
#define AG_ERR_PENDING     9  /* serving: query not finished yet */
#define AG_ERR_NO_DEVICE  10  /* no HIP device visible */

const char *ag_last_error(void);
const char *ag_version(void);

/* ---- context ------------------------------------------------------------ */
typedef struct ag_ctx ag_ctx;

/* device < 0 selects the current HIP device. */
int ag_ctx_create(int device, ag_ctx **out);
int ag_ctx_destroy(ag_ctx *ctx);
int ag_ctx_synchronize(ag_ctx *ctx);
/* hipStream_t of the context, as an opaque pointer. */
void *ag_ctx_stream(ag_ctx *ctx);
/* Process-wide default context on the current device (created on first use;
 * the reference-named drop-in functions use it). */
int ag_default_ctx(ag_ctx **out);

/* ---- device-resident graph ---------------------------------------------- */
typedef struct ag_graph ag_graph;

/* sort modes (sortRun.h); flags may add: */
#define AG_SORT_RADIX_SRC        0u
#define AG_SORT_RADIX_SRC_DST    1u
#define AG_SORT_COUNT_SRC        2u
#define AG_SORT_COUNT_SRC_DST    3u
#define AG_BUILD_INVERSE_FROM_INPUT 0x100u /* inverse from the input order, not the sorted one */
#define AG_BUILD_NO_INVERSE         0x200u /* forward CSR only */
#define AG_BUILD_DEVICE_INPUT       0x400u /* src/dst are device pointers */

/* Builds forward + inverse CSR on the GPU from an edge list (host pointers
 * unless AG_BUILD_DEVICE_INPUT).  num_vertices 0 means max id + 1.  With
 * num_vertices given, an id >= num_vertices is found by the sort's first
 * pass and reported (AG_ERR_INVALID, no graph) when the build ends. */
int ag_csr_build(ag_ctx *ctx, const uint32_t *src, const uint32_t *dst, uint32_t num_edges,
                 uint32_t num_vertices, uint32_t flags, ag_graph **out);

/* Uploads an existing CSR (host arrays in the reference layout) holding only
 * what pull PageRank reads: forward out-degree, inverse degree / edges_idx /
 * dest (capienv.c:409,421-426). */
int ag_graph_from_csr(ag_ctx *ctx, uint32_t num_vertices, uint32_t num_edges,
                      const uint32_t *fwd_out_degree, const uint32_t *inv_out_degree,
                      const uint32_t *inv_edges_idx, const uint32_t *inv_dest,
                      ag_graph **out);

int ag_graph_free(ag_graph *g);
uint32_t ag_graph_num_vertices(const ag_graph *g);
uint32_t ag_graph_num_edges(const ag_graph *g);
/* Device bytes held by the graph (CSR + gather layout once built). */
uint64_t ag_graph_device_bytes(const ag_graph *g);

/* Device pointers of the CSR arrays (NULL where not built). */
typedef struct ag_csr_view
{
    uint32_t num_vertices, num_edges;
    const uint32_t *out_degree, *in_degree, *edges_idx, *sorted_src, *sorted_dest;
    const uint32_t *inv_out_degree, *inv_in_degree, *inv_edges_idx, *inv_sorted_src,
        *inv_sorted_dest;
    const uint32_t *offsets, *inv_offsets; /* exclusive scans, N + 1 entries */
} ag_csr_view;
int ag_graph_view(const ag_graph *g, ag_csr_view *out);

/* Copies CSR arrays to host memory; any pointer may be NULL (skipped).
 * Vertex arrays hold N entries, edge arrays E. */
int ag_graph_download(ag_ctx *ctx, const ag_graph *g,
                      uint32_t *out_degree, uint32_t *in_degree, uint32_t *edges_idx,
                      uint32_t *sorted_src, uint32_t *sorted_dest,
                      uint32_t *inv_out_degree, uint32_t *inv_in_degree,
                      uint32_t *inv_edges_idx, uint32_t *inv_sorted_src,
                      uint32_t *inv_sorted_dest);

/* CSR binary serialization (no counterpart file format in the reference: its
 * graphCSR-binary layout, README.md:295-306, lives in the absent OpenGraph
 * module).  Writes the graph's CSR arrays to `path` in the AGCSRv1 layout:
 * a 32-byte header (magic "AGCSRv1\0", N, E, section mask, reserved,
 * FNV-1a-64 of the payload) followed by the present sections as
 * little-endian uint32 arrays (bit 0..9: sorted src, sorted dest,
 * out_degree, edges_idx, offsets, then the same five for the inverse). */
int ag_graph_save(ag_ctx *ctx, const ag_graph *g, const char *path);
/* Loads an AGCSRv1 file onto the device: checks the size against the header,
 * the checksum, offsets[N] == E and vertex ids < N (AG_ERR_INVALID on any
 * mismatch, AG_ERR_IO on file errors). */
int ag_graph_load(ag_ctx *ctx, const char *path, ag_graph **out);
/* Host-only check of an AGCSRv1 file (header, size, checksum): no device. */
int ag_csr_file_info(const char *path, uint32_t *num_vertices, uint32_t *num_edges,
                     uint32_t *sections);

/* Stable sort of (key, val) pairs by key (keys < key_limit) on the device:
 * the radix passes used by the builder, exposed for sortRunAlgorithms.  All
 * four pointers are host memory; in-place is allowed. */
int ag_sort_pairs_host(ag_ctx *ctx, const uint32_t *keys, const uint32_t *vals, uint32_t n,
                       uint32_t key_limit, uint32_t *keys_out, uint32_t *vals_out);

/* Vertex mapping of a key-sorted array (host in/out): degree and edges_idx. */
int ag_vertex_map_host(ag_ctx *ctx, const uint32_t *sorted_keys, uint32_t n, uint32_t num_vertices,
                       uint32_t *degree, uint32_t *edges_idx);

/* ---- pull PageRank ------------------------------------------------------ */
#define AG_PR_REFERENCE 0u /* pageRank.c:942-1108 formulation */
#define AG_PR_STANDARD  1u /* x0 = 1/N, dangling mass spread uniformly, FP64 dangling sum */
#define AG_PR_DATA_DRIVEN 2u /* pageRank.c:2810-2968 worklist pull: only rows with an updated
                                in-neighbour are recomputed, a row takes its new rank only when
                                the change is >= epsilon (FP32 storage only) */

#define AG_PREC_FP32 0u    /* contributions stored FP32 */
#define AG_PREC_BF16 1u    /* stored BF16 (RNE), FP32 accumulate */
#define AG_PREC_FP16 2u    /* stored FP16 (RNE), FP32 accumulate */

#define AG_PR_FLAG_FINISH_FP32 0x1u /* low precision: switch to FP32 storage when the residual stalls */
#define AG_PR_FLAG_TIME_SWEEPS 0x2u /* time the sweep kernels alone (HIP events around each) */
#define AG_PR_FLAG_EXTRAPOLATE 0x4u /* AG_PR_STANDARD (and personalized PageRank through
                                       ag_ppr_single_opts): Aitken-type extrapolated steps
                                       once the residual falls by a steady ratio; the run
                                       stops only on a plain sweep's residual */
#define AG_PR_FLAG_WARM_START  0x8u /* REFERENCE / STANDARD / DATA_DRIVEN: start from the rank
                                       vector the caller passes in the output buffer (`ranks`
                                       / `d_ranks`, N floats, read before it is overwritten)
                                       instead of the reference's uniform start.  The vector
                                       is in the output's scale (the ranks a previous call
                                       returned): the sweeps read it times N in REFERENCE and
                                       DATA_DRIVEN modes, as is in STANDARD mode, where the
                                       first sweep's dangling mass is summed from it.  No
                                       counterpart in the reference (pageRank.c:942-1108
                                       always starts at base).  In REFERENCE mode with N a
                                       power of two, k sweeps then a warm start with m more
                                       equal k + m sweeps bit for bit; DATA_DRIVEN restarts
                                       with every vertex on the worklist */

typedef struct ag_pr_options
{
    uint32_t iterations; /* sweep cap */
    uint32_t mode;       /* AG_PR_* */
    uint32_t precision;  /* AG_PREC_* */
    uint32_t flags;      /* AG_PR_FLAG_* */
    double epsilon;      /* REFERENCE / DATA_DRIVEN: per-vertex active threshold (0: fixed sweeps);
                            STANDARD: stop when the L1 residual <= epsilon (0: fixed) */
    float damp;          /* 0.85f */
    float reserved;
} ag_pr_options;

void ag_pr_options_default(ag_pr_options *opt);

typedef struct ag_pr_result
{
    uint32_t iterations;      /* reported count (reference semantics in REFERENCE mode) */
    uint32_t sweeps;          /* sweeps executed on the device */
    uint32_t active_vertices; /* active count of the last sweep */
    uint32_t fp32_switch_sweep; /* first FP32-stored sweep (precision modes), else 0 */
    double error_total;       /* REFERENCE: mean |delta|; DATA_DRIVEN: |delta|/N summed over
                                 the last worklist; STANDARD: L1 residual */
    double pr_sum;            /* sum of returned ranks (FP64) */
    double loop_ms;           /* device time of the sweep loop (HIP events); when the whole
                                 fixed-count loop is replayed as one hipGraph (epsilon <= 0,
                                 no AG_PR_FLAG_TIME_SWEEPS), the replay incl. init + normalise */
    double total_ms;          /* device time init + loop + normalise */
    double plan_ms;           /* gather-layout build time (0 if cached) */
    double sweep_kernel_ms;   /* AG_PR_FLAG_TIME_SWEEPS: device time of the sweep kernels
                                 alone, summed over sweeps (finalize launches excluded) */
} ag_pr_result;

/* Runs PageRank on a resident graph.  ranks (host, N floats) may be NULL
 * unless AG_PR_FLAG_WARM_START is set (it then holds the start on entry);
 * active_iter / err_iter (host, `iterations` entries) may be NULL. */
int ag_pagerank_pull(ag_ctx *ctx, ag_graph *g, const ag_pr_options *opt, float *ranks,
                     uint32_t *active_iter, double *err_iter, ag_pr_result *res);

/* Same, writing ranks to a device buffer (N floats).  The sweep loop (init,
 * sweeps, normalise) is captured into hipGraphs cached on the graph and
 * replayed on later calls with the same arguments; AG_NO_GRAPH=1 in the
 * environment runs it eagerly.  AG_PR_FLAG_TIME_SWEEPS always runs eagerly. */
int ag_pagerank_pull_device(ag_ctx *ctx, ag_graph *g, const ag_pr_options *opt,
                            float *d_ranks, ag_pr_result *res);

/* Allocation-free form of ag_pagerank_pull_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It calls no hipMalloc / hipFree, creates no event or graph, and takes its
 * per-call buffers from a static 16 MiB device array that each GPU the library
 * runs on holds until the process ends (about 1.3 M vertices at 20 sweeps;
 * larger graphs are refused with AG_ERR_INVALID).  The graph's gather layout
 * must exist (ag_pagerank_prepare).  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU race.  loop_ms / total_ms come from
 * the host clock; AG_PR_FLAG_TIME_SWEEPS is refused. */
int ag_pagerank_pull_device_noalloc(ag_ctx *ctx, ag_graph *g, const ag_pr_options *opt,
                                    float *d_ranks, ag_pr_result *res);

/* ---- locality relabelling (the reference's -l/-L/-O options, README.md:362-411) ----
 * labels[v] = new label of vertex v: the EdgeList label_array the reference's
 * algorithms read (BFS.c:301, SSSP.c:583).  The reference's implementation is
 * in the unavailable OpenGraph reorder module; these modes restate the
 * published algorithms.  Degrees count the edge list (out: as src, in: as
 * dst); avg = E / N (uint32); a hub has degree > avg. */
#define AG_REORDER_NONE            0u  /* identity */
#define AG_REORDER_OUT_DEGREE      1u  /* descending degree, ties by id */
#define AG_REORDER_IN_DEGREE       2u
#define AG_REORDER_INOUT_DEGREE    3u
#define AG_REORDER_DBG_OUT         4u  /* 8 degree groups at avg * 2^k, hottest first, id order inside */
#define AG_REORDER_DBG_IN          5u
#define AG_REORDER_HUBSORT_OUT     6u  /* hubs by descending degree, then the rest in id order */
#define AG_REORDER_HUBSORT_IN      7u
#define AG_REORDER_HUBCLUSTER_OUT  8u  /* hubs in id order, then the rest in id order */
#define AG_REORDER_HUBCLUSTER_IN   9u
#define AG_REORDER_RANDOM         10u  /* seeded random permutation */

/* Labels for `mode` from a host edge list (ids < num_vertices). */
int ag_reorder_labels(ag_ctx *ctx, const uint32_t *src, const uint32_t *dst, uint32_t num_edges,
                      uint32_t num_vertices, uint32_t mode, uint32_t seed, uint32_t *labels);
/* Same on device arrays. */
int ag_reorder_labels_device(ag_ctx *ctx, const uint32_t *d_src, const uint32_t *d_dst,
                             uint32_t num_edges, uint32_t num_vertices, uint32_t mode,
                             uint32_t seed, uint32_t *d_labels);
/* src[e] = labels[src[e]], dst[e] = labels[dst[e]] in place (host arrays).
 * Refuses labels that are not a permutation of [0, num_vertices) and ids out
 * of range (AG_ERR_INVALID), leaving the arrays unchanged. */
int ag_relabel_edges(ag_ctx *ctx, const uint32_t *labels, uint32_t num_vertices, uint32_t *src,
                     uint32_t *dst, uint32_t num_edges);
/* Same on device arrays. */
int ag_relabel_edges_device(ag_ctx *ctx, const uint32_t *d_labels, uint32_t num_vertices,
                            uint32_t *d_src, uint32_t *d_dst, uint32_t num_edges);

/* ---- SPMV pull (SPMV.c:622-750, SPMVPullGraphCSR) ---------------------- */
typedef struct ag_spmv_result
{
    uint32_t iterations; /* sweeps run (the reference restarts its compute unit per iteration) */
    uint32_t reserved;
    double y_sum;        /* FP64 sum of y */
    double loop_ms;      /* device time of the sweeps (HIP events) */
    double total_ms;     /* device time including input relabel and output un-relabel */
    double plan_ms;      /* gather-layout build time (0 if cached) */
} ag_spmv_result;

/* y[v] = sum of x[u] over v's in-edges u -> v (unweighted, the reference's
 * WEIGHTED 0 build), FP32 accumulation, repeated `iterations` times on the
 * same input (0: y is all zero, as the reference's output buffer starts).
 * Host buffers of N floats. */
int ag_spmv_pull(ag_ctx *ctx, ag_graph *g, const float *x, float *y, uint32_t iterations,
                 ag_spmv_result *res);

/* Same on device buffers (N floats each, not aliased). */
int ag_spmv_pull_device(ag_ctx *ctx, ag_graph *g, const float *d_x, float *d_y,
                        uint32_t iterations, ag_spmv_result *res);

/* Builds (or rebuilds) the degree-binned gather layout of g now, so that a
 * timed call does not include it.  Returns its build time in *ms. */
int ag_pagerank_prepare(ag_ctx *ctx, ag_graph *g, double *ms);

/* Gather-layout statistics (for tuning and reports). */
typedef struct ag_plan_info
{
    uint32_t n_hub, n_mid, n_low, n_zero; /* vertices per in-degree bin */
    uint32_t hub_chunks;                  /* workgroup chunks over hub rows */
    uint32_t low_max, hub_min;            /* bin thresholds (in-degree) */
    uint64_t sell_slots;                  /* padded slots in the low-degree slices */
    uint64_t plan_bytes;
    uint32_t hub_chunk;                   /* edges per hub chunk chosen for this graph (the
                                             hub rows' summation order depends on it) */
    uint32_t layout;                      /* AG_LAYOUT_TABLES or AG_LAYOUT_ROWS */
    uint32_t tables;                      /* partitioned layout: LDS tables (cold excluded) */
    uint32_t table_workgroups;            /* partitioned layout: table-kernel workgroups */
    uint64_t subrows;                     /* partitioned layout: (row, table) sub-rows */
    uint64_t cold_edges;                  /* partitioned layout: padded index entries gathered
                                             from global memory (sources past the last LDS
                                             table) */
    uint32_t lane_max;                    /* partitioned layout: one-lane sub-row limit */
    uint32_t max_tables;                  /* partitioned layout: LDS table limit */
    uint32_t autotuned;                   /* 1 when chosen by ag_pagerank_autotune */
} ag_plan_info;
int ag_pagerank_plan_info(const ag_graph *g, ag_plan_info *info);

/* Kernel tuning knobs (0 keeps the default).  Applies to plans built later. */
typedef struct ag_tuning
{
    uint32_t low_max;     /* in-degree <= low_max: thread per vertex */
    uint32_t hub_min;     /* in-degree >= hub_min: workgroup chunks */
    uint32_t hub_chunk;   /* edges per hub workgroup chunk; AG_TUNE_AUTO (the default):
                             4,096 when the hub rows make at least 8 such chunks per
                             compute unit, else 2,048 */
    uint32_t hot_bytes;   /* LDS bytes per sweep workgroup caching the hottest
                             contributions (multiple of 16; 1 disables the table) */
    uint32_t layout;      /* gather layout of the pull sweep: AG_LAYOUT_AUTO (default),
                             AG_LAYOUT_TABLES or AG_LAYOUT_ROWS; 0 keeps the current one */
    uint32_t max_tables;  /* partitioned layout: at most this many LDS tables (1..62) */
    uint32_t lane_max;    /* partitioned layout: sub-rows of up to this many edges are
                             summed by one lane (multiple of 4, 4..64; default 16) */
} ag_tuning;
#define AG_TUNE_AUTO 0xFFFFFFFFu
/* Partitioned tables: source labels are cut into LDS-sized tables (40,440
 * sources each), every workgroup of the sweep holds one table and sums each
 * row's in-edges from it; a combine kernel adds each row's partial sums in
 * table order.  Rows: the degree-binned row layout with one LDS table of the
 * hottest sources (its low-degree rows sum in the reference's edge order,
 * bit-exact).  Auto: tables when the sources outnumber one table, else rows
 * (whose LDS table then holds every source). */
#define AG_LAYOUT_TABLES 1u
#define AG_LAYOUT_ROWS   2u
#define AG_LAYOUT_AUTO   3u
int ag_set_tuning(const ag_tuning *t);
int ag_get_tuning(ag_tuning *t);

/* Chooses g's gather layout by timing candidates on this device (no
 * counterpart in the reference, whose gather has no layout to tune): from g's
 * current tuning (the process-wide one unless autotuned before), one pass of
 * coordinate search over max_tables {8,12,16,20,24,31,40,48,62} and lane_max
 * {8,12,16,24} (partitioned tables; low_max {16,32,64,128} for the row
 * layout), hub_min {4096,6144,8192,12288} and hub_chunk {auto,1024,2048,4096}.
 * Each candidate layout is built and timed over `sweeps` (1..1000) FP32 pull
 * sweeps (the faster of two graph replays); one replaces the best only when
 * at least 1% faster.  The fastest becomes g's layout and g's own tuning (later
 * rebuilds use it), reported by ag_pagerank_plan_info; *best (optional)
 * receives it and *ms (optional) the wall time of the search.  Builds the
 * layout up to 22 times (about 20 ms each at 4 M vertices and 67 M edges) and
 * holds two layouts at once; it must not overlap other calls on g, nor run
 * while a server holds g (its batches point at the layout).  A layout fixes
 * the summation order of the ranks (every result stays within the parity
 * bar); two searches may choose differently when candidates time within
 * noise, so a run that must be bit-reproducible sets the reported tuning
 * with ag_set_tuning instead of searching. */
int ag_pagerank_autotune(ag_ctx *ctx, ag_graph *g, uint32_t sweeps, ag_tuning *best, double *ms);

/* ---- batched personalized PageRank -------------------------------------- */
typedef struct ag_ppr_query
{
    const uint32_t *seeds; /* teleport set, mass split evenly (duplicates add) */
    uint32_t nseeds;
    uint32_t max_iter;     /* 0: 1000 */
    double tol;            /* stop at L1 residual <= tol */
} ag_ppr_query;

typedef struct ag_ppr_stat
{
    uint32_t iterations;
    uint32_t status;       /* AG_OK or AG_ERR_* */
    double residual;
    double latency_ms;     /* serving: submit -> done */
    uint32_t fine_sweep;   /* 16-bit storage: sweeps done before the column read full
                              FP32 contributions (0: FP32 throughout) */
    uint32_t reserved;
} ag_ppr_stat;

typedef struct ag_ppr_options
{
    uint32_t precision;       /* AG_PREC_FP32, or AG_PREC_BF16: each FP32 contribution
                                 split into a high (BF16) and a low 16-bit plane; a
                                 column reads the high plane alone until its residual
                                 stalls, then both (exact FP32) until it converges */
    uint32_t cols_per_wave;   /* 0: autotuned from {4, 8, 16, 32, 64} */
    float damp;               /* 0.85f */
    uint32_t use_graph_replay;/* 1: capture sweeps in a hipGraph and replay */
    uint32_t extrapolate;     /* 1: a column whose residual falls by a steady ratio rho
                                 takes the step x + rho / (1 - rho) (T(x) - x)
                                 (Aitken-type); the stop rule still reads a plain
                                 sweep's residual.  0 (default): power iteration */
} ag_ppr_options;
void ag_ppr_options_default(ag_ppr_options *opt);

/* Runs nq queries as multi-right-hand-side columns with per-column retirement
 * (at most 64 columns at a time; later queries take the slots that free up).
 * x (host, nq * N floats, row q = query q) may be NULL; stats (nq) may be NULL.
 * AG_PREC_FP16 is refused here (AG_ERR_INVALID); ag_ppr_single accepts it. */
int ag_ppr_batch(ag_ctx *ctx, ag_graph *g, const ag_ppr_query *q, uint32_t nq,
                 const ag_ppr_options *opt, float *x, ag_ppr_stat *stats, double *elapsed_ms);

/* One query with the single-vector sweep kernel (the one-query-at-a-time
 * path): x (host, N floats) may be NULL.  precision AG_PREC_BF16/FP16 stores
 * contributions in 16 bits and finishes in FP32 once the residual stalls. */
int ag_ppr_single(ag_ctx *ctx, ag_graph *g, const ag_ppr_query *q, uint32_t precision, float *x,
                  ag_ppr_stat *stat);
/* ag_ppr_single with options: precision, damp and extrapolate are honoured
 * (cols_per_wave and use_graph_replay do not apply to one vector). */
int ag_ppr_single_opts(ag_ctx *ctx, ag_graph *g, const ag_ppr_query *q, const ag_ppr_options *opt,
                       float *x, ag_ppr_stat *stat);

/* Chooses cols_per_wave for g by timing {4, 8, 16, 32, 64} (BF16: from 8) on a
 * full 64-column batch; result cached per graph and precision. */
int ag_ppr_autotune(ag_ctx *ctx, ag_graph *g, uint32_t precision, uint32_t *cols_per_wave);
/* Device bytes a batch of `capacity` columns (8..64, multiple of 8) holds on g:
 * what the serving loop reserves against its HBM budget per admitted batch. */
int ag_ppr_batch_bytes(ag_ctx *ctx, ag_graph *g, uint32_t capacity, uint32_t precision,
                       uint64_t *bytes);

/* ---- serving loop -------------------------------------------------------- */
typedef struct ag_server ag_server;

typedef struct ag_serve_config
{
    uint64_t hbm_budget_bytes; /* admission limit for column state (0: 8 GiB) */
    uint32_t max_columns;      /* per-graph column slots in flight (0: 64) */
    uint32_t precision;        /* AG_PREC_FP32 or AG_PREC_BF16 (as ag_ppr_options) */
    uint32_t cols_per_wave;    /* 0: autotune */
    uint32_t inject_fail_alloc;/* test hook: refuse the n-th admission (0: off) */
    uint32_t inject_fail_query;/* test hook: fail ticket n (0: off) */
    float damp;
    uint32_t extrapolate;      /* as ag_ppr_options.extrapolate (0: off) */
} ag_serve_config;
void ag_serve_config_default(ag_serve_config *cfg);

int ag_serve_create(ag_ctx *ctx, const ag_serve_config *cfg, ag_server **out);
int ag_serve_destroy(ag_server *s);
/* Registers a resident graph (not owned). */
int ag_serve_add_graph(ag_server *s, ag_graph *g, uint32_t *graph_id);
/* Queues a query; it is admitted into a column slot when memory and slots free. */
int ag_serve_submit(ag_server *s, uint32_t graph_id, const uint32_t *seeds, uint32_t nseeds,
                    double tol, uint32_t max_iter, uint64_t *ticket);
int ag_serve_cancel(ag_server *s, uint64_t ticket);
/* Advances the scheduler: admits, runs one sweep for every graph with active
 * columns, retires converged columns.  *active receives the number of queries
 * still queued or running. */
int ag_serve_step(ag_server *s, uint32_t *active);
/* Result of a finished query (x: N floats, may be NULL).  AG_ERR_PENDING if
 * not finished.  The result is released after a successful poll. */
int ag_serve_poll(ag_server *s, uint64_t ticket, float *x, uint32_t n, ag_ppr_stat *st);
/* Zero-copy view of a finished query's result: *x points at *n floats (vertex
 * order) in pinned host memory written by the GPU, valid until
 * ag_serve_release(ticket).  AG_ERR_PENDING while the query runs; the query's
 * status code once it has ended (then *x is NULL unless it succeeded). */
int ag_serve_peek(ag_server *s, uint64_t ticket, const float **x, uint32_t *n, ag_ppr_stat *st);
/* Frees a finished query and its result buffer (after ag_serve_peek). */
int ag_serve_release(ag_server *s, uint64_t ticket);
/* Bytes of HBM currently reserved for column state. */
uint64_t ag_serve_reserved_bytes(const ag_server *s);

/* ---- synthetic inputs (host, deterministic) ------------------------------ */
/* Uniform random directed edges: MT19937(seed); per edge src then dst, each
 * floor(r * N / 2^32).  Self-loops and duplicates kept. */
int ag_gen_uniform(uint32_t seed, uint32_t num_vertices, uint32_t num_edges, uint32_t *src,
                   uint32_t *dst);
/* R-MAT (a, b, c, d = 1 - a - b - c), 2^scale vertices; edge e draws one
 * 32-bit MT19937 word per level (most significant bit first) from the
 * stream of its 65,536-edge block, seeded with seed ^ (block * 0x9E3779B9),
 * and picks quadrant a/b/c/d as (src,dst) bit (0,0)/(0,1)/(1,0)/(1,1);
 * permute != 0
 * relabels vertices by a Fisher-Yates shuffle from MT19937(seed + 1).
 * threads <= 0: all hardware threads.  Output is independent of threads. */
int ag_gen_rmat(uint32_t seed, uint32_t scale, uint32_t edge_factor, double a, double b,
                double c, int permute, uint32_t *src, uint32_t *dst, int threads);

#ifdef __cplusplus
}
#endif

#endif
