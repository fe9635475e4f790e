/*
 * SPMV over a GraphCSR: drop-in for the reference's
 * 02_capi_graph/include/algorithms/capi/SPMV.h:19-35,51-52.
 * struct SPMVStats is member-for-member the reference's (:19-28).
 *
 * SPMVPullGraphCSR keeps the reference's semantics (SPMV.c:622-750) for its
 * default unweighted build (graphConfig.h WEIGHTED 0): vector_input[v] =
 * 1/out_degree(v), or 0.001 for a vertex without out-edges; vector_output[v]
 * = sum of vector_input[u] over the in-edges u -> v, FP32; the product is run
 * `iterations` times on the same input; the printed sum rounds each output to
 * two decimals.  The gather runs as a HIP kernel on the default device.
 */
#ifndef SPMV_H
#define SPMV_H

#include <stdint.h>

#include "graphConfig.h"
#include "graphCSR.h"

struct SPMVStats
{
    uint32_t iterations;
    struct Arguments *arguments;
    uint32_t num_vertices;
    float *vector_output;
    float *vector_input;
    double time_total;
};

#ifdef __cplusplus
extern "C" {
#endif

/* Both vectors zeroed, iterations 0 (SPMV.c:51-73). */
struct SPMVStats *newSPMVStatsGraphCSR(struct GraphCSR *graph);
void freeSPMVStats(struct SPMVStats *stats);

/* Dispatch on arguments->pushpull (SPMV.c:596-620): 0 and every variant
 * outside this library's scope (push, fixed point) run the pull kernel, as
 * the reference's `default:` branch does. */
struct SPMVStats *SPMVGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct SPMVStats *SPMVPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);

#ifdef __cplusplus
}
#endif

#endif
