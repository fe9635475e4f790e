/*
 * Pull PageRank over a GraphCSR: drop-in for the reference's
 * 02_capi_graph/include/algorithms/capi/pageRank.h:13-37,69-70.
 * struct PageRankStats is member-for-member the reference's (:19-30).
 *
 * pageRankPullGraphCSR keeps the reference's semantics
 * (pageRank.c:942-1108): ranks start at base_pr = 1 - damp, contributions use
 * the forward out-degree and are 0 for vertices without out-edges, the loop
 * stops after `iterations` sweeps or when no vertex changed by >= epsilon
 * (the breaking sweep is not counted), and pageRanks holds pr/N on return.
 * The gather and update run as HIP kernels on the default device.
 */
#ifndef PAGERANK_H
#define PAGERANK_H

#include <stdint.h>

#include "graphConfig.h"
#include "graphCSR.h"

#define Damp 0.85f

struct PageRankStats
{
    float damp;
    float base_pr;
    uint32_t iterations;
    uint32_t num_vertices;
    uint32_t activeVertices;
    uint32_t *realRanks;
    float *pageRanks;
    double time_total;
    double error_total;
};

#ifdef __cplusplus
extern "C" {
#endif

struct PageRankStats *newPageRankStatsGraphCSR(struct GraphCSR *graph);
void freePageRankStats(struct PageRankStats *stats);

/* Dispatch on arguments->pushpull (pageRank.c:874-938): 6 runs the
 * data-driven pull; 0 and every variant outside this library's scope run the
 * pull kernel, as the reference's `default:` branch does. */
struct PageRankStats *pageRankGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:2810-2968 (declared at pageRank.h:84): worklist pull.  Only
 * vertices with an in-neighbour updated in the previous sweep are recomputed;
 * a vertex takes its new rank only when |new - old| >= epsilon.  The first
 * sweep recomputes every vertex; the loop stops when no vertex is updated. */
struct PageRankStats *pageRankDataDrivenPullGraphCSR(struct Arguments *arguments,
                                                     struct GraphCSR *graph);

void pageRankPrint(float *pageRankArray, uint32_t num_vertices);

#ifdef __cplusplus
}
#endif

#endif
