/*
 * Drop-in check: a C caller written against the reference's API only
 * (graphCSR.h / edgeList.h / sortRun.h / pageRank.h), as the reference's
 * algorithms build graphs (SSSP.c:485-488,518-520,548-563) and run PageRank
 * (pageRank.c:874-938).  Two graphs are built from the same edge-list file:
 *   A: graphCSRPreProcessingStep(arguments)              (one call)
 *   B: newEdgeList + sortRunAlgorithms + graphCSRNew +
 *      graphCSRAssignEdgeList(fwd) / (inverse)           (the SSSP.c sequence)
 * Both run pageRankGraphCSR; ranks and CSR arrays are written to <out>.bin
 * for the Python test to compare with the oracle.
 * usage: dropin_pagerank <edge list> <format 0|1|2> <iterations> <epsilon> <out>
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "edgeList.h"
#include "graphCSR.h"
#include "graphConfig.h"
#include "pageRank.h"
#include "sortRun.h"

static int same(const uint32_t *a, const uint32_t *b, uint32_t n)
{
    return n == 0 || memcmp(a, b, (size_t)n * 4) == 0;
}

int main(int argc, char **argv)
{
    if (argc < 6) {
        fprintf(stderr, "usage: %s <edge list> <format> <iterations> <epsilon> <out> [pushpull]\n",
                argv[0]);
        return 2;
    }
    struct Arguments args;
    argumentsSetDefaults(&args);
    args.fnameb = argv[1];
    args.fnameb_format = (uint32_t)atoi(argv[2]);
    args.iterations = (uint32_t)atoi(argv[3]);
    args.epsilon = atof(argv[4]);
    args.pushpull = argc > 6 ? (uint32_t)atoi(argv[6]) : 0; /* 6: data-driven pull */
    args.sort = 0;

    struct GraphCSR *a = graphCSRPreProcessingStep(&args);
    if (!a) { fprintf(stderr, "graphCSRPreProcessingStep failed\n"); return 1; }

    /* B: the builder call sequence of SSSP.c */
    struct EdgeList *in = readEdgeListsFile(args.fnameb, args.fnameb_format, 0, 0);
    uint32_t V = in->num_vertices, E = in->num_edges;
    struct GraphCSR *b = graphCSRNew(V, E, 1);
    struct EdgeList *fw = newEdgeList(E), *iv = newEdgeList(E);
    fw->num_vertices = iv->num_vertices = V;
    memcpy(fw->edges_array_src, in->edges_array_src, (size_t)E * 4);
    memcpy(fw->edges_array_dest, in->edges_array_dest, (size_t)E * 4);
    fw = sortRunAlgorithms(fw, 0);
    for (uint32_t e = 0; e < E; e++) { /* inverse from the sorted list */
        iv->edges_array_src[e] = fw->edges_array_dest[e];
        iv->edges_array_dest[e] = fw->edges_array_src[e];
    }
    iv = sortRunAlgorithms(iv, 0);
    graphCSRAssignEdgeList(b, fw, 0);
    graphCSRAssignEdgeList(b, iv, 1);
    freeEdgeList(in);

    int csr_same = a->num_vertices == b->num_vertices &&
        same(a->vertices->out_degree, b->vertices->out_degree, V) &&
        same(a->vertices->in_degree, b->vertices->in_degree, V) &&
        same(a->vertices->edges_idx, b->vertices->edges_idx, V) &&
        same(a->sorted_edges_array->edges_array_src, b->sorted_edges_array->edges_array_src, E) &&
        same(a->sorted_edges_array->edges_array_dest, b->sorted_edges_array->edges_array_dest, E) &&
        same(a->inverse_vertices->out_degree, b->inverse_vertices->out_degree, V) &&
        same(a->inverse_vertices->edges_idx, b->inverse_vertices->edges_idx, V) &&
        same(a->inverse_sorted_edges_array->edges_array_dest,
             b->inverse_sorted_edges_array->edges_array_dest, E);

    struct PageRankStats *sa = pageRankGraphCSR(&args, a);
    struct PageRankStats *sb = pageRankGraphCSR(&args, b);
    if (!sa || !sb) { fprintf(stderr, "pageRankGraphCSR failed\n"); return 1; }

    FILE *f = fopen(argv[5], "wb");
    uint32_t hdr[4] = {V, E, sa->iterations, (uint32_t)csr_same};
    fwrite(hdr, 4, 4, f);
    fwrite(&sa->error_total, 8, 1, f);
    fwrite(&sb->error_total, 8, 1, f);
    fwrite(&sb->iterations, 4, 1, f);
    fwrite(sa->pageRanks, 4, V, f);
    fwrite(sb->pageRanks, 4, V, f);
    fwrite(a->sorted_edges_array->edges_array_dest, 4, E, f);
    fwrite(a->inverse_sorted_edges_array->edges_array_dest, 4, E, f);
    fwrite(a->inverse_vertices->edges_idx, 4, V, f);
    fclose(f);
    printf("dropin: V=%u E=%u iterations=%u csr_same=%d error_total=%.13f time_total=%f\n", V, E,
           sa->iterations, csr_same, sa->error_total, sa->time_total);
    freePageRankStats(sa);
    freePageRankStats(sb);
    graphCSRFree(a);
    graphCSRFree(b);
    return 0;
}
