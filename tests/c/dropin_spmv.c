/*
 * Drop-in check for SPMV: a C caller written against the reference's API only
 * (graphCSR.h / SPMV.h) builds a GraphCSR with graphCSRPreProcessingStep and
 * runs SPMVGraphCSR (SPMV.c:596-620) twice on it (the second call reuses the
 * resident device graph).  Writes V, iterations, vector_input and both
 * vector_outputs to <out> for the Python test to compare with the oracle.
 * usage: dropin_spmv <edge list> <format 0|1|2> <iterations> <out>
 */
#include <stdio.h>
#include <stdlib.h>

#include "SPMV.h"
#include "graphCSR.h"
#include "graphConfig.h"

int main(int argc, char **argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s <edge list> <format> <iterations> <out>\n", argv[0]);
        return 2;
    }
    struct Arguments args;
    argumentsSetDefaults(&args);
    args.fnameb = argv[1];
    args.fnameb_format = (uint32_t)atoi(argv[2]);
    args.iterations = (uint32_t)atoi(argv[3]);
    args.pushpull = 0;
    args.sort = 0;

    struct GraphCSR *g = graphCSRPreProcessingStep(&args);
    if (!g) { fprintf(stderr, "graphCSRPreProcessingStep failed\n"); return 1; }
    struct SPMVStats *a = SPMVGraphCSR(&args, g);
    struct SPMVStats *b = SPMVGraphCSR(&args, g);
    if (!a || !b) { fprintf(stderr, "SPMVGraphCSR failed\n"); return 1; }

    FILE *f = fopen(argv[4], "wb");
    uint32_t hdr[2] = {a->num_vertices, a->iterations};
    fwrite(hdr, 4, 2, f);
    fwrite(a->vector_input, 4, a->num_vertices, f);
    fwrite(a->vector_output, 4, a->num_vertices, f);
    fwrite(b->vector_output, 4, b->num_vertices, f);
    fclose(f);
    printf("dropin_spmv: V=%u iterations=%u time_total=%f\n", a->num_vertices, a->iterations,
           a->time_total);
    freeSPMVStats(a);
    freeSPMVStats(b);
    graphCSRFree(g);
    return 0;
}
