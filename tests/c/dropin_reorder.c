/*
 * Drop-in check for the relabelling options: a C caller against the
 * reference's API only sets Arguments.lmode / lmode_l2 / lmode_l3 (and fnamel
 * for mode 11), builds with graphCSRPreProcessingStep and runs
 * pageRankGraphCSR.  Writes N, E, label_array and inverse_label_array of both
 * sorted edge lists, the sorted and inverse-sorted dest arrays and the ranks.
 * usage: dropin_reorder <edge list> <format> <l1> <l2> <l3> <label file|-> <out>
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "graphCSR.h"
#include "graphConfig.h"
#include "pageRank.h"

int main(int argc, char **argv)
{
    if (argc < 8) {
        fprintf(stderr, "usage: %s <edge list> <format> <l1> <l2> <l3> <label file|-> <out>\n",
                argv[0]);
        return 2;
    }
    struct Arguments args;
    argumentsSetDefaults(&args);
    args.fnameb = argv[1];
    args.fnameb_format = (uint32_t)atoi(argv[2]);
    args.lmode = (uint32_t)atoi(argv[3]);
    args.lmode_l2 = (uint32_t)atoi(argv[4]);
    args.lmode_l3 = (uint32_t)atoi(argv[5]);
    args.fnamel = strcmp(argv[6], "-") ? argv[6] : NULL;
    args.iterations = 20;
    args.epsilon = 0.0;

    struct GraphCSR *g = graphCSRPreProcessingStep(&args);
    if (!g) { fprintf(stderr, "graphCSRPreProcessingStep failed\n"); return 1; }
    struct PageRankStats *st = pageRankGraphCSR(&args, g);
    if (!st) { fprintf(stderr, "pageRankGraphCSR failed\n"); return 1; }
    uint32_t V = g->num_vertices, E = g->num_edges;
    FILE *f = fopen(argv[7], "wb");
    uint32_t hdr[2] = {V, E};
    fwrite(hdr, 4, 2, f);
    fwrite(g->sorted_edges_array->label_array, 4, V, f);
    fwrite(g->sorted_edges_array->inverse_label_array, 4, V, f);
    fwrite(g->inverse_sorted_edges_array->label_array, 4, V, f);
    fwrite(g->inverse_sorted_edges_array->inverse_label_array, 4, V, f);
    fwrite(g->sorted_edges_array->edges_array_dest, 4, E, f);
    fwrite(g->inverse_sorted_edges_array->edges_array_dest, 4, E, f);
    fwrite(st->pageRanks, 4, V, f);
    fclose(f);
    printf("dropin_reorder: V=%u E=%u layers %u %u %u\n", V, E, args.lmode, args.lmode_l2,
           args.lmode_l3);
    freePageRankStats(st);
    graphCSRFree(g);
    return 0;
}
