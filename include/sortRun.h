/*
 * sortRunAlgorithms: drop-in for OpenGraph's edge-list sort (call sites
 * SSSP.c:548-553, bellmanFord.c:700-748; modes README.md:434-439).
 *   sort 0  radix sort by src            (stable)
 *   sort 1  radix sort by src then dst   (lexicographic)
 *   sort 2  count sort by src            (stable)
 *   sort 3  count sort by src then dst   (lexicographic)
 * 0 and 2 produce the same permutation, as do 1 and 3 (both algorithms are
 * stable).  Here every mode runs as GPU radix passes (hand-written HIP
 * kernels); the list is sorted in place and returned.  num_vertices must be
 * set (bellmanFord.c:698-699); keys >= num_vertices are an error (NULL).
 */
#ifndef SORTRUN_H
#define SORTRUN_H

#include <stdint.h>
#include "edgeList.h"

#ifdef __cplusplus
extern "C" {
#endif

struct EdgeList *sortRunAlgorithms(struct EdgeList *edgeList, uint32_t sort);

#ifdef __cplusplus
}
#endif

#endif
