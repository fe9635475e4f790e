/*
 * struct EdgeList and its I/O: drop-in for the OpenGraph edge list used by the
 * reference (call sites SSSP.c:488,518-520; bellmanFord.c:698-699; field
 * census in 02_capi_graph/src).  The OpenGraph header is absent, so member
 * order and the reader's signature are inferred from those call sites.
 *
 * Binary format (04_test_graphs/TEST/v51_e1021/graph.bin and siblings): little-endian int32
 * (src, dst) pairs, no header.  Weighted (graph.wbin): int32 (src, dst, w)
 * triples.  Text (graph): whitespace-separated "src dst" lines.
 */
#ifndef EDGELIST_H
#define EDGELIST_H

#include <stdint.h>
#include "graphConfig.h"

struct EdgeList
{
    uint32_t num_edges;
    uint32_t num_vertices;
    uint32_t avg_degree;
    uint32_t *edges_array_src;
    uint32_t *edges_array_dest;
#if WEIGHTED
    float *edges_array_weight;
    float max_weight;
#endif
    uint32_t *label_array;          /* identity labels (no reordering) */
    uint32_t *inverse_label_array;
};

#ifdef __cplusplus
extern "C" {
#endif

struct EdgeList *newEdgeList(uint32_t num_edges);
void freeEdgeList(struct EdgeList *edgeList);

/* fnameb_format: 0 text, 1 binary pairs, 2 weighted binary triples (weights
 * are read and dropped, WEIGHTED is 0).  inverse != 0 swaps src and dst.
 * symmetric != 0 appends the reversed copy of every edge.  num_vertices is
 * max id + 1.  Returns NULL and sets agLastError() text on I/O failure. */
struct EdgeList *readEdgeListsFile(const char *fname, uint32_t fnameb_format,
                                   uint8_t inverse, uint32_t symmetric);
struct EdgeList *readEdgeListsbin(const char *fname, uint8_t inverse, uint32_t symmetric,
                                  uint32_t weighted);
struct EdgeList *readEdgeListstxt(const char *fname, uint8_t inverse, uint32_t symmetric,
                                  uint32_t weighted);
int writeEdgeListToBinaryFile(const struct EdgeList *edgeList, const char *fname);

#ifdef __cplusplus
}
#endif

#endif
