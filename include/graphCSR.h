/*
 * struct GraphCSR and its builder: drop-in for the OpenGraph CSR used by the
 * reference (fields read at pageRank.c:958-969, capienv.c:401-433; builder
 * call sequence SSSP.c:485-488,548-563).  The OpenGraph header is absent:
 * field names are the reference's, member order and the private device
 * handle at the end are this library's.
 */
#ifndef GRAPHCSR_H
#define GRAPHCSR_H

#include <stdint.h>
#include "graphConfig.h"
#include "edgeList.h"
#include "vertex.h"

struct GraphCSR
{
    uint32_t num_edges;
    uint32_t num_vertices;
    uint32_t avg_degree;
#if WEIGHTED
    float max_weight;
#endif
    struct Vertex *vertices;
    struct EdgeList *sorted_edges_array;
#if DIRECTED
    struct Vertex *inverse_vertices;
    struct EdgeList *inverse_sorted_edges_array;
#endif
    /* private: device-resident copy built by this library (NULL until the GPU
     * builder or the first PageRank call creates it); owned by the graph. */
    void *ag_device;
};

#ifdef __cplusplus
extern "C" {
#endif

/* Allocates the struct and its Vertex arrays (inverse != 0: also the inverse
 * side); edge arrays are attached by graphCSRAssignEdgeList. */
struct GraphCSR *graphCSRNew(uint32_t V, uint32_t E, uint8_t inverse);

/* Takes ownership of a SORTED edge list and fills the matching Vertex arrays
 * (inverse 0: vertices/sorted_edges_array; 1: inverse_*).  The vertex
 * mapping runs on the GPU. */
struct GraphCSR *graphCSRAssignEdgeList(struct GraphCSR *graph, struct EdgeList *edgeList,
                                        uint8_t inverse);

/* Read arguments->fnameb, sort (arguments->sort), build forward and inverse
 * CSR on the GPU, and return the host GraphCSR with its device copy kept
 * resident.  The inverse list is the forward-sorted list with endpoints
 * swapped (SSSP.c:518-520), sorted the same way. */
struct GraphCSR *graphCSRPreProcessingStep(struct Arguments *arguments);

/* Same, from an in-memory edge list (the list is not modified). */
struct GraphCSR *graphCSRFromEdgeList(const struct EdgeList *edgeList, uint32_t sort);

void graphCSRFree(struct GraphCSR *graph);
void graphCSRPrint(const struct GraphCSR *graph);

#ifdef __cplusplus
}
#endif

#endif
