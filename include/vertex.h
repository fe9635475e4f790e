/*
 * struct Vertex: per-vertex CSR arrays.  Field names as read by the reference
 * at capienv.c:409-411 and pageRank.c:1027,2901-2902 (the OpenGraph header
 * that defines it is absent; member order is this library's).
 */
#ifndef VERTEX_H
#define VERTEX_H

#include <stdint.h>

struct Vertex
{
    uint32_t *out_degree; /* segment length in the owning sorted edge array */
    uint32_t *in_degree;  /* the other direction's out_degree */
    uint32_t *edges_idx;  /* segment start; 0 when the segment is empty */
};

#ifdef __cplusplus
extern "C" {
#endif

struct Vertex *newVertexArray(uint32_t num_vertices);
void freeVertexArray(struct Vertex *vertices);

#ifdef __cplusplus
}
#endif

#endif
