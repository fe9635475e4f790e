// Host-side consistency check of one direction of a CSR read from a file,
// before any of it reaches the device: the gather and push kernels walk
// [edges_idx[v], edges_idx[v] + degree[v]) and [offsets[v], offsets[v+1])
// without bounds checks, so every such range must lie inside [0, E).
// Header-only and free of HIP so that a CPU unit test can include it.
#pragma once

#include <cstdint>

namespace ag {

// degree / edges_idx / offsets may each be null (section absent).  Returns
// nullptr when the present arrays are consistent, else what is wrong.
inline const char *csr_direction_check(uint32_t N, uint32_t E, const uint32_t *degree,
                                       const uint32_t *edges_idx, const uint32_t *offsets)
{
    if (offsets) {
        if (offsets[0] != 0) return "offsets[0] is not 0";
        if (offsets[N] != E) return "offsets[N] is not E";
        for (uint32_t v = 0; v < N; v++)
            if (offsets[v + 1] < offsets[v]) return "offsets decrease";
        if (degree)
            for (uint32_t v = 0; v < N; v++)
                if (offsets[v + 1] - offsets[v] != degree[v])
                    return "offsets[v+1] - offsets[v] differs from degree[v]";
    }
    if (degree) {
        uint64_t total = 0;
        for (uint32_t v = 0; v < N; v++) total += degree[v];
        if (total != E) return "degrees do not sum to E";
        if (edges_idx)
            for (uint32_t v = 0; v < N; v++)
                if (degree[v] && (uint64_t)edges_idx[v] + degree[v] > E)
                    return "edges_idx[v] + degree[v] exceeds E";
    } else if (edges_idx) {
        return "edges_idx without degrees";
    }
    return nullptr;
}

} // namespace ag
