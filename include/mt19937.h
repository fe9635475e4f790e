/*
 * MT19937 state used by struct Arguments (graphConfig.h:40 of the reference
 * names the type mt19937state; its header lives in the absent OpenGraph
 * submodule, so the layout here is this library's own).  The generator is
 * the published MT19937 (Matsumoto & Nishimura, 1998): seed 5489 gives
 * 3499211612 as its first output.
 */
#ifndef MT19937_H
#define MT19937_H

#include <stdint.h>

typedef struct mt19937state
{
    uint32_t mt[624];
    int mti;
} mt19937state;

#ifdef __cplusplus
extern "C" {
#endif

/* Names as called at accel-graph.c:329 and test_open-graph-match.c:110. */
void initializeMersenneState(mt19937state *state, uint32_t seed);
uint32_t generateRandInt(mt19937state *state);

#ifdef __cplusplus
}
#endif

#endif
