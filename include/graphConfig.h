/*
 * Run-time arguments of the hot path: drop-in for the reference's
 * 02_capi_graph/include/config/graphConfig.h:1-53 (struct Arguments, same
 * member names, types and order; WEIGHTED 0 and DIRECTED 1 as at :7-8).
 *
 * Members the GPU path reads: iterations, epsilon, pushpull, sort, verbosity,
 * fnameb, fnameb_format, symmetric, weighted.  ker_numThreads, algo_numThreads,
 * pre_numThreads and the CAPI afu_config / cu_config words are accepted and
 * ignored (they size OpenMP pools and program the FPGA in the reference).
 */
#ifndef GRAPHCONFIG_H
#define GRAPHCONFIG_H

#include <stdint.h>
#include "mt19937.h"

#define WEIGHTED 0
#define DIRECTED 1

struct Arguments
{
    int wflag;
    int xflag;
    int sflag;
    int Sflag;
    int dflag;
    uint32_t binSize;
    uint32_t verbosity;
    uint32_t iterations;
    uint32_t trials;
    double epsilon;
    int source;
    uint32_t algorithm;
    uint32_t datastructure;
    uint32_t pushpull;
    uint32_t sort;
    uint32_t lmode;
    uint32_t lmode_l2;
    uint32_t lmode_l3;
    uint32_t mmode;
    uint32_t symmetric;
    uint32_t weighted;
    uint32_t delta;
    int pre_numThreads;
    int algo_numThreads;
    int ker_numThreads;
    char *fnameb;
    char *fnamel;
    uint32_t fnameb_format;
    uint32_t convert_format;
    mt19937state mt19937var;
    uint32_t cache_size;

    /* CAPI integration words: accepted, ignored by the GPU path */
    uint64_t afu_config;
    uint64_t afu_config_2;
    uint64_t cu_config;
    uint64_t cu_config_2;
    uint64_t cu_config_3;
    uint64_t cu_config_4;
};

#ifdef __cplusplus
extern "C" {
#endif

/* Fills the defaults of the reference CLI (accel-graph.c:288-329):
 * iterations 20, epsilon 1e-4, pushpull 0, sort 0, binSize 512, cache_size
 * 32768, fnameb_format 1, convert_format 1, MT19937 seed 27491095. */
void argumentsSetDefaults(struct Arguments *arguments);

#ifdef __cplusplus
}
#endif

#endif
