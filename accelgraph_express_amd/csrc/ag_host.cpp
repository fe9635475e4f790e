// Host-side pieces of the library: MT19937, synthetic edge-list generators,
// edge-list file I/O and struct Arguments defaults.  No GPU work here.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "accelgraph_amd.h"
#include "edgeList.h"
#include "graphConfig.h"
#include "mt19937.h"

namespace ag {
void set_last_error(int code, const char *msg);
}

// ---- MT19937 (Matsumoto & Nishimura, ACM TOMACS 8(1), 1998) ------------------
namespace {
constexpr int kMtN = 624, kMtM = 397;
constexpr uint32_t kMatrixA = 0x9908b0dfu, kUpper = 0x80000000u, kLower = 0x7fffffffu;

void mt_seed(mt19937state *s, uint32_t seed)
{
    s->mt[0] = seed;
    for (int i = 1; i < kMtN; i++)
        s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i;
    s->mti = kMtN;
}

uint32_t mt_next(mt19937state *s)
{
    if (s->mti >= kMtN) {
        uint32_t *mt = s->mt;
        int k = 0;
        for (; k < kMtN - kMtM; k++) {
            const uint32_t y = (mt[k] & kUpper) | (mt[k + 1] & kLower);
            mt[k] = mt[k + kMtM] ^ (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
        }
        for (; k < kMtN - 1; k++) {
            const uint32_t y = (mt[k] & kUpper) | (mt[k + 1] & kLower);
            mt[k] = mt[k + (kMtM - kMtN)] ^ (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
        }
        const uint32_t y = (mt[kMtN - 1] & kUpper) | (mt[0] & kLower);
        mt[kMtN - 1] = mt[kMtM - 1] ^ (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
        s->mti = 0;
    }
    uint32_t y = s->mt[s->mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

inline uint32_t scale_u32(uint32_t r, uint32_t n) { return (uint32_t)(((uint64_t)r * n) >> 32); }

void err(int code, const char *msg) { ag::set_last_error(code, msg); }

template <typename F>
void parallel_for(uint64_t n, int threads, F &&f)
{
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = (int)std::min<uint64_t>((uint64_t)threads, std::max<uint64_t>(n, 1));
    if (threads <= 1) {
        for (uint64_t i = 0; i < n; i++) f(i);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) {
        pool.emplace_back([&, t] {
            for (uint64_t i = t; i < n; i += (uint64_t)threads) f(i);
        });
    }
    for (auto &th : pool) th.join();
}
} // namespace

extern "C" {

void initializeMersenneState(mt19937state *state, uint32_t seed) { mt_seed(state, seed); }
uint32_t generateRandInt(mt19937state *state) { return mt_next(state); }

void argumentsSetDefaults(struct Arguments *a)
{
    std::memset(a, 0, sizeof(*a));
    a->binSize = 512;
    a->iterations = 20;
    a->trials = 1;
    a->epsilon = 0.0001;
    a->cache_size = 32768;
    a->delta = 1;
    a->pre_numThreads = a->algo_numThreads = a->ker_numThreads =
        (int)std::max(1u, std::thread::hardware_concurrency());
    a->fnameb_format = 1;
    a->convert_format = 1;
    a->afu_config = a->afu_config_2 = 0x01;
    a->cu_config = a->cu_config_2 = a->cu_config_3 = a->cu_config_4 = 0x01;
    initializeMersenneState(&a->mt19937var, 27491095u);
}

int ag_gen_uniform(uint32_t seed, uint32_t n, uint32_t m, uint32_t *src, uint32_t *dst)
{
    if (m && (!src || !dst)) { err(AG_ERR_INVALID, "null output"); return AG_ERR_INVALID; }
    if (m && n == 0) { err(AG_ERR_INVALID, "edges need vertices"); return AG_ERR_INVALID; }
    mt19937state s;
    mt_seed(&s, seed);
    for (uint32_t e = 0; e < m; e++) {
        src[e] = scale_u32(mt_next(&s), n);
        dst[e] = scale_u32(mt_next(&s), n);
    }
    err(AG_OK, "");
    return AG_OK;
}

int ag_gen_rmat(uint32_t seed, uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                int permute, uint32_t *src, uint32_t *dst, int threads)
{
    if (scale == 0 || scale > 31) { err(AG_ERR_INVALID, "scale must be in 1..31"); return AG_ERR_INVALID; }
    if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0)) {
        err(AG_ERR_INVALID, "R-MAT probabilities must be >= 0 with a+b+c <= 1");
        return AG_ERR_INVALID;
    }
    const uint64_t n = 1ull << scale;
    const uint64_t m = n * edge_factor;
    if (m > 0xffffffffull) { err(AG_ERR_INVALID, "edge count exceeds 2^32-1"); return AG_ERR_INVALID; }
    if (m && (!src || !dst)) { err(AG_ERR_INVALID, "null output"); return AG_ERR_INVALID; }
    const double two32 = 4294967296.0;
    const uint64_t ta = (uint64_t)(a * two32), tab = (uint64_t)((a + b) * two32),
                   tabc = (uint64_t)((a + b + c) * two32);
    constexpr uint64_t kBlockEdges = 65536;
    const uint64_t nblocks = (m + kBlockEdges - 1) / kBlockEdges;
    parallel_for(nblocks, threads, [&](uint64_t blk) {
        mt19937state s;
        mt_seed(&s, seed ^ (uint32_t)(blk * 0x9E3779B9ull));
        const uint64_t e0 = blk * kBlockEdges, e1 = std::min(m, e0 + kBlockEdges);
        for (uint64_t e = e0; e < e1; e++) {
            uint32_t u = 0, v = 0;
            for (uint32_t l = 0; l < scale; l++) {
                const uint64_t r = mt_next(&s);
                const uint32_t bit = scale - 1 - l;
                if (r < ta) {
                } else if (r < tab) {
                    v |= 1u << bit;
                } else if (r < tabc) {
                    u |= 1u << bit;
                } else {
                    u |= 1u << bit;
                    v |= 1u << bit;
                }
            }
            src[e] = u;
            dst[e] = v;
        }
    });
    if (permute) {
        std::vector<uint32_t> perm(n);
        for (uint64_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
        mt19937state s;
        mt_seed(&s, seed + 1u);
        for (uint64_t i = n - 1; i >= 1; i--) {
            const uint64_t j = ((uint64_t)mt_next(&s) * (i + 1)) >> 32;
            std::swap(perm[i], perm[j]);
        }
        const uint64_t chunk = 1u << 20;
        parallel_for((m + chunk - 1) / chunk, threads, [&](uint64_t q) {
            const uint64_t e0 = q * chunk, e1 = std::min(m, e0 + chunk);
            for (uint64_t e = e0; e < e1; e++) {
                src[e] = perm[src[e]];
                dst[e] = perm[dst[e]];
            }
        });
    }
    err(AG_OK, "");
    return AG_OK;
}

// ---- EdgeList ------------------------------------------------------------------
struct EdgeList *newEdgeList(uint32_t num_edges)
{
    struct EdgeList *el = (struct EdgeList *)std::calloc(1, sizeof(struct EdgeList));
    if (!el) return nullptr;
    el->num_edges = num_edges;
    el->edges_array_src = (uint32_t *)std::malloc((size_t)std::max<uint32_t>(num_edges, 1) * 4);
    el->edges_array_dest = (uint32_t *)std::malloc((size_t)std::max<uint32_t>(num_edges, 1) * 4);
    if (!el->edges_array_src || !el->edges_array_dest) {
        std::free(el->edges_array_src);
        std::free(el->edges_array_dest);
        std::free(el);
        return nullptr;
    }
    return el;
}

void freeEdgeList(struct EdgeList *el)
{
    if (!el) return;
    std::free(el->edges_array_src);
    std::free(el->edges_array_dest);
    std::free(el->label_array);
    std::free(el->inverse_label_array);
    std::free(el);
}

static void finish_edge_list(struct EdgeList *el)
{
    uint32_t mx = 0;
    for (uint32_t e = 0; e < el->num_edges; e++)
        mx = std::max(mx, std::max(el->edges_array_src[e], el->edges_array_dest[e]));
    el->num_vertices = el->num_edges ? mx + 1 : 0;
    el->avg_degree = el->num_vertices ? el->num_edges / el->num_vertices : 0;
}

static struct EdgeList *build_list(std::vector<uint32_t> &s, std::vector<uint32_t> &d,
                                   uint8_t inverse, uint32_t symmetric)
{
    const uint64_t m = s.size();
    const uint64_t total = symmetric ? 2 * m : m;
    if (total > 0xffffffffull) { err(AG_ERR_INVALID, "edge count exceeds 2^32-1"); return nullptr; }
    struct EdgeList *el = newEdgeList((uint32_t)total);
    if (!el) { err(AG_ERR_OOM, "edge list allocation failed"); return nullptr; }
    for (uint64_t e = 0; e < m; e++) {
        el->edges_array_src[e] = inverse ? d[e] : s[e];
        el->edges_array_dest[e] = inverse ? s[e] : d[e];
        if (symmetric) {
            el->edges_array_src[m + e] = el->edges_array_dest[e];
            el->edges_array_dest[m + e] = el->edges_array_src[e];
        }
    }
    finish_edge_list(el);
    err(AG_OK, "");
    return el;
}

struct EdgeList *readEdgeListsFile(const char *fname, uint32_t format, uint8_t inverse,
                                   uint32_t symmetric)
{
    char msg[512];
    if (!fname) { err(AG_ERR_INVALID, "null file name"); return nullptr; }
    FILE *f = std::fopen(fname, "rb");
    if (!f) {
        std::snprintf(msg, sizeof(msg), "cannot open %s: %s", fname, std::strerror(errno));
        err(AG_ERR_IO, msg);
        return nullptr;
    }
    std::vector<uint32_t> s, d;
    if (format == 0) {
        long long a, b;
        while (std::fscanf(f, "%lld %lld", &a, &b) == 2) {
            if (a < 0 || b < 0 || a > 0xfffffffeLL || b > 0xfffffffeLL) {
                std::fclose(f);
                std::snprintf(msg, sizeof(msg), "%s: vertex id out of range", fname);
                err(AG_ERR_IO, msg);
                return nullptr;
            }
            s.push_back((uint32_t)a);
            d.push_back((uint32_t)b);
        }
    } else {
        const int words = (format == 2) ? 3 : 2;
        std::fseek(f, 0, SEEK_END);
        const long long bytes = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        if (bytes < 0 || bytes % (4 * words) != 0) {
            std::fclose(f);
            std::snprintf(msg, sizeof(msg), "%s: size %lld is not a multiple of %d bytes", fname,
                          bytes, 4 * words);
            err(AG_ERR_IO, msg);
            return nullptr;
        }
        const uint64_t m = (uint64_t)bytes / (4 * words);
        std::vector<int32_t> buf((size_t)m * words);
        if (m && std::fread(buf.data(), 4, buf.size(), f) != buf.size()) {
            std::fclose(f);
            std::snprintf(msg, sizeof(msg), "%s: short read", fname);
            err(AG_ERR_IO, msg);
            return nullptr;
        }
        s.resize(m);
        d.resize(m);
        for (uint64_t e = 0; e < m; e++) {
            const int32_t a = buf[e * words], b = buf[e * words + 1];
            if (a < 0 || b < 0) {
                std::fclose(f);
                std::snprintf(msg, sizeof(msg), "%s: negative vertex id at edge %llu", fname,
                              (unsigned long long)e);
                err(AG_ERR_IO, msg);
                return nullptr;
            }
            s[e] = (uint32_t)a;
            d[e] = (uint32_t)b;
        }
    }
    std::fclose(f);
    return build_list(s, d, inverse, symmetric);
}

struct EdgeList *readEdgeListsbin(const char *fname, uint8_t inverse, uint32_t symmetric,
                                  uint32_t weighted)
{
    return readEdgeListsFile(fname, weighted ? 2u : 1u, inverse, symmetric);
}

struct EdgeList *readEdgeListstxt(const char *fname, uint8_t inverse, uint32_t symmetric,
                                  uint32_t weighted)
{
    (void)weighted;
    return readEdgeListsFile(fname, 0u, inverse, symmetric);
}

int writeEdgeListToBinaryFile(const struct EdgeList *el, const char *fname)
{
    if (!el || !fname) { err(AG_ERR_INVALID, "null argument"); return AG_ERR_INVALID; }
    FILE *f = std::fopen(fname, "wb");
    if (!f) { err(AG_ERR_IO, "cannot open output file"); return AG_ERR_IO; }
    std::vector<uint32_t> buf((size_t)el->num_edges * 2);
    for (uint32_t e = 0; e < el->num_edges; e++) {
        buf[2 * (size_t)e] = el->edges_array_src[e];
        buf[2 * (size_t)e + 1] = el->edges_array_dest[e];
    }
    const bool ok = buf.empty() || std::fwrite(buf.data(), 4, buf.size(), f) == buf.size();
    std::fclose(f);
    if (!ok) { err(AG_ERR_IO, "short write"); return AG_ERR_IO; }
    err(AG_OK, "");
    return AG_OK;
}

} // extern "C"
