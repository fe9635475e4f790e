// Context, error reporting and the device CSR builder.
//
// ag_csr_build is the device form of OpenGraph's graph construction as the
// reference calls it (SSSP.c:485-488,518-520,548-563): sort the edge list by
// source (sortRunAlgorithms), map vertices (graphCSRAssignEdgeList(.., 0)),
// swap endpoints of the forward-sorted list, sort again and map the inverse
// side (graphCSRAssignEdgeList(.., 1)).
#include <cstdarg>
#include <cstring>
#include <mutex>

#include "ag_internal.h"

namespace ag {

namespace {
thread_local int t_last_code = AG_OK;
thread_local char t_last_msg[1024] = "";
std::mutex g_default_mu;
ag_ctx *g_default = nullptr;
Tuning g_tuning;
} // namespace

void set_last_error(int code, const char *msg)
{
    t_last_code = code;
    std::snprintf(t_last_msg, sizeof(t_last_msg), "%s", msg ? msg : "");
}

void fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Error(code, buf);
}

Tuning &tuning() { return g_tuning; }

void check_ctx(ag_ctx *ctx)
{
    if (!ctx) fail(AG_ERR_INVALID, "null context");
    AG_HIP(hipSetDevice(ctx->device));
}

static ag_ctx *make_ctx(int device)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        fail(AG_ERR_NO_DEVICE, "no HIP device visible: the GPU path has no CPU fallback");
    if (device < 0) AG_HIP(hipGetDevice(&device));
    if (device >= count) fail(AG_ERR_INVALID, "device %d out of range (%d devices)", device, count);
    AG_HIP(hipSetDevice(device));
    std::unique_ptr<ag_ctx> c(new ag_ctx());
    c->device = device;
    hipDeviceProp_t prop;
    AG_HIP(hipGetDeviceProperties(&prop, device));
    c->num_cus = prop.multiProcessorCount;
    c->arch = prop.gcnArchName;
    AG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    noalloc_scratch_init(device);
    return c.release();
}

ag_ctx *default_ctx()
{
    std::lock_guard<std::mutex> lk(g_default_mu);
    if (!g_default) g_default = make_ctx(-1);
    return g_default;
}

} // namespace ag

using namespace ag;

uint64_t ag_graph::bytes() const
{
    uint64_t b = fwd_src.bytes + fwd_dst.bytes + fwd_outdeg.bytes + fwd_idx.bytes +
                 fwd_off.bytes + inv_src.bytes + inv_dst.bytes + inv_outdeg.bytes +
                 inv_idx.bytes + inv_off.bytes;
    ag_plan_info info{};
    if (plan && ag_pagerank_plan_info(this, &info) == AG_OK) b += info.plan_bytes;
    return b;
}

extern "C" {

const char *ag_last_error(void) { return t_last_msg; }
const char *ag_version(void) { return "accelgraph-amd 0.1 (gfx950)"; }

int ag_ctx_create(int device, ag_ctx **out)
{
    return guard([&] {
        if (!out) fail(AG_ERR_INVALID, "null out");
        *out = make_ctx(device);
    });
}

int ag_ctx_destroy(ag_ctx *ctx)
{
    return guard([&] {
        if (!ctx) return;
        {
            std::lock_guard<std::mutex> lk(g_default_mu);
            if (ctx == g_default) g_default = nullptr;
        }
        (void)hipSetDevice(ctx->device);
        if (ctx->stream) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipStreamDestroy(ctx->stream);
        }
        delete ctx;
    });
}

int ag_ctx_synchronize(ag_ctx *ctx)
{
    return guard([&] {
        check_ctx(ctx);
        AG_HIP(hipStreamSynchronize(ctx->stream));
    });
}

void *ag_ctx_stream(ag_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int ag_default_ctx(ag_ctx **out)
{
    return guard([&] {
        if (!out) fail(AG_ERR_INVALID, "null out");
        *out = default_ctx();
    });
}

int ag_set_tuning(const ag_tuning *t)
{
    return guard([&] {
        if (!t) fail(AG_ERR_INVALID, "null tuning");
        Tuning n = g_tuning;
        if (t->low_max) n.low_max = t->low_max;
        if (t->hub_min) n.hub_min = t->hub_min;
        if (t->hub_chunk) n.hub_chunk = t->hub_chunk;
        if (t->hot_bytes) n.hot_bytes = t->hot_bytes == 1 ? 0u : t->hot_bytes;
        if (t->layout) n.layout = t->layout;
        if (t->max_tables) n.max_tables = t->max_tables;
        if (t->lane_max) n.lane_max = t->lane_max;
        if (n.layout != AG_LAYOUT_TABLES && n.layout != AG_LAYOUT_ROWS && n.layout != AG_LAYOUT_AUTO)
            fail(AG_ERR_INVALID, "layout must be AG_LAYOUT_AUTO, AG_LAYOUT_TABLES or AG_LAYOUT_ROWS");
        if (n.max_tables < 1 || n.max_tables > 62) fail(AG_ERR_INVALID, "max_tables must be 1..62");
        if (n.lane_max < 4 || n.lane_max > 64 || n.lane_max % 4)
            fail(AG_ERR_INVALID, "lane_max must be a multiple of 4 in 4..64");
        if (n.hot_bytes % 16 || n.hot_bytes > 160 * 1024 - 2048)
            fail(AG_ERR_INVALID, "hot_bytes must be a multiple of 16 and <= 161792");
        if (n.low_max >= n.hub_min) fail(AG_ERR_INVALID, "low_max must be < hub_min");
        if (n.hub_chunk < 256) fail(AG_ERR_INVALID, "hub_chunk must be >= 256");
        g_tuning = n;
    });
}

int ag_get_tuning(ag_tuning *t)
{
    return guard([&] {
        if (!t) fail(AG_ERR_INVALID, "null tuning");
        t->low_max = g_tuning.low_max;
        t->hub_min = g_tuning.hub_min;
        t->hub_chunk = g_tuning.hub_chunk;
        t->hot_bytes = g_tuning.hot_bytes;
        t->layout = g_tuning.layout;
        t->max_tables = g_tuning.max_tables;
        t->lane_max = g_tuning.lane_max;
    });
}

int ag_csr_build(ag_ctx *ctx, const uint32_t *src, const uint32_t *dst, uint32_t E, uint32_t N,
                 uint32_t flags, ag_graph **out)
{
    return guard([&] {
        PhaseRange phase_("ag_csr_build");
        check_ctx(ctx);
        if (!out) fail(AG_ERR_INVALID, "null out");
        if (E && (!src || !dst)) fail(AG_ERR_INVALID, "null edge arrays");
        const uint32_t mode = flags & 0xffu;
        if (mode > 3) fail(AG_ERR_INVALID, "sort mode %u not in 0..3", mode);
        hipStream_t s = ctx->stream;
        std::unique_ptr<ag_graph> g(new ag_graph());
        g->E = E;
        std::lock_guard<std::recursive_mutex> lk(ctx->scratch_mu);

        const uint32_t *ds = src, *dd = dst;
        if (!(flags & AG_BUILD_DEVICE_INPUT) && E) {
            void *in_src = scratch(ctx, kScrInSrc, (size_t)E * 4);
            void *in_dst = scratch(ctx, kScrInDst, (size_t)E * 4);
            AG_HIP(hipMemcpyAsync(in_src, src, (size_t)E * 4, hipMemcpyHostToDevice, s));
            AG_HIP(hipMemcpyAsync(in_dst, dst, (size_t)E * 4, hipMemcpyHostToDevice, s));
            ds = static_cast<const uint32_t *>(in_src);
            dd = static_cast<const uint32_t *>(in_dst);
        }
        // With num_vertices given, the ids are checked after the build: the
        // first pass of each direction's sort takes the largest key on the
        // way (dmax), and the vertex mapping skips ids >= N until then.
        // Without it, N = max id + 1 is needed before the sort.
        uint32_t *dmax = nullptr;
        if (N == 0) {
            const uint32_t maxid = device_max2(ctx, ds, dd, E);
            if (E && maxid == 0xffffffffu) fail(AG_ERR_INVALID, "vertex id 0xffffffff");
            N = E ? maxid + 1 : 0;
        } else if (E) {
            dmax = scratch_as<uint32_t>(ctx, kScrMax, 1);
            AG_HIP(hipMemsetAsync(dmax, 0, 4, s));
            if (flags & AG_BUILD_NO_INVERSE) device_max_async(ctx, dd, nullptr, E, dmax);
        }
        g->N = N;
        const uint32_t bits = bits_for(N);
        const bool by_src_dst = (mode == 1 || mode == 3);

        // the sorted arrays are allocated once the sort's first pass is
        // queued, so that hipMalloc's host time overlaps device work
        auto sort_into = [&](const uint32_t *k, const uint32_t *v, DevBuf &ko, DevBuf &vo) {
            auto out = [&](uint32_t **kk, uint32_t **vv) {
                ko.alloc((size_t)E * 4);
                vo.alloc((size_t)E * 4);
                *kk = ko.as<uint32_t>();
                *vv = vo.as<uint32_t>();
            };
            if (by_src_dst) {
                uint32_t *ta = static_cast<uint32_t *>(scratch(ctx, kScrPairA, (size_t)E * 4));
                uint32_t *tb = static_cast<uint32_t *>(scratch(ctx, kScrPairB, (size_t)E * 4));
                radix_sort_pairs(ctx, v, k, tb, ta, E, bits);
                radix_sort_pairs_late(ctx, ta, tb, E, bits, out, dmax);
            } else {
                radix_sort_pairs_late(ctx, k, v, E, bits, out, dmax);
            }
        };

        std::unique_ptr<PhaseRange> half(new PhaseRange("csr.forward"));
        sort_into(ds, dd, g->fwd_src, g->fwd_dst);
        g->fwd_outdeg.alloc((size_t)N * 4);
        g->fwd_idx.alloc((size_t)N * 4);
        g->fwd_off.alloc((size_t)(N + 1) * 4);
        vertex_map(ctx, g->fwd_src.as<uint32_t>(), E, N, g->fwd_outdeg.as<uint32_t>(),
                   g->fwd_idx.as<uint32_t>(), g->fwd_off.as<uint32_t>());
        g->has_full_fwd = true;

        half.reset();
        if (!(flags & AG_BUILD_NO_INVERSE)) {
            PhaseRange inv_phase("csr.inverse");
            const bool from_input = (flags & AG_BUILD_INVERSE_FROM_INPUT) != 0;
            const uint32_t *ik = from_input ? dd : g->fwd_dst.as<uint32_t>();
            const uint32_t *iv = from_input ? ds : g->fwd_src.as<uint32_t>();
            sort_into(ik, iv, g->inv_src, g->inv_dst);
            g->inv_outdeg.alloc((size_t)N * 4);
            g->inv_idx.alloc((size_t)N * 4);
            g->inv_off.alloc((size_t)(N + 1) * 4);
            vertex_map(ctx, g->inv_src.as<uint32_t>(), E, N, g->inv_outdeg.as<uint32_t>(),
                       g->inv_idx.as<uint32_t>(), g->inv_off.as<uint32_t>());
            g->has_inverse_full = true;
        }
        uint32_t maxid = 0;
        if (dmax) AG_HIP(hipMemcpyAsync(&maxid, dmax, 4, hipMemcpyDeviceToHost, s));
        AG_HIP(hipStreamSynchronize(s));
        if (dmax && maxid >= N) fail(AG_ERR_INVALID, "vertex id %u >= num_vertices %u", maxid, N);
        *out = g.release();
    });
}

int ag_graph_from_csr(ag_ctx *ctx, uint32_t N, uint32_t E, const uint32_t *fwd_out_degree,
                      const uint32_t *inv_out_degree, const uint32_t *inv_edges_idx,
                      const uint32_t *inv_dest, ag_graph **out)
{
    return guard([&] {
        PhaseRange phase_("ag_graph_from_csr");
        check_ctx(ctx);
        if (!out) fail(AG_ERR_INVALID, "null out");
        if (N && (!fwd_out_degree || !inv_out_degree || !inv_edges_idx))
            fail(AG_ERR_INVALID, "null vertex arrays");
        if (E && !inv_dest) fail(AG_ERR_INVALID, "null inverse edge array");
        // Every gather reads [edges_idx, edges_idx + degree): check the ranges
        // here so that no kernel can read outside the edge array.
        uint64_t fsum = 0, isum = 0;
        for (uint32_t v = 0; v < N; v++) {
            if ((uint64_t)inv_edges_idx[v] + inv_out_degree[v] > E)
                fail(AG_ERR_INVALID, "inverse segment of vertex %u exceeds num_edges", v);
            fsum += fwd_out_degree[v];
            isum += inv_out_degree[v];
        }
        if (isum != E) fail(AG_ERR_INVALID, "inverse degrees sum to %llu, not %u",
                            (unsigned long long)isum, E);
        (void)fsum;
        hipStream_t s = ctx->stream;
        std::unique_ptr<ag_graph> g(new ag_graph());
        g->N = N;
        g->E = E;
        g->fwd_outdeg.alloc((size_t)N * 4);
        g->inv_outdeg.alloc((size_t)N * 4);
        g->inv_idx.alloc((size_t)N * 4);
        g->inv_dst.alloc((size_t)E * 4);
        AG_HIP(hipMemcpyAsync(g->fwd_outdeg.p, fwd_out_degree, (size_t)N * 4,
                              hipMemcpyHostToDevice, s));
        AG_HIP(hipMemcpyAsync(g->inv_outdeg.p, inv_out_degree, (size_t)N * 4,
                              hipMemcpyHostToDevice, s));
        AG_HIP(hipMemcpyAsync(g->inv_idx.p, inv_edges_idx, (size_t)N * 4,
                              hipMemcpyHostToDevice, s));
        if (E)
            AG_HIP(hipMemcpyAsync(g->inv_dst.p, inv_dest, (size_t)E * 4, hipMemcpyHostToDevice,
                                  s));
        if (E) {
            const uint32_t m = device_max1(ctx, g->inv_dst.as<uint32_t>(), E);
            if (m >= N) fail(AG_ERR_INVALID, "inverse edge endpoint %u >= num_vertices %u", m, N);
        }
        AG_HIP(hipStreamSynchronize(s));
        *out = g.release();
    });
}

int ag_graph_free(ag_graph *g)
{
    return guard([&] { delete g; });
}

uint32_t ag_graph_num_vertices(const ag_graph *g) { return g ? g->N : 0; }
uint32_t ag_graph_num_edges(const ag_graph *g) { return g ? g->E : 0; }
uint64_t ag_graph_device_bytes(const ag_graph *g) { return g ? g->bytes() : 0; }

int ag_graph_view(const ag_graph *g, ag_csr_view *v)
{
    return guard([&] {
        if (!g || !v) fail(AG_ERR_INVALID, "null argument");
        std::memset(v, 0, sizeof(*v));
        v->num_vertices = g->N;
        v->num_edges = g->E;
        v->out_degree = g->fwd_outdeg.as<uint32_t>();
        v->in_degree = g->inv_outdeg.as<uint32_t>();
        v->edges_idx = g->fwd_idx.as<uint32_t>();
        v->sorted_src = g->fwd_src.as<uint32_t>();
        v->sorted_dest = g->fwd_dst.as<uint32_t>();
        v->inv_out_degree = g->inv_outdeg.as<uint32_t>();
        v->inv_in_degree = g->fwd_outdeg.as<uint32_t>();
        v->inv_edges_idx = g->inv_idx.as<uint32_t>();
        v->inv_sorted_src = g->inv_src.as<uint32_t>();
        v->inv_sorted_dest = g->inv_dst.as<uint32_t>();
        v->offsets = g->fwd_off.as<uint32_t>();
        v->inv_offsets = g->inv_off.as<uint32_t>();
    });
}

int ag_graph_download(ag_ctx *ctx, const ag_graph *g, uint32_t *out_degree, uint32_t *in_degree,
                      uint32_t *edges_idx, uint32_t *sorted_src, uint32_t *sorted_dest,
                      uint32_t *inv_out_degree, uint32_t *inv_in_degree,
                      uint32_t *inv_edges_idx, uint32_t *inv_sorted_src,
                      uint32_t *inv_sorted_dest)
{
    return guard([&] {
        PhaseRange phase_("ag_graph_download");
        check_ctx(ctx);
        if (!g) fail(AG_ERR_INVALID, "null graph");
        hipStream_t s = ctx->stream;
        auto cp = [&](uint32_t *dst, const DevBuf &b, uint32_t n, const char *what) {
            if (!dst || n == 0) return;
            if (!b.p) fail(AG_ERR_INVALID, "%s not built for this graph", what);
            AG_HIP(hipMemcpyAsync(dst, b.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        };
        cp(out_degree, g->fwd_outdeg, g->N, "out_degree");
        cp(in_degree, g->inv_outdeg, g->N, "in_degree");
        cp(edges_idx, g->fwd_idx, g->N, "edges_idx");
        cp(sorted_src, g->fwd_src, g->E, "sorted src");
        cp(sorted_dest, g->fwd_dst, g->E, "sorted dest");
        cp(inv_out_degree, g->inv_outdeg, g->N, "inverse out_degree");
        cp(inv_in_degree, g->fwd_outdeg, g->N, "inverse in_degree");
        cp(inv_edges_idx, g->inv_idx, g->N, "inverse edges_idx");
        cp(inv_sorted_src, g->inv_src, g->E, "inverse sorted src");
        cp(inv_sorted_dest, g->inv_dst, g->E, "inverse sorted dest");
        AG_HIP(hipStreamSynchronize(s));
    });
}

int ag_sort_pairs_host(ag_ctx *ctx, const uint32_t *keys, const uint32_t *vals, uint32_t n,
                       uint32_t key_limit, uint32_t *keys_out, uint32_t *vals_out)
{
    return guard([&] {
        check_ctx(ctx);
        if (n && (!keys || !vals || !keys_out || !vals_out)) fail(AG_ERR_INVALID, "null array");
        if (n == 0) return;
        hipStream_t s = ctx->stream;
        DevBuf k((size_t)n * 4), v((size_t)n * 4), ko((size_t)n * 4), vo((size_t)n * 4);
        AG_HIP(hipMemcpyAsync(k.p, keys, (size_t)n * 4, hipMemcpyHostToDevice, s));
        AG_HIP(hipMemcpyAsync(v.p, vals, (size_t)n * 4, hipMemcpyHostToDevice, s));
        const uint32_t m = device_max1(ctx, k.as<uint32_t>(), n);
        if (key_limit && m >= key_limit)
            fail(AG_ERR_INVALID, "key %u >= key limit %u", m, key_limit);
        radix_sort_pairs(ctx, k.as<uint32_t>(), v.as<uint32_t>(), ko.as<uint32_t>(),
                         vo.as<uint32_t>(), n, bits_for((uint64_t)m + 1));
        AG_HIP(hipMemcpyAsync(keys_out, ko.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        AG_HIP(hipMemcpyAsync(vals_out, vo.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        AG_HIP(hipStreamSynchronize(s));
    });
}

int ag_vertex_map_host(ag_ctx *ctx, const uint32_t *sorted_keys, uint32_t n, uint32_t N,
                       uint32_t *degree, uint32_t *edges_idx)
{
    return guard([&] {
        check_ctx(ctx);
        if (N && (!degree || !edges_idx)) fail(AG_ERR_INVALID, "null output");
        if (n && !sorted_keys) fail(AG_ERR_INVALID, "null keys");
        if (N == 0) {
            if (n) fail(AG_ERR_INVALID, "edges but no vertices");
            return;
        }
        hipStream_t s = ctx->stream;
        DevBuf k((size_t)n * 4), d((size_t)N * 4), x((size_t)N * 4);
        if (n) AG_HIP(hipMemcpyAsync(k.p, sorted_keys, (size_t)n * 4, hipMemcpyHostToDevice, s));
        if (n) {
            const uint32_t m = device_max1(ctx, k.as<uint32_t>(), n);
            if (m >= N) fail(AG_ERR_INVALID, "key %u >= num_vertices %u", m, N);
        }
        vertex_map(ctx, k.as<uint32_t>(), n, N, d.as<uint32_t>(), x.as<uint32_t>(), nullptr);
        AG_HIP(hipMemcpyAsync(degree, d.p, (size_t)N * 4, hipMemcpyDeviceToHost, s));
        AG_HIP(hipMemcpyAsync(edges_idx, x.p, (size_t)N * 4, hipMemcpyDeviceToHost, s));
        AG_HIP(hipStreamSynchronize(s));
    });
}

} // extern "C"
