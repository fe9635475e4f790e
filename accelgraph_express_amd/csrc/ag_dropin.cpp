// Reference-named drop-in API (graphCSR.h, sortRun.h, pageRank.h, vertex.h)
// implemented over the accelgraph_amd.h C ABI.  Every compute step runs on the
// GPU; when no device is visible the calls print the reason and return NULL
// (there is no CPU fallback).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "accelgraph_amd.h"
#include "graphCSR.h"
#include "SPMV.h"
#include "pageRank.h"
#include "sortRun.h"
#include "vertex.h"

namespace {

// Seed of the random relabelling (mode 10): the reference CLI's MT19937
// seed (accel-graph.c, argumentsSetDefaults).
constexpr uint32_t kRelabelSeed = 27491095u;

bool quiet()
{
    const char *q = std::getenv("AG_QUIET");
    return q && *q && *q != '0';
}

void report(const char *where)
{
    std::fprintf(stderr, "[accelgraph-amd] %s failed: %s\n", where, ag_last_error());
}

ag_ctx *ctx_or_null(const char *where)
{
    ag_ctx *c = nullptr;
    if (ag_default_ctx(&c) != AG_OK) {
        report(where);
        return nullptr;
    }
    return c;
}

void drop_device(struct GraphCSR *g)
{
    if (g && g->ag_device) {
        ag_graph_free((ag_graph *)g->ag_device);
        g->ag_device = nullptr;
    }
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

} // namespace

extern "C" {

struct Vertex *newVertexArray(uint32_t n)
{
    struct Vertex *v = (struct Vertex *)std::calloc(1, sizeof(struct Vertex));
    if (!v) return nullptr;
    const size_t b = (size_t)std::max<uint32_t>(n, 1) * 4;
    v->out_degree = (uint32_t *)std::calloc(1, b);
    v->in_degree = (uint32_t *)std::calloc(1, b);
    v->edges_idx = (uint32_t *)std::calloc(1, b);
    if (!v->out_degree || !v->in_degree || !v->edges_idx) {
        freeVertexArray(v);
        return nullptr;
    }
    return v;
}

void freeVertexArray(struct Vertex *v)
{
    if (!v) return;
    std::free(v->out_degree);
    std::free(v->in_degree);
    std::free(v->edges_idx);
    std::free(v);
}

// sortRunAlgorithms(el, sort): stable sort by src (0, 2) or by (src, dst)
// (1, 3) on the GPU, in place (SSSP.c:548-553).
struct EdgeList *sortRunAlgorithms(struct EdgeList *el, uint32_t sort)
{
    if (!el) return nullptr;
    if (sort > 3) {
        std::fprintf(stderr, "[accelgraph-amd] sortRunAlgorithms: sort mode %u not in 0..3\n", sort);
        return nullptr;
    }
    ag_ctx *c = ctx_or_null("sortRunAlgorithms");
    if (!c) return nullptr;
    const uint32_t E = el->num_edges;
    uint32_t limit = el->num_vertices;
    if (E == 0) return el;
    if (sort == 1 || sort == 3) {
        if (ag_sort_pairs_host(c, el->edges_array_dest, el->edges_array_src, E, limit,
                               el->edges_array_dest, el->edges_array_src) != AG_OK) {
            report("sortRunAlgorithms");
            return nullptr;
        }
    }
    if (ag_sort_pairs_host(c, el->edges_array_src, el->edges_array_dest, E, limit,
                           el->edges_array_src, el->edges_array_dest) != AG_OK) {
        report("sortRunAlgorithms");
        return nullptr;
    }
    return el;
}

struct GraphCSR *graphCSRNew(uint32_t V, uint32_t E, uint8_t inverse)
{
    struct GraphCSR *g = (struct GraphCSR *)std::calloc(1, sizeof(struct GraphCSR));
    if (!g) return nullptr;
    g->num_vertices = V;
    g->num_edges = E;
    g->avg_degree = V ? E / V : 0;
    g->vertices = newVertexArray(V);
#if DIRECTED
    if (inverse) g->inverse_vertices = newVertexArray(V);
#endif
    if (!g->vertices
#if DIRECTED
        || (inverse && !g->inverse_vertices)
#endif
    ) {
        graphCSRFree(g);
        return nullptr;
    }
    return g;
}

// graphCSRAssignEdgeList(g, el, inverse) (SSSP.c:557-563): takes ownership of
// the sorted list and maps vertices (degree / edges_idx) on the GPU.
struct GraphCSR *graphCSRAssignEdgeList(struct GraphCSR *g, struct EdgeList *el, uint8_t inverse)
{
    if (!g || !el) return nullptr;
    ag_ctx *c = ctx_or_null("graphCSRAssignEdgeList");
    if (!c) return nullptr;
    struct Vertex *v = g->vertices;
#if DIRECTED
    if (inverse) {
        if (!g->inverse_vertices) g->inverse_vertices = newVertexArray(g->num_vertices);
        v = g->inverse_vertices;
    }
#endif
    if (!v) return nullptr;
    if (ag_vertex_map_host(c, el->edges_array_src, el->num_edges, g->num_vertices, v->out_degree,
                           v->edges_idx) != AG_OK) {
        report("graphCSRAssignEdgeList");
        return nullptr;
    }
    drop_device(g);
#if DIRECTED
    if (inverse) {
        if (g->inverse_sorted_edges_array && g->inverse_sorted_edges_array != el)
            freeEdgeList(g->inverse_sorted_edges_array);
        g->inverse_sorted_edges_array = el;
    } else
#endif
    {
        if (g->sorted_edges_array && g->sorted_edges_array != el) freeEdgeList(g->sorted_edges_array);
        g->sorted_edges_array = el;
    }
    g->num_edges = el->num_edges;
    g->avg_degree = g->num_vertices ? g->num_edges / g->num_vertices : 0;
#if DIRECTED
    if (g->inverse_vertices && g->sorted_edges_array && g->inverse_sorted_edges_array) {
        std::memcpy(g->vertices->in_degree, g->inverse_vertices->out_degree,
                    (size_t)g->num_vertices * 4);
        std::memcpy(g->inverse_vertices->in_degree, g->vertices->out_degree,
                    (size_t)g->num_vertices * 4);
    }
#endif
    return g;
}

struct GraphCSR *graphCSRFromEdgeList(const struct EdgeList *el, uint32_t sort)
{
    if (!el) return nullptr;
    ag_ctx *c = ctx_or_null("graphCSRFromEdgeList");
    if (!c) return nullptr;
    ag_graph *dg = nullptr;
    if (ag_csr_build(c, el->edges_array_src, el->edges_array_dest, el->num_edges, el->num_vertices,
                     sort & 0xffu, &dg) != AG_OK) {
        report("graphCSRFromEdgeList");
        return nullptr;
    }
    const uint32_t N = ag_graph_num_vertices(dg), E = el->num_edges;
    struct GraphCSR *g = graphCSRNew(N, E, 1);
    struct EdgeList *fw = newEdgeList(E);
    struct EdgeList *iv = newEdgeList(E);
    if (!g || !fw || !iv) {
        ag_graph_free(dg);
        graphCSRFree(g);
        freeEdgeList(fw);
        freeEdgeList(iv);
        return nullptr;
    }
    fw->num_vertices = iv->num_vertices = N;
    fw->avg_degree = iv->avg_degree = N ? E / N : 0;
    if (ag_graph_download(c, dg, g->vertices->out_degree, g->vertices->in_degree,
                          g->vertices->edges_idx, fw->edges_array_src, fw->edges_array_dest,
                          g->inverse_vertices->out_degree, g->inverse_vertices->in_degree,
                          g->inverse_vertices->edges_idx, iv->edges_array_src,
                          iv->edges_array_dest) != AG_OK) {
        report("graphCSRFromEdgeList");
        ag_graph_free(dg);
        graphCSRFree(g);
        freeEdgeList(fw);
        freeEdgeList(iv);
        return nullptr;
    }
    g->sorted_edges_array = fw;
    g->inverse_sorted_edges_array = iv;
    g->ag_device = dg;
    return g;
}

} // extern "C"

namespace {

// Mode 11 of the relabelling options: labels from a text file, one unsigned
// integer per vertex, vertex v's new label on line v (the format of the
// reference's 04_test_graphs/TEST/graphbrew/graph_Rabbit.org.labels).
bool read_label_file(const char *path, uint32_t N, std::vector<uint32_t> &lab)
{
    FILE *f = path ? std::fopen(path, "r") : nullptr;
    if (!f) {
        std::fprintf(stderr, "[accelgraph-amd] relabelling: cannot open label file %s\n",
                     path ? path : "(null: arguments->fnamel)");
        return false;
    }
    lab.assign(N, 0);
    uint32_t n = 0;
    unsigned long x;
    while (n < N && std::fscanf(f, "%lu", &x) == 1) lab[n++] = (uint32_t)x;
    std::fclose(f);
    if (n != N) {
        std::fprintf(stderr, "[accelgraph-amd] relabelling: %s holds %u labels, the graph %u vertices\n",
                     path, n, N);
        return false;
    }
    return true;
}

// The -l / -L / -O layers (Arguments.lmode, lmode_l2, lmode_l3): each layer
// relabels the edge list left by the previous one; `total` composes them
// (total[v] = label of original vertex v after every layer).
bool apply_relabel_layers(ag_ctx *c, const struct Arguments *a, struct EdgeList *el,
                          std::vector<uint32_t> &total)
{
    const uint32_t N = el->num_vertices, E = el->num_edges;
    total.resize(N);
    for (uint32_t v = 0; v < N; v++) total[v] = v;
    const uint32_t modes[3] = {a->lmode, a->lmode_l2, a->lmode_l3};
    for (uint32_t m : modes) {
        if (m == 0) continue;
        std::vector<uint32_t> lab(N);
        if (m == 11) {
            if (!read_label_file(a->fnamel, N, lab)) return false;
        } else if (ag_reorder_labels(c, el->edges_array_src, el->edges_array_dest, E, N, m,
                                     kRelabelSeed, lab.data()) != AG_OK) {
            report("graphCSRPreProcessingStep (relabel)");
            return false;
        }
        if (ag_relabel_edges(c, lab.data(), N, el->edges_array_src, el->edges_array_dest, E) !=
            AG_OK) {
            report("graphCSRPreProcessingStep (relabel)");
            return false;
        }
        for (uint32_t v = 0; v < N; v++) total[v] = lab[total[v]];
    }
    return true;
}

bool set_labels(struct EdgeList *el, const std::vector<uint32_t> &total)
{
    const size_t N = total.size();
    el->label_array = (uint32_t *)std::malloc(std::max<size_t>(N, 1) * 4);
    el->inverse_label_array = (uint32_t *)std::malloc(std::max<size_t>(N, 1) * 4);
    if (!el->label_array || !el->inverse_label_array) return false;
    for (size_t v = 0; v < N; v++) {
        el->label_array[v] = total[v];
        el->inverse_label_array[total[v]] = (uint32_t)v;
    }
    return true;
}

} // namespace

extern "C" {

// Reads the edge list, applies the relabelling layers, builds both CSR
// directions on the GPU, and fills label_array (original id -> label) and
// inverse_label_array on both sorted edge lists (identity without layers).
struct GraphCSR *graphCSRPreProcessingStep(struct Arguments *arguments)
{
    if (!arguments || !arguments->fnameb) return nullptr;
    ag_ctx *c = ctx_or_null("graphCSRPreProcessingStep");
    if (!c) return nullptr;
    struct EdgeList *el = readEdgeListsFile(arguments->fnameb, arguments->fnameb_format, 0,
                                            arguments->symmetric);
    if (!el) {
        report("graphCSRPreProcessingStep");
        return nullptr;
    }
    std::vector<uint32_t> total;
    if (!apply_relabel_layers(c, arguments, el, total)) {
        freeEdgeList(el);
        return nullptr;
    }
    struct GraphCSR *g = graphCSRFromEdgeList(el, arguments->sort);
    freeEdgeList(el);
    if (g && total.size() == g->num_vertices &&
        !(set_labels(g->sorted_edges_array, total) &&
          set_labels(g->inverse_sorted_edges_array, total))) {
        graphCSRFree(g);
        return nullptr;
    }
    return g;
}

void graphCSRFree(struct GraphCSR *g)
{
    if (!g) return;
    drop_device(g);
    freeVertexArray(g->vertices);
    freeEdgeList(g->sorted_edges_array);
#if DIRECTED
    freeVertexArray(g->inverse_vertices);
    freeEdgeList(g->inverse_sorted_edges_array);
#endif
    std::free(g);
}

void graphCSRPrint(const struct GraphCSR *g)
{
    if (!g) return;
    std::printf(" -----------------------------------------------------\n");
    std::printf("| %-51s | \n", "GraphCSR (MI355X device-resident)");
    std::printf(" -----------------------------------------------------\n");
    std::printf("| %-25s | %-23u | \n", "Number of Vertices (V)", g->num_vertices);
    std::printf("| %-25s | %-23u | \n", "Number of Edges (E)", g->num_edges);
    std::printf("| %-25s | %-23u | \n", "Average Degree (D)", g->avg_degree);
    std::printf(" -----------------------------------------------------\n");
}

// newPageRankStatsGraphCSR (pageRank.c:50-77): damp, base_pr = 1 - damp,
// pageRanks = base_pr, realRanks = identity.
struct PageRankStats *newPageRankStatsGraphCSR(struct GraphCSR *graph)
{
    if (!graph) return nullptr;
    struct PageRankStats *st = (struct PageRankStats *)std::calloc(1, sizeof(struct PageRankStats));
    if (!st) return nullptr;
    st->damp = Damp;
    st->base_pr = 1.0f - st->damp;
    st->num_vertices = graph->num_vertices;
    const size_t n = (size_t)std::max<uint32_t>(graph->num_vertices, 1);
    st->realRanks = (uint32_t *)std::malloc(n * 4);
    st->pageRanks = (float *)std::malloc(n * 4);
    if (!st->realRanks || !st->pageRanks) {
        freePageRankStats(st);
        return nullptr;
    }
    for (uint32_t v = 0; v < graph->num_vertices; v++) {
        st->pageRanks[v] = st->base_pr;
        st->realRanks[v] = v;
    }
    return st;
}

void freePageRankStats(struct PageRankStats *st)
{
    if (!st) return;
    std::free(st->realRanks);
    std::free(st->pageRanks);
    std::free(st);
}

struct PageRankStats *pageRankGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    // pageRank.c:874-938: case 6 is the data-driven pull; case 0 and
    // `default` are the pull kernel.  The push, fixed-point and quantised
    // variants are outside this library and fall through to pull like the
    // reference's default branch.
    if (arguments && arguments->pushpull == 6) return pageRankDataDrivenPullGraphCSR(arguments, graph);
    return pageRankPullGraphCSR(arguments, graph);
}

} // extern "C"

namespace {

// The device copy of a host GraphCSR, uploaded on first use and kept in
// graph->ag_device (released by graphCSRFree).  The gathers need the forward
// out-degree and the inverse (DIRECTED, graphConfig.h) degree / index / dest
// arrays, as the reference's WED hands the compute units (capienv.c:396-438).
ag_graph *resident(ag_ctx *c, struct GraphCSR *graph, const char *fn)
{
    ag_graph *dg = (ag_graph *)graph->ag_device;
    if (dg) return dg;
#if DIRECTED
    struct Vertex *iv = graph->inverse_vertices;
    struct EdgeList *ie = graph->inverse_sorted_edges_array;
#else
    struct Vertex *iv = graph->vertices;
    struct EdgeList *ie = graph->sorted_edges_array;
#endif
    if (!iv || !ie || !graph->vertices) {
        std::fprintf(stderr, "[accelgraph-amd] %s: graph has no inverse CSR\n", fn);
        return nullptr;
    }
    if (ag_graph_from_csr(c, graph->num_vertices, graph->num_edges, graph->vertices->out_degree,
                          iv->out_degree, iv->edges_idx, ie->edges_array_dest, &dg) != AG_OK) {
        report(fn);
        return nullptr;
    }
    graph->ag_device = dg;
    return dg;
}

// Shared body of the pull entry points: time_total spans the device copy of
// the CSR (when not resident yet), the gather-layout build and the sweep loop,
// as the reference's spans WED mapping, AFU setup and loop.
struct PageRankStats *run_pull(struct Arguments *arguments, struct GraphCSR *graph,
                               uint32_t mode, const char *fn, const char *title)
{
    if (!arguments || !graph) return nullptr;
    ag_ctx *c = ctx_or_null(fn);
    if (!c) return nullptr;
    struct PageRankStats *st = newPageRankStatsGraphCSR(graph);
    if (!st) return nullptr;
    const bool q = quiet();
    if (!q) {
        std::printf(" -----------------------------------------------------\n");
        std::printf("| %-51s | \n", "          ---->>> MI355X (gfx950) <<<----");
        std::printf(" -----------------------------------------------------\n");
        std::printf("| %-51s | \n", title);
        std::printf(" -----------------------------------------------------\n");
        std::printf("| %-51.13lf | \n", arguments->epsilon);
        std::printf(" -----------------------------------------------------\n");
        std::printf("| %-10s | %-8s | %-15s | %-9s | \n", "Iteration", "Active", "Error", "Time (S)");
        std::printf(" -----------------------------------------------------\n");
    }
    const double t0 = now_s();
    ag_graph *dg = resident(c, graph, fn);
    if (!dg) {
        freePageRankStats(st);
        return nullptr;
    }
    ag_pr_options o;
    ag_pr_options_default(&o);
    o.iterations = arguments->iterations;
    o.epsilon = arguments->epsilon;
    o.mode = mode;
    o.precision = AG_PREC_FP32;
    std::vector<uint32_t> act(std::max<uint32_t>(o.iterations, 1));
    std::vector<double> err(std::max<uint32_t>(o.iterations, 1));
    ag_pr_result r;
    if (ag_pagerank_pull(c, dg, &o, st->pageRanks, act.data(), err.data(), &r) != AG_OK) {
        report(fn);
        freePageRankStats(st);
        return nullptr;
    }
    const double t1 = now_s();
    st->iterations = r.iterations;
    st->activeVertices = r.active_vertices;
    st->error_total = r.error_total;
    st->time_total = t1 - t0;
    if (!q) {
        const double per = r.sweeps ? r.loop_ms / 1e3 / r.sweeps : 0.0;
        for (uint32_t i = 0; i < r.sweeps; i++)
            std::printf("| %-10u | %-8u | %-15.13lf | %-9f | \n", i, act[i], err[i], per);
        std::printf(" -----------------------------------------------------\n");
        std::printf("| %-10s | %-8s | %-15s | %-9s | \n", "Iterations", "PR Sum", "Error", "Time (S)");
        std::printf(" -----------------------------------------------------\n");
        std::printf("| %-10u | %-8lf | %-15.13lf | %-9f | \n", st->iterations, r.pr_sum,
                    st->error_total, st->time_total);
        std::printf(" -----------------------------------------------------\n");
    }
    return st;
}

} // namespace

extern "C" {

// pageRankPullGraphCSR (pageRank.c:942-1108) on the GPU.
struct PageRankStats *pageRankPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_pull(arguments, graph, AG_PR_REFERENCE, "pageRankPullGraphCSR",
                    "Starting Page Rank Pull (tolerance/epsilon)");
}

// pageRankDataDrivenPullGraphCSR (pageRank.c:2810-2968) on the GPU: the
// worklist pull.  Same table and PageRankStats; activeVertices is the number
// of vertices that took a new rank in the last sweep.
struct PageRankStats *pageRankDataDrivenPullGraphCSR(struct Arguments *arguments,
                                                     struct GraphCSR *graph)
{
    return run_pull(arguments, graph, AG_PR_DATA_DRIVEN, "pageRankDataDrivenPullGraphCSR",
                    "Starting Page Rank Pull DD (tolerance/epsilon)");
}

// newSPMVStatsGraphCSR (SPMV.c:51-73): both vectors zeroed.
struct SPMVStats *newSPMVStatsGraphCSR(struct GraphCSR *graph)
{
    if (!graph) return nullptr;
    struct SPMVStats *st = (struct SPMVStats *)std::calloc(1, sizeof(struct SPMVStats));
    if (!st) return nullptr;
    st->num_vertices = graph->num_vertices;
    const size_t n = (size_t)std::max<uint32_t>(graph->num_vertices, 1);
    st->vector_output = (float *)std::calloc(n, 4);
    st->vector_input = (float *)std::calloc(n, 4);
    if (!st->vector_output || !st->vector_input) {
        freeSPMVStats(st);
        return nullptr;
    }
    return st;
}

void freeSPMVStats(struct SPMVStats *st)
{
    if (!st) return;
    std::free(st->vector_output);
    std::free(st->vector_input);
    std::free(st);
}

struct SPMVStats *SPMVGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    // SPMV.c:596-620: case 0 and `default` are the pull kernel; push and
    // fixed point are outside this library and fall through to pull.
    return SPMVPullGraphCSR(arguments, graph);
}

// SPMVPullGraphCSR (SPMV.c:622-750) on the GPU.  time_total spans the device
// copy of the CSR (when not resident yet), the gather-layout build and the
// `iterations` products, as the reference's spans AFU setup and its loop.
struct SPMVStats *SPMVPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    if (!arguments || !graph) return nullptr;
    ag_ctx *c = ctx_or_null("SPMVPullGraphCSR");
    if (!c) return nullptr;
    struct SPMVStats *st = newSPMVStatsGraphCSR(graph);
    if (!st) return nullptr;
    const bool q = quiet();
    if (!q) {
        std::printf(" -----------------------------------------------------\n");
        std::printf("| %-51s | \n", "          ---->>> MI355X (gfx950) <<<----");
        std::printf(" -----------------------------------------------------\n");
        std::printf("| %-51s | \n", "Starting SPMV-PULL");
        std::printf(" -----------------------------------------------------\n");
        std::printf("| %-21s | %-27s | \n", "Iteration", "Time (S)");
        std::printf(" -----------------------------------------------------\n");
    }
    // benchmark input vector (SPMV.c:699-706)
    for (uint32_t v = 0; v < graph->num_vertices; v++) {
        if (graph->vertices->out_degree[v])
            st->vector_input[v] = (1.0f / graph->vertices->out_degree[v]);
        else
            st->vector_input[v] = 0.001f;
    }
    const double t0 = now_s();
    ag_graph *dg = resident(c, graph, "SPMVPullGraphCSR");
    if (!dg) {
        freeSPMVStats(st);
        return nullptr;
    }
    ag_spmv_result r;
    if (ag_spmv_pull(c, dg, st->vector_input, st->vector_output, arguments->iterations, &r) !=
        AG_OK) {
        report("SPMVPullGraphCSR");
        freeSPMVStats(st);
        return nullptr;
    }
    st->iterations = arguments->iterations;
    st->time_total = now_s() - t0;
    if (!q) {
        const double per = r.iterations ? r.loop_ms / 1e3 / r.iterations : 0.0;
        for (uint32_t i = 0; i < st->iterations; i++)
            std::printf("| %-21u | %-27f | \n", i, per);
        double sum = 0.0; // SPMV.c:735-739
        for (uint32_t v = 0; v < graph->num_vertices; v++)
            sum += ((int)(st->vector_output[v] * 100 + .5) / 100.0);
        std::printf(" -----------------------------------------------------\n");
        std::printf("| %-15s | %-15s | %-15s | \n", "Iterations", "Sum", "Time (S)");
        std::printf(" -----------------------------------------------------\n");
        std::printf("| %-15u | %-15lf | %-15f | \n", st->iterations, sum, st->time_total);
        std::printf(" -----------------------------------------------------\n");
    }
    return st;
}

void pageRankPrint(float *pr, uint32_t n)
{
    for (uint32_t v = 0; v < n; v++) std::printf("Rank[%u]=%f \n", v, pr[v]);
}

} // extern "C"
