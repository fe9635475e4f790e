// Serving loop for personalized-PageRank queries over several resident graphs.
//
// Continuous batching: every graph with work owns one PprBatch (column slots
// in HBM).  ag_serve_step admits queued queries into free slots (FIFO, by HBM
// budget: a batch's column state is reserved when the batch is created; a
// drained batch stays resident until another graph needs its memory), runs
// sweeps for every graph with active columns, and retires converged columns
// immediately, so queries that arrive while others are running join the next
// sweep.  Host code only; the sweeps are the
// kernels of ag_ppr.hip.  The reference has no serving path.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <mutex>
#include <unordered_map>

#include "ag_admit.h"
#include "ag_ppr.h"

namespace {

using Clock = std::chrono::steady_clock;

enum class QState { Queued, Running, Done, Failed, Cancelled };

// A finished query's result: pinned, device-mapped host memory that the GPU
// fills in vertex order; `ready` is recorded after the writing kernel.
struct ResultBuf {
    float *host = nullptr;
    float *dev = nullptr;
    uint32_t n = 0;
    hipEvent_t ready = nullptr;
};

struct Query {
    uint64_t ticket = 0;
    uint32_t graph = 0;
    std::vector<uint32_t> seeds;
    double tol = 1e-4;
    uint32_t max_iter = 1000;
    QState state = QState::Queued;
    int code = AG_OK;
    ResultBuf *res = nullptr;
    ag_ppr_stat stat{};
    Clock::time_point t_submit;
};

struct GraphEntry {
    ag_graph *g = nullptr;
    std::unique_ptr<ag::PprBatch> batch;
    uint64_t reserved = 0;
    uint32_t cpw = 0;
    uint32_t stream = 0; // index into the server's streams
};

// Streams of the server's batches: graph i's batch runs on stream i % 4, so
// the sweeps that one step queues for several graphs overlap on the device
// (a small graph's sweeps need not wait behind a large one's).  Default
// priority, like every stream of the library.
constexpr uint32_t kServeStreams = 4;

} // namespace

struct ag_server {
    ag_ctx *ctx = nullptr;
    ag_serve_config cfg{};
    std::vector<GraphEntry> graphs;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, Query> queries;
    uint64_t next_ticket = 1;
    uint64_t reserved = 0;
    uint32_t alloc_attempts = 0;
    // result buffers: all of them, and the free ones by length
    std::vector<std::unique_ptr<ResultBuf>> bufs;
    std::unordered_map<uint32_t, std::vector<ResultBuf *>> free_bufs;
    std::mutex mu;
    hipStream_t streams[kServeStreams] = {};
    ~ag_server()
    {
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        for (hipStream_t st : streams)
            if (st) {
                (void)hipStreamSynchronize(st);
                (void)hipStreamDestroy(st);
            }
        for (auto &b : bufs) {
            if (b->ready) (void)hipEventDestroy(b->ready);
            if (b->host) (void)hipHostFree(b->host);
        }
    }
};

using namespace ag;

namespace {

// Frees a graph's batch: its column buffers are returned to the device (not
// kept on the graph for a later batch), so the budget counts only memory
// that is actually held.
void release_batch(ag_server *s, GraphEntry &ge)
{
    ge.batch.reset();
    ppr_release_kept(ge.g);
    s->reserved -= ge.reserved;
    ge.reserved = 0;
}

// Creates the graph's batch with the largest capacity (multiple of 8, at most
// max_columns rounded up to 8) whose column state fits the budget.  Drained
// batches of other graphs stay resident until their memory is needed: they
// are evicted (largest first) before the capacity is reduced.  False: no room.
bool ensure_batch(ag_server *s, GraphEntry &ge)
{
    if (ge.batch) return true;
    const uint32_t want = std::min<uint32_t>(64, std::max<uint32_t>(8, (s->cfg.max_columns + 7) / 8 * 8));
    std::vector<AdmitHolder> holders;
    std::vector<GraphEntry *> who;
    for (auto &o : s->graphs)
        if (&o != &ge && o.batch) {
            holders.push_back({o.reserved, o.batch->active() == 0});
            who.push_back(&o);
        }
    const AdmitPlan plan = plan_admission(
        s->cfg.hbm_budget_bytes, s->reserved, want, holders,
        [&](uint32_t cap) { return PprBatch::bytes_for(s->ctx, ge.g, cap, s->cfg.precision); });
    if (plan.cap == 0) return false;
    for (size_t i : plan.evict) release_batch(s, *who[i]);
    const uint32_t cap = plan.cap;
    s->alloc_attempts++;
    if (s->cfg.inject_fail_alloc && s->alloc_attempts == s->cfg.inject_fail_alloc)
        return false; // injected allocation failure: the query stays queued
    // The autotune batch has the admitted capacity and the batch below takes
    // its buffers over, so tuning stays inside the reservation.
    if (!ge.cpw) ge.cpw = s->cfg.cols_per_wave
                              ? s->cfg.cols_per_wave
                              : ppr_autotune(s->ctx, ge.g, s->cfg.precision, cap, true);
    ge.batch.reset(new PprBatch(s->ctx, ge.g, cap, ge.cpw, s->cfg.precision, s->cfg.damp,
                                s->cfg.extrapolate != 0, true, s->streams[ge.stream]));
    ge.reserved = PprBatch::bytes_for(s->ctx, ge.g, cap, s->cfg.precision);
    s->reserved += ge.reserved;
    return true;
}

ResultBuf *take_buf(ag_server *s, uint32_t n)
{
    auto &fl = s->free_bufs[n];
    if (!fl.empty()) {
        ResultBuf *b = fl.back();
        fl.pop_back();
        return b;
    }
    std::unique_ptr<ResultBuf> b(new ResultBuf());
    b->n = n;
    if (hipHostMalloc((void **)&b->host, (size_t)n * 4, hipHostMallocMapped) != hipSuccess)
        fail(AG_ERR_OOM, "pinned result buffer of %u floats", n);
    AG_HIP(hipHostGetDevicePointer((void **)&b->dev, b->host, 0));
    AG_HIP(hipEventCreateWithFlags(&b->ready, hipEventDisableTiming));
    s->bufs.push_back(std::move(b));
    return s->bufs.back().get();
}

void give_buf(ag_server *s, ResultBuf *b)
{
    if (b) s->free_bufs[b->n].push_back(b);
}

void finish(Query &q, QState st, int code)
{
    q.state = st;
    q.code = code;
    q.stat.status = (uint32_t)code;
    q.stat.latency_ms =
        std::chrono::duration<double, std::milli>(Clock::now() - q.t_submit).count();
}

void admit(ag_server *s)
{
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        Query &q = s->queries.at(*it);
        if (q.state != QState::Queued) { it = s->queue.erase(it); continue; }
        if (s->cfg.inject_fail_query && q.ticket == s->cfg.inject_fail_query) {
            finish(q, QState::Failed, AG_ERR_INJECTED);
            it = s->queue.erase(it);
            continue;
        }
        GraphEntry &ge = s->graphs[q.graph];
        if (!ensure_batch(s, ge) || ge.batch->free_slots() == 0) { ++it; continue; }
        PprSeedSet set = ge.batch->make_seeds(q.seeds.data(), (uint32_t)q.seeds.size(), q.tol,
                                              q.max_iter);
        ge.batch->admit(set, q.ticket);
        q.state = QState::Running;
        it = s->queue.erase(it);
    }
}

bool graph_has_queued(ag_server *s, uint32_t gi)
{
    for (uint64_t t : s->queue) {
        auto it = s->queries.find(t);
        if (it != s->queries.end() && it->second.graph == gi && it->second.state == QState::Queued)
            return true;
    }
    return false;
}

void retire(ag_server *s, uint32_t gi)
{
    GraphEntry &ge = s->graphs[gi];
    if (!ge.batch) return;
    // results are written by the GPU straight into pinned host buffers, in
    // stream order, without waiting; poll / peek wait on the buffer's event
    auto fin = ge.batch->retire(
        [&](uint64_t tag) -> float * {
            auto qi = s->queries.find(tag);
            if (qi == s->queries.end() || qi->second.state != QState::Running) return nullptr;
            qi->second.res = take_buf(s, ge.g->N);
            return qi->second.res->dev;
        },
        !graph_has_queued(s, gi), true);
    for (auto &f : fin) {
        auto qi = s->queries.find(f.tag);
        if (qi == s->queries.end() || qi->second.state != QState::Running) continue;
        if (qi->second.res) AG_HIP(hipEventRecord(qi->second.res->ready, ge.batch->stream()));
        qi->second.stat.iterations = f.iters;
        qi->second.stat.residual = f.residual;
        qi->second.stat.fine_sweep = f.fine_at;
        finish(qi->second, QState::Done, AG_OK);
    }
}


} // namespace

extern "C" {

void ag_serve_config_default(ag_serve_config *c)
{
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->hbm_budget_bytes = 8ull << 30;
    c->max_columns = 64;
    c->precision = AG_PREC_BF16;
    c->damp = 0.85f;
}

int ag_serve_create(ag_ctx *ctx, const ag_serve_config *cfg, ag_server **out)
{
    return guard([&] {
        check_ctx(ctx);
        if (!out) fail(AG_ERR_INVALID, "null out");
        std::unique_ptr<ag_server> s(new ag_server());
        s->ctx = ctx;
        if (cfg) s->cfg = *cfg; else ag_serve_config_default(&s->cfg);
        if (!s->cfg.hbm_budget_bytes) s->cfg.hbm_budget_bytes = 8ull << 30;
        if (!s->cfg.max_columns) s->cfg.max_columns = 64;
        if (s->cfg.damp == 0.0f) s->cfg.damp = 0.85f;
        if (s->cfg.precision != AG_PREC_FP32 && s->cfg.precision != AG_PREC_BF16)
            fail(AG_ERR_INVALID, "served columns store FP32 or BF16 (precision %u)", s->cfg.precision);
        AG_HIP(hipSetDevice(ctx->device));
        for (hipStream_t &st : s->streams) AG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        *out = s.release();
    });
}

int ag_serve_destroy(ag_server *s)
{
    return guard([&] {
        if (!s) return;
        {
            // the server's column buffers go back to the device, not to the graphs
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->ctx) AG_HIP(hipSetDevice(s->ctx->device));
            for (auto &ge : s->graphs)
                if (ge.batch) release_batch(s, ge);
        }
        delete s;
    });
}

int ag_serve_add_graph(ag_server *s, ag_graph *g, uint32_t *id)
{
    return guard([&] {
        if (!s || !g) fail(AG_ERR_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        if (g->N == 0) fail(AG_ERR_INVALID, "graph has no vertices");
        ensure_plan(s->ctx, g, nullptr);
        GraphEntry ge;
        ge.g = g;
        ge.stream = (uint32_t)(s->graphs.size() % kServeStreams);
        s->graphs.push_back(std::move(ge));
        if (id) *id = (uint32_t)(s->graphs.size() - 1);
    });
}

int ag_serve_submit(ag_server *s, uint32_t gid, const uint32_t *seeds, uint32_t nseeds, double tol,
                    uint32_t max_iter, uint64_t *ticket)
{
    return guard([&] {
        PhaseRange phase_("ag_serve_submit");
        if (!s || !ticket) fail(AG_ERR_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        if (gid >= s->graphs.size()) fail(AG_ERR_NOT_FOUND, "unknown graph id %u", gid);
        if (!seeds || nseeds == 0) fail(AG_ERR_INVALID, "a query needs at least one seed");
        const uint32_t N = s->graphs[gid].g->N;
        for (uint32_t i = 0; i < nseeds; i++)
            if (seeds[i] >= N) fail(AG_ERR_INVALID, "seed %u >= num_vertices %u", seeds[i], N);
        if (!(tol > 0.0)) fail(AG_ERR_INVALID, "tolerance must be > 0");
        Query q;
        q.ticket = s->next_ticket++;
        q.graph = gid;
        q.seeds.assign(seeds, seeds + nseeds);
        q.tol = tol;
        q.max_iter = max_iter ? max_iter : 1000;
        q.t_submit = Clock::now();
        s->queue.push_back(q.ticket);
        *ticket = q.ticket;
        s->queries.emplace(q.ticket, std::move(q));
    });
}

int ag_serve_cancel(ag_server *s, uint64_t ticket)
{
    return guard([&] {
        if (!s) fail(AG_ERR_INVALID, "null server");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->queries.find(ticket);
        if (it == s->queries.end()) fail(AG_ERR_NOT_FOUND, "unknown ticket");
        Query &q = it->second;
        if (q.state == QState::Running && s->graphs[q.graph].batch)
            s->graphs[q.graph].batch->cancel(ticket);
        if (q.state == QState::Queued || q.state == QState::Running)
            finish(q, QState::Cancelled, AG_ERR_CANCELLED);
    });
}

int ag_serve_step(ag_server *s, uint32_t *active)
{
    return guard([&] {
        PhaseRange phase_("ag_serve_step");
        if (!s) fail(AG_ERR_INVALID, "null server");
        std::lock_guard<std::mutex> lk(s->mu);
        check_ctx(s->ctx);
        admit(s);
        for (uint32_t gi = 0; gi < s->graphs.size(); gi++) {
            GraphEntry &ge = s->graphs[gi];
            // sweeps of all graphs are queued first, each graph's on its own
            // stream so that they overlap; retire() waits per graph.
            // Two sweeps per step keep admissions prompt; a full batch with
            // queries waiting admits nothing before its first column
            // finishes, so it sweeps until then.
            if (!ge.batch || !ge.batch->active()) continue;
            uint32_t n = 2;
            if (ge.batch->free_slots() == 0 && graph_has_queued(s, gi))
                n = std::max(2u, ge.batch->sweeps_hint());
            ge.batch->sweep(n, true, false);
        }
        for (uint32_t gi = 0; gi < s->graphs.size(); gi++) retire(s, gi);
        admit(s);
        if (active) {
            uint32_t n = 0;
            for (auto &kv : s->queries)
                if (kv.second.state == QState::Queued || kv.second.state == QState::Running) n++;
            *active = n;
        }
    });
}

int ag_serve_poll(ag_server *s, uint64_t ticket, float *x, uint32_t n, ag_ppr_stat *st)
{
    if (!s) return AG_ERR_INVALID;
    int code = AG_OK;
    int rc = guard([&] {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->queries.find(ticket);
        if (it == s->queries.end()) fail(AG_ERR_NOT_FOUND, "unknown ticket");
        Query &q = it->second;
        if (q.state == QState::Queued || q.state == QState::Running) {
            code = AG_ERR_PENDING;
            return;
        }
        if (st) *st = q.stat;
        if (q.state == QState::Done && x && q.res) {
            if (n < q.res->n) fail(AG_ERR_INVALID, "output holds %u < %u values", n, q.res->n);
            AG_HIP(hipEventSynchronize(q.res->ready));
            std::memcpy(x, q.res->host, (size_t)q.res->n * 4);
        }
        code = q.code;
        give_buf(s, q.res);
        s->queries.erase(it);
    });
    return rc != AG_OK ? rc : code;
}

int ag_serve_peek(ag_server *s, uint64_t ticket, const float **x, uint32_t *n, ag_ppr_stat *st)
{
    if (!s) return AG_ERR_INVALID;
    int code = AG_OK;
    int rc = guard([&] {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->queries.find(ticket);
        if (it == s->queries.end()) fail(AG_ERR_NOT_FOUND, "unknown ticket");
        Query &q = it->second;
        if (q.state == QState::Queued || q.state == QState::Running) {
            code = AG_ERR_PENDING;
            return;
        }
        if (st) *st = q.stat;
        if (x) *x = nullptr;
        if (n) *n = 0;
        if (q.state == QState::Done && q.res) {
            AG_HIP(hipEventSynchronize(q.res->ready));
            if (x) *x = q.res->host;
            if (n) *n = q.res->n;
        }
        code = q.code;
    });
    return rc != AG_OK ? rc : code;
}

int ag_serve_release(ag_server *s, uint64_t ticket)
{
    return guard([&] {
        if (!s) fail(AG_ERR_INVALID, "null server");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->queries.find(ticket);
        if (it == s->queries.end()) fail(AG_ERR_NOT_FOUND, "unknown ticket");
        Query &q = it->second;
        if (q.state == QState::Queued || q.state == QState::Running)
            fail(AG_ERR_PENDING, "ticket %llu has not finished", (unsigned long long)ticket);
        give_buf(s, q.res);
        s->queries.erase(it);
    });
}

uint64_t ag_serve_reserved_bytes(const ag_server *s) { return s ? s->reserved : 0; }

} // extern "C"
