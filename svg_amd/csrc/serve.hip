// Serving loop on a device-resident SVG: continuous batching of SSSD / MSAD
// requests into shared frontier rounds (solve.h).
//
// The request slots are split into groups ("streams"), each a slot engine
// with its own HIP stream and worker thread.  Between rounds a worker admits
// queued requests into its free slots (up to max_batch per round); a slot is
// the unit of device memory, so the number of slots follows the memory
// budget.  Requests therefore join while others are mid-search, share every
// expand launch of their group, and leave as soon as their own search
// converges.  The groups' rounds run concurrently on the device: the passes
// of one round are short and latency-bound, so a second group's passes fill
// the time the first group's waves spend waiting on memory, and each round
// carries only its own group's frontier.  The reference answers one query at
// a time on the host (SVG.cu:369-413; kernel.cu:63-92).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <system_error>
#include <thread>
#include <unordered_map>
#include <vector>

#include "solve.h"
#include "svg_amd.h"
#include "svg_capi_internal.h"

using namespace svg;
using Clock = std::chrono::steady_clock;

namespace {

struct Ticket {
    int kind = Q_SSSD;
    int s = 0, t = 0;
    std::vector<int> src;
    std::vector<double> w;
    double* distOut = nullptr;
    int32_t* srcidOut = nullptr;
    Clock::time_point submit, done;
    int state = 0;           // 0 queued, 1 running, 2 done, 3 failed
    long long id = -1;
    svg_done_fn cb = nullptr;   // completion callback (the ticket is then never waited for)
    void* user = nullptr;
};

// A callback due once the server mutex is released.
struct DoneCall {
    svg_done_fn cb;
    void* user;
    long long id;
    int status;
    double ms;
};

void run_calls(const std::vector<DoneCall>& calls)
{
    for (const DoneCall& c : calls) c.cb(c.id, c.status, c.ms, c.user);
}

// One slot group: a slot engine, its stream and worker, and the pinned
// buffers its results leave through.
struct Group {
    SolverState st;
    int slots = 0;
    std::thread worker;
    std::vector<Ticket*> slotTicket;      // guarded by the server mutex
    int running = 0;                      // slots searching (guarded by the server mutex)
    std::atomic<long long> launches{0};
    // pipelining: two batches of rounds in flight, their slot records copied
    // into alternating pinned buffers; finished MSAD results leave through a
    // second stream into per-slot pinned buffers
    SlotParam* hs[2] = {nullptr, nullptr};
    hipEvent_t batchEv[2] = {nullptr, nullptr};
    hipStream_t stream = nullptr;         // admissions, rounds, slot records
    hipStream_t fetchStream = nullptr;
    std::vector<hipEvent_t> fetchEv;
    double* fDist = nullptr;              // slots * V (pinned)
    int32_t* fSrc = nullptr;              // slots * V (pinned)
};

}  // namespace

struct svg_server_s {
    svg_graph_s* g = nullptr;
    std::vector<std::unique_ptr<Group>> groups;
    int slots = 0;
    size_t slotBytes = 0;
    int maxBatch = 0;
    size_t queueCap = 0;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<Ticket*> queue;
    // Live tickets by id.  A ticket is freed once its result is consumed:
    // by svg_server_wait (or replay), or right after its callback ran.
    std::unordered_map<long long, std::unique_ptr<Ticket>> tickets;
    long long nextTicket = 0;
    bool stopping = false;
    std::string error;
    // stats
    std::vector<double> lat;
    long long refused = 0;
    int maxBatchSeen = 0;
    Clock::time_point firstSubmit, lastDone;
    bool anySubmit = false;
    long long nKind[2] = {0, 0}, roundsKind[2] = {0, 0}, settledKind[2] = {0, 0}, relaxedKind[2] = {0, 0};
};

namespace {

enum : int { SLOT_FREE = 0, SLOT_RUNNING = 1, SLOT_FETCHING = 2 };

// Marks T failed (3) or done (2); a ticket with a callback is freed here and
// its call queued for after the mutex is released.  Caller holds S->mu.
void settle(svg_server_s* S, Ticket* T, int state, std::vector<DoneCall>& calls)
{
    T->state = state;
    if (!T->cb) return;
    const double ms = state == 2 ? std::chrono::duration<double, std::milli>(T->done - T->submit).count() : -1.0;
    calls.push_back(DoneCall{T->cb, T->user, T->id, state == 2 ? SVG_OK : SVG_E_HIP, ms});
    S->tickets.erase(T->id);
}

void fail_all(svg_server_s* S, const std::string& e)
{
    std::vector<DoneCall> calls;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->error.empty()) S->error = e;
        for (auto& G : S->groups)
            for (Ticket*& T : G->slotTicket)
                if (T) { settle(S, T, 3, calls); T = nullptr; }
        for (Ticket* T : S->queue) settle(S, T, 3, calls);
        S->queue.clear();
        S->cvDone.notify_all();
    }
    run_calls(calls);
}

void finish_ticket(svg_server_s* S, Group& G, int q, const SlotParam& sp)
{
    solver_release_slot(G.st, q);
    const Clock::time_point now = Clock::now();
    std::vector<DoneCall> calls;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        Ticket* T = G.slotTicket[q];
        if (!T) return;                       // failed meanwhile
        const int kd = T->kind == Q_MSAD ? 1 : 0;
        S->nKind[kd] += 1;
        S->roundsKind[kd] += sp.rounds;
        S->settledKind[kd] += sp.settled;
        S->relaxedKind[kd] += sp.relaxed;
        T->done = now;
        S->lat.push_back(std::chrono::duration<double, std::milli>(now - T->submit).count());
        S->lastDone = now;
        G.slotTicket[q] = nullptr;
        settle(S, T, 2, calls);
        S->cvDone.notify_all();
    }
    run_calls(calls);
}

// A group takes a queued request only while no other group with a free slot
// is searching fewer requests, so concurrent requests spread over the groups.
bool least_loaded(const svg_server_s* S, const Group* G, int mine)
{
    for (const auto& O : S->groups)
        if (O.get() != G && O->running < O->slots && O->running < mine) return false;
    return true;
}

}  // namespace

// The worker keeps its stream busy: it queues batch k of rounds, then reads
// the slot records of batch k-1 (complete by then) while batch k runs.  A
// request admitted between two batches joins the next one; its record is
// trusted from the first batch queued after its admission.
static void worker_body(svg_server_s* S, Group* Gp)
{
    Group& G = *Gp;
    svg_ensure_device();
    const DevMesh& m = S->g->mesh->d;
    std::vector<int> state((size_t)G.slots, SLOT_FREE);
    std::vector<long long> admitBatch((size_t)G.slots, 0);
    std::vector<SlotParam> fetchRec((size_t)G.slots);
    long long k = 0;                 // batches queued so far
    long long nextRead = 0;          // batches before this one have been read
    const size_t V = (size_t)G.st.V;
    for (;;) {
        // results that have arrived in pinned memory (a slot's ticket is only
        // changed by this worker, or cleared by fail_all)
        for (int q = 0; q < G.slots; ++q) {
            if (state[q] != SLOT_FETCHING || hipEventQuery(G.fetchEv[q]) != hipSuccess) continue;
            Ticket* T;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                T = G.slotTicket[q];
            }
            if (T && T->distOut) std::memcpy(T->distOut, G.fDist + q * V, sizeof(double) * V);
            if (T && T->srcidOut) std::memcpy(T->srcidOut, G.fSrc + q * V, sizeof(int32_t) * V);
            finish_ticket(S, G, q, fetchRec[q]);
            state[q] = SLOT_FREE;
        }
        // admissions into free slots
        std::vector<Admission> adm;
        int running = 0, fetching = 0;
        for (int q = 0; q < G.slots; ++q) { running += state[q] == SLOT_RUNNING; fetching += state[q] == SLOT_FETCHING; }
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (!S->error.empty()) return;
            G.running = running;
            if (running == 0 && fetching == 0 && nextRead == k && (S->queue.empty() || !least_loaded(S, &G, 0))) {
                if (S->stopping && S->queue.empty()) break;
                S->cvWork.wait_for(lk, std::chrono::milliseconds(2));
                continue;
            }
            for (int q = 0; q < G.slots && !S->queue.empty() && (int)adm.size() < S->maxBatch; ++q) {
                if (state[q] != SLOT_FREE) continue;
                if (!least_loaded(S, &G, running + (int)adm.size())) break;
                Ticket* T = S->queue.front();
                S->queue.pop_front();
                Admission a;
                a.slot = q; a.kind = T->kind; a.target = T->kind == Q_SSSD ? T->t : -1;
                a.wantParent = T->kind == Q_MSAD && T->srcidOut != nullptr;
                if (T->kind == Q_SSSD) { a.src.push_back(T->s); a.w.push_back(0.0); }
                else { a.src = T->src; a.w = T->w; }
                T->state = 1;
                G.slotTicket[q] = T;
                adm.push_back(std::move(a));
            }
            G.running = running + (int)adm.size();
            int total = 0;
            for (const auto& O : S->groups) total += O->running;
            S->maxBatchSeen = std::max(S->maxBatchSeen, total);
        }
        if (!adm.empty()) {
            std::string e = solver_admit_many(G.st, m, adm);
            if (!e.empty()) { fail_all(S, e); return; }
            for (const Admission& a : adm) { state[a.slot] = SLOT_RUNNING; admitBatch[a.slot] = k; }
            running += (int)adm.size();
        }
        // queue the next batch while the previous one may still run
        if (running > 0) {
            std::string e = solver_enqueue_rounds(G.st, S->g->d, m, G.st.roundsPerSync);
            if (e.empty()) e = solver_enqueue_status(G.st, G.hs[k & 1]);
            if (e.empty() && hipEventRecord(G.batchEv[k & 1], G.stream) != hipSuccess) e = "event record failed";
            if (!e.empty()) { fail_all(S, e); return; }
            G.launches.store(G.st.launches, std::memory_order_relaxed);
            ++k;
        }
        // read the records of the batch before the one just queued (or of the
        // last batch, when nothing new was queued)
        const long long upto = running > 0 ? k - 2 : k - 1;
        for (; nextRead <= upto; ++nextRead) {
            const int pb = (int)(nextRead & 1);
            if (hipEventSynchronize(G.batchEv[pb]) != hipSuccess) { fail_all(S, "batch wait failed"); return; }
            const SlotParam* rec = G.hs[pb];
            if (rec[G.slots].kind) { fail_all(S, "solver frontier overflow"); return; }
            for (int q = 0; q < G.slots; ++q) {
                if (state[q] != SLOT_RUNNING || admitBatch[q] > nextRead || !rec[q].done) continue;
                Ticket* T;
                {
                    std::lock_guard<std::mutex> lk(S->mu);
                    T = G.slotTicket[q];
                }
                if (!T) return;                   // failed meanwhile
                if (T->kind == Q_SSSD) {
                    if (T->distOut) *T->distOut = rec[q].result;
                    finish_ticket(S, G, q, rec[q]);
                    state[q] = SLOT_FREE;
                } else {
                    std::string e;
                    if (hipStreamWaitEvent(G.fetchStream, G.batchEv[pb], 0) != hipSuccess) e = "stream wait failed";
                    if (e.empty())
                        e = solver_enqueue_fetch(G.st, q, G.fetchStream, T->distOut ? G.fDist + q * V : nullptr,
                                                 T->srcidOut ? G.fSrc + q * V : nullptr, nullptr);
                    if (e.empty() && hipEventRecord(G.fetchEv[q], G.fetchStream) != hipSuccess)
                        e = "event record failed";
                    if (!e.empty()) { fail_all(S, e); return; }
                    fetchRec[q] = rec[q];
                    state[q] = SLOT_FETCHING;
                }
            }
        }
        // only results in flight: wait for the oldest rather than spin
        if (running == 0 && nextRead == k && fetching > 0) {
            for (int q = 0; q < G.slots; ++q)
                if (state[q] == SLOT_FETCHING) { hipEventSynchronize(G.fetchEv[q]); break; }
        }
    }
}

// A worker never lets an exception end the process: it fails the server's
// requests (their waiters and callbacks see the error) and stops.
static void worker_loop(svg_server_s* S, Group* Gp)
{
    try {
        worker_body(S, Gp);
    } catch (const std::exception& e) {
        try { fail_all(S, std::string("server worker: ") + e.what()); } catch (...) {}
    } catch (...) {
        try { fail_all(S, "server worker failed"); } catch (...) {}
    }
}

static void group_free(Group& G)
{
    solver_destroy(G.st);
    for (int i = 0; i < 2; ++i) {
        if (G.hs[i]) hipHostFree(G.hs[i]);
        if (G.batchEv[i]) hipEventDestroy(G.batchEv[i]);
    }
    for (hipEvent_t e : G.fetchEv) if (e) hipEventDestroy(e);
    if (G.fetchStream) hipStreamDestroy(G.fetchStream);
    if (G.stream) hipStreamDestroy(G.stream);
    if (G.fDist) hipHostFree(G.fDist);
    if (G.fSrc) hipHostFree(G.fSrc);
}

static void server_free(svg_server_s* S)
{
    for (auto& G : S->groups) group_free(*G);
    delete S;
}

extern "C" int svg_server_start(svg_graph_t g, const svg_server_cfg* cfg, svg_server_t* out)
{
    return svg_guarded<int>([&]() -> int {
        if (!g || !out) return svg_set_error(SVG_E_ARG, "null argument");
        int rc = svg_ensure_device();
        if (rc) return rc;
        svg_server_s* S = new (std::nothrow) svg_server_s();
        if (!S) return svg_set_error(SVG_E_MEM, "host allocation failed");
        S->g = g;
        int slots = cfg && cfg->max_inflight > 0 ? cfg->max_inflight : 64;
        const size_t perSlot = solver_bytes_per_slot(g->d.V);
        size_t freeB = 0, totB = 0;
        if (hipMemGetInfo(&freeB, &totB) != hipSuccess) { delete S; return svg_set_error(SVG_E_HIP, "hipMemGetInfo failed"); }
        size_t budget = cfg && cfg->mem_budget > 0 ? (size_t)cfg->mem_budget : freeB / 4;
        const int memSlots = (int)std::min<size_t>(4096, budget / perSlot);
        if (memSlots < 1) { delete S; return svg_set_error(SVG_E_MEM, "memory budget below one request slot"); }
        slots = std::min(slots, memSlots);
        const int nGroups = std::max(1, std::min(cfg && cfg->streams > 0 ? cfg->streams : 4, std::min(slots, 8)));
        S->slots = slots;
        S->slotBytes = perSlot;
        S->maxBatch = cfg && cfg->max_batch > 0 ? cfg->max_batch : slots;
        S->queueCap = cfg && cfg->queue_cap > 0 ? (size_t)cfg->queue_cap : (size_t)std::max(1024, 64 * slots);
        const double scale = cfg && cfg->delta_scale > 0.0 ? cfg->delta_scale : 1.0;
        const double delta = graph_mean_arc(g->d) * scale;
        for (int gi = 0; gi < nGroups; ++gi) {
            S->groups.emplace_back(new Group());
            Group& G = *S->groups.back();
            G.slots = slots / nGroups + (gi < slots % nGroups ? 1 : 0);
            std::string e;
            if (hipStreamCreateWithFlags(&G.stream, hipStreamNonBlocking) != hipSuccess) e = "stream creation failed";
            G.st.stream = G.stream;
            if (e.empty()) e = solver_create(g->d, g->mesh->d, G.slots, delta, G.st);
            if (!e.empty()) { server_free(S); return svg_set_error(SVG_E_MEM, e); }
            if (cfg && cfg->rounds_per_sync > 0) G.st.roundsPerSync = std::min(cfg->rounds_per_sync, 64);
            // the groups' expand passes share the device: together about two
            // machine-fills of waves
            G.st.expandBlocks = std::max(192, std::min(768, 1536 / nGroups));
            if (cfg && cfg->chain_rows != 0) G.st.chainBudget = std::max(0, std::min(cfg->chain_rows, 64));
            if (cfg && cfg->msad_delta_scale > 0.0) G.st.deltaMsad = delta / scale * cfg->msad_delta_scale;
            bool ok = true;
            for (int i = 0; i < 2 && ok; ++i) {
                ok = hipHostMalloc(&G.hs[i], sizeof(SlotParam) * (size_t)(G.slots + 1), hipHostMallocDefault) == hipSuccess &&
                     hipEventCreateWithFlags(&G.batchEv[i], hipEventDisableTiming) == hipSuccess;
            }
            ok = ok && hipStreamCreateWithFlags(&G.fetchStream, hipStreamNonBlocking) == hipSuccess;
            G.fetchEv.assign((size_t)G.slots, nullptr);
            for (int q = 0; q < G.slots && ok; ++q)
                ok = hipEventCreateWithFlags(&G.fetchEv[q], hipEventDisableTiming) == hipSuccess;
            ok = ok && hipHostMalloc(&G.fDist, sizeof(double) * (size_t)G.slots * g->d.V, hipHostMallocDefault) == hipSuccess &&
                 hipHostMalloc(&G.fSrc, sizeof(int32_t) * (size_t)G.slots * g->d.V, hipHostMallocDefault) == hipSuccess;
            if (!ok) {
                server_free(S);
                return svg_set_error(SVG_E_MEM, "server staging allocation failed");
            }
            G.slotTicket.assign((size_t)G.slots, nullptr);
        }
        try {
            for (auto& G : S->groups) G->worker = std::thread(worker_loop, S, G.get());
        } catch (const std::system_error& e) {
            // the workers already started end before the server is freed
            {
                std::lock_guard<std::mutex> lk(S->mu);
                S->stopping = true;
            }
            S->cvWork.notify_all();
            for (auto& G : S->groups)
                if (G->worker.joinable()) G->worker.join();
            server_free(S);
            return svg_set_error(SVG_E_INTERNAL, std::string("server worker start: ") + e.what());
        }
        *out = S;
        return SVG_OK;
    });
}

static long long submit_impl(svg_server_t S, const svg_request* r, svg_done_fn cb, void* user)
{
    if (!S || !r) return svg_set_error(SVG_E_ARG, "null argument");
    const int V = S->g->d.V;
    std::unique_ptr<Ticket> T(new Ticket());
    T->kind = r->kind == 1 ? Q_MSAD : Q_SSSD;
    if (T->kind == Q_SSSD) {
        if (r->s < 0 || r->t < 0 || r->s >= V || r->t >= V) return svg_set_error(SVG_E_ARG, "vertex out of range");
        T->s = r->s; T->t = r->t;
    } else {
        if (!r->src || !r->w || r->ns <= 0) return svg_set_error(SVG_E_ARG, "MSAD needs sources and weights");
        for (int i = 0; i < r->ns; ++i) {
            if (r->src[i] < 0 || r->src[i] >= V) return svg_set_error(SVG_E_ARG, "source out of range");
            if (!std::isfinite(r->w[i]) || std::signbit(r->w[i])) return svg_set_error(SVG_E_ARG, kBadWeight);
        }
        T->src.assign(r->src, r->src + r->ns);
        T->w.assign(r->w, r->w + r->ns);
    }
    T->distOut = r->dist_out;
    T->srcidOut = r->srcid_out;
    T->cb = cb;
    T->user = user;
    T->submit = Clock::now();
    std::lock_guard<std::mutex> lk(S->mu);
    if (!S->error.empty()) return svg_set_error(SVG_E_HIP, "server failed: " + S->error);
    if (S->queue.size() >= S->queueCap) { S->refused++; return svg_set_error(SVG_E_BUSY, "queue full"); }
    if (!S->anySubmit) { S->firstSubmit = T->submit; S->anySubmit = true; }
    const long long id = S->nextTicket++;
    T->id = id;
    S->queue.push_back(T.get());
    S->tickets.emplace(id, std::move(T));
    S->cvWork.notify_all();
    return id;
}

extern "C" long long svg_server_submit(svg_server_t S, const svg_request* r)
{
    return svg_guarded<long long>([&]() -> long long {
        return submit_impl(S, r, nullptr, nullptr);
    });
}

extern "C" long long svg_server_submit_cb(svg_server_t S, const svg_request* r, svg_done_fn cb, void* user)
{
    return svg_guarded<long long>([&]() -> long long {
        if (!cb) return svg_set_error(SVG_E_ARG, "null callback");
        return submit_impl(S, r, cb, user);
    });
}

// Waits for a ticket without a callback and frees it; *done receives its
// completion time.
static int wait_ticket(svg_server_t S, long long ticket, Clock::time_point* submit, Clock::time_point* done)
{
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->tickets.find(ticket);
    if (it == S->tickets.end() || it->second->cb)
        return svg_set_error(SVG_E_ARG, "unknown ticket (never issued, already waited for, or submitted with a callback)");
    Ticket* T = it->second.get();
    S->cvDone.wait(lk, [&] { return T->state >= 2; });
    const int st = T->state;
    if (submit) *submit = T->submit;
    if (done) *done = T->done;
    S->tickets.erase(ticket);
    if (st == 3) return svg_set_error(SVG_E_HIP, "request failed: " + S->error);
    return SVG_OK;
}

extern "C" int svg_server_wait(svg_server_t S, long long ticket, double* latency_ms)
{
    return svg_guarded<int>([&]() -> int {
        if (!S) return svg_set_error(SVG_E_ARG, "null server");
        Clock::time_point a, b;
        const int rc = wait_ticket(S, ticket, &a, &b);
        if (rc) return rc;
        if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(b - a).count();
        return SVG_OK;
    });
}

static double pct(std::vector<double> v, double p)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double r = p * (double)(v.size() - 1);
    const size_t i = (size_t)r;
    const double f = r - (double)i;
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

extern "C" int svg_server_stats_get(svg_server_t S, svg_server_stats* st)
{
    return svg_guarded<int>([&]() -> int {
        if (!S || !st) return svg_set_error(SVG_E_ARG, "null argument");
        std::lock_guard<std::mutex> lk(S->mu);
        std::memset(st, 0, sizeof(*st));
        st->completed = (long long)S->lat.size();
        st->refused = S->refused;
        st->p50_ms = pct(S->lat, 0.50);
        st->p99_ms = pct(S->lat, 0.99);
        double sum = 0.0, mx = 0.0;
        for (double x : S->lat) { sum += x; mx = std::max(mx, x); }
        st->mean_ms = S->lat.empty() ? 0.0 : sum / (double)S->lat.size();
        st->max_ms = mx;
        const double span = S->anySubmit ? std::chrono::duration<double>(S->lastDone - S->firstSubmit).count() : 0.0;
        st->qps = span > 0 ? (double)S->lat.size() / span : 0.0;
        st->max_batch_seen = S->maxBatchSeen;
        st->slots = S->slots;
        st->slot_bytes = (long long)S->slotBytes;
        st->launches = 0;
        for (const auto& G : S->groups) st->launches += G->launches.load(std::memory_order_relaxed);
        for (int k = 0; k < 2; ++k) {
            st->n_kind[k] = S->nKind[k];
            st->rounds_kind[k] = S->roundsKind[k];
            st->settled_kind[k] = S->settledKind[k];
            st->relaxed_kind[k] = S->relaxedKind[k];
        }
        return SVG_OK;
    });
}

extern "C" int svg_server_stop(svg_server_t S)
{
    return svg_guarded<int>([&]() -> int {
        if (!S) return SVG_OK;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->stopping = true;
        }
        S->cvWork.notify_all();
        for (auto& G : S->groups)
            if (G->worker.joinable()) G->worker.join();
        svg_ensure_device();
        server_free(S);
        return SVG_OK;
    });
}

extern "C" int svg_server_replay(svg_server_t S, const svg_request* reqs, const double* arrival_s, int n,
                                 double* latency_ms)
{
    return svg_guarded<int>([&]() -> int {
        if (!S || (!reqs && n > 0) || (!arrival_s && n > 0)) return svg_set_error(SVG_E_ARG, "null argument");
        std::vector<long long> ids((size_t)n, -1);
        std::vector<Clock::time_point> sched((size_t)n);
        const Clock::time_point t0 = Clock::now() + std::chrono::milliseconds(5);
        for (int i = 0; i < n; ++i) {
            sched[i] = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[i]));
            std::this_thread::sleep_until(sched[i]);
            const long long id = svg_server_submit(S, &reqs[i]);
            ids[i] = id;
        }
        for (int i = 0; i < n; ++i) {
            if (ids[i] < 0) { if (latency_ms) latency_ms[i] = -1.0; continue; }
            Clock::time_point done;
            int rc = wait_ticket(S, ids[i], nullptr, &done);
            if (rc) return rc;
            if (latency_ms) latency_ms[i] = std::chrono::duration<double, std::milli>(done - sched[i]).count();
        }
        return SVG_OK;
    });
}
