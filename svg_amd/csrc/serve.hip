// Serving loop on a device-resident SVG: continuous batching of SSSD / MSAD
// requests into shared frontier rounds (solve.h).
//
// A worker thread owns the slot engine.  Between rounds it admits queued
// requests into free slots (up to max_batch per round); a slot is the unit of
// device memory, so the number of slots follows the memory budget.  Requests
// therefore join while others are mid-search, share every expand launch with
// them, and leave as soon as their own search converges.  The reference
// answers one query at a time on the host (SVG.cu:369-413; kernel.cu:63-92).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
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
};


}  // namespace

struct svg_server_s {
    svg_graph_s* g = nullptr;
    SolverState st;
    int slots = 0;
    int maxBatch = 0;
    size_t queueCap = 0;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<long long> queue;
    std::vector<std::unique_ptr<Ticket>> tickets;
    std::vector<long long> slotTicket;
    bool stopping = false;
    std::string error;
    // stats
    std::vector<double> lat;
    long long refused = 0;
    int maxBatchSeen = 0;
    Clock::time_point firstSubmit, lastDone;
    bool anySubmit = false;
    long long nKind[2] = {0, 0}, roundsKind[2] = {0, 0}, settledKind[2] = {0, 0}, relaxedKind[2] = {0, 0};
    // pipelining: two batches of rounds in flight, their slot records copied
    // into alternating pinned buffers; finished MSAD results leave through a
    // second stream into per-slot pinned buffers
    SlotParam* hs[2] = {nullptr, nullptr};
    hipEvent_t batchEv[2] = {nullptr, nullptr};
    hipStream_t fetchStream = nullptr;
    std::vector<hipEvent_t> fetchEv;
    double* fDist = nullptr;        // slots * V (pinned)
    int32_t* fSrc = nullptr;        // slots * V (pinned)
};

namespace {

enum : int { SLOT_FREE = 0, SLOT_RUNNING = 1, SLOT_FETCHING = 2 };

void fail_all(svg_server_s* S, const std::string& e)
{
    std::lock_guard<std::mutex> lk(S->mu);
    S->error = e;
    for (int q = 0; q < S->slots; ++q)
        if (S->slotTicket[q] != -1) { S->tickets[S->slotTicket[q]]->state = 3; S->slotTicket[q] = -1; }
    for (long long id : S->queue) S->tickets[id]->state = 3;
    S->queue.clear();
    S->cvDone.notify_all();
}

void finish_ticket(svg_server_s* S, int q, const SlotParam& sp)
{
    Ticket& T = *S->tickets[S->slotTicket[q]];
    const int kd = T.kind == Q_MSAD ? 1 : 0;
    S->nKind[kd] += 1;
    S->roundsKind[kd] += sp.rounds;
    S->settledKind[kd] += sp.settled;
    S->relaxedKind[kd] += sp.relaxed;
    solver_release_slot(S->st, q);
    const Clock::time_point now = Clock::now();
    std::lock_guard<std::mutex> lk(S->mu);
    T.done = now;
    T.state = 2;
    S->lat.push_back(std::chrono::duration<double, std::milli>(now - T.submit).count());
    S->lastDone = now;
    S->slotTicket[q] = -1;
    S->cvDone.notify_all();
}

}  // namespace

// The worker keeps the device busy: it queues batch k of rounds, then reads
// the slot records of batch k-1 (complete by then) while batch k runs.  A
// request admitted between two batches joins the next one; its record is
// trusted from the first batch queued after its admission.
static void worker_loop(svg_server_s* S)
{
    svg_ensure_device();
    const DevMesh& m = S->g->mesh->d;
    std::vector<int> state((size_t)S->slots, SLOT_FREE);
    std::vector<long long> admitBatch((size_t)S->slots, 0);
    std::vector<SlotParam> fetchRec((size_t)S->slots);
    long long k = 0;                 // batches queued so far
    long long nextRead = 0;          // batches before this one have been read
    const size_t V = (size_t)S->st.V;
    for (;;) {
        // results that have arrived in pinned memory
        for (int q = 0; q < S->slots; ++q) {
            if (state[q] != SLOT_FETCHING || hipEventQuery(S->fetchEv[q]) != hipSuccess) continue;
            Ticket& T = *S->tickets[S->slotTicket[q]];
            if (T.distOut) std::memcpy(T.distOut, S->fDist + q * V, sizeof(double) * V);
            if (T.srcidOut) std::memcpy(T.srcidOut, S->fSrc + q * V, sizeof(int32_t) * V);
            finish_ticket(S, q, fetchRec[q]);
            state[q] = SLOT_FREE;
        }
        // admissions into free slots
        std::vector<Admission> adm;
        int running = 0, fetching = 0;
        for (int q = 0; q < S->slots; ++q) { running += state[q] == SLOT_RUNNING; fetching += state[q] == SLOT_FETCHING; }
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (running == 0 && fetching == 0 && nextRead == k && S->queue.empty()) {
                if (S->stopping) break;
                S->cvWork.wait_for(lk, std::chrono::milliseconds(2));
                continue;
            }
            for (int q = 0; q < S->slots && !S->queue.empty() && (int)adm.size() < S->maxBatch; ++q) {
                if (state[q] != SLOT_FREE) continue;
                const long long id = S->queue.front();
                S->queue.pop_front();
                Ticket& T = *S->tickets[id];
                Admission a;
                a.slot = q; a.kind = T.kind; a.target = T.kind == Q_SSSD ? T.t : -1;
                a.wantParent = T.kind == Q_MSAD && T.srcidOut != nullptr;
                if (T.kind == Q_SSSD) { a.src.push_back(T.s); a.w.push_back(0.0); }
                else { a.src = T.src; a.w = T.w; }
                T.state = 1;
                S->slotTicket[q] = id;
                adm.push_back(std::move(a));
            }
        }
        if (!adm.empty()) {
            std::string e = solver_admit_many(S->st, m, adm);
            if (!e.empty()) { fail_all(S, e); return; }
            for (const Admission& a : adm) { state[a.slot] = SLOT_RUNNING; admitBatch[a.slot] = k; }
            running += (int)adm.size();
            S->maxBatchSeen = std::max(S->maxBatchSeen, running);
        }
        // queue the next batch while the previous one may still run
        if (running > 0) {
            std::string e = solver_enqueue_rounds(S->st, S->g->d, m, S->st.roundsPerSync);
            if (e.empty()) e = solver_enqueue_status(S->st, S->hs[k & 1]);
            if (e.empty() && hipEventRecord(S->batchEv[k & 1], 0) != hipSuccess) e = "event record failed";
            if (!e.empty()) { fail_all(S, e); return; }
            ++k;
        }
        // read the records of the batch before the one just queued (or of the
        // last batch, when nothing new was queued)
        const long long upto = running > 0 ? k - 2 : k - 1;
        for (; nextRead <= upto; ++nextRead) {
            const int pb = (int)(nextRead & 1);
            if (hipEventSynchronize(S->batchEv[pb]) != hipSuccess) { fail_all(S, "batch wait failed"); return; }
            const SlotParam* rec = S->hs[pb];
            if (rec[S->slots].kind) { fail_all(S, "solver frontier overflow"); return; }
            for (int q = 0; q < S->slots; ++q) {
                if (state[q] != SLOT_RUNNING || admitBatch[q] > nextRead || !rec[q].done) continue;
                Ticket& T = *S->tickets[S->slotTicket[q]];
                if (T.kind == Q_SSSD) {
                    if (T.distOut) *T.distOut = rec[q].result;
                    finish_ticket(S, q, rec[q]);
                    state[q] = SLOT_FREE;
                } else {
                    std::string e;
                    if (hipStreamWaitEvent(S->fetchStream, S->batchEv[pb], 0) != hipSuccess) e = "stream wait failed";
                    if (e.empty())
                        e = solver_enqueue_fetch(S->st, q, S->fetchStream, T.distOut ? S->fDist + q * V : nullptr,
                                                 T.srcidOut ? S->fSrc + q * V : nullptr, nullptr);
                    if (e.empty() && hipEventRecord(S->fetchEv[q], S->fetchStream) != hipSuccess)
                        e = "event record failed";
                    if (!e.empty()) { fail_all(S, e); return; }
                    fetchRec[q] = rec[q];
                    state[q] = SLOT_FETCHING;
                }
            }
        }
        // only results in flight: wait for the oldest rather than spin
        if (running == 0 && nextRead == k && fetching > 0) {
            for (int q = 0; q < S->slots; ++q)
                if (state[q] == SLOT_FETCHING) { hipEventSynchronize(S->fetchEv[q]); break; }
        }
    }
}

static void server_free(svg_server_s* S)
{
    solver_destroy(S->st);
    for (int i = 0; i < 2; ++i) {
        if (S->hs[i]) hipHostFree(S->hs[i]);
        if (S->batchEv[i]) hipEventDestroy(S->batchEv[i]);
    }
    for (hipEvent_t e : S->fetchEv) if (e) hipEventDestroy(e);
    if (S->fetchStream) hipStreamDestroy(S->fetchStream);
    if (S->fDist) hipHostFree(S->fDist);
    if (S->fSrc) hipHostFree(S->fSrc);
    delete S;
}

extern "C" int svg_server_start(svg_graph_t g, const svg_server_cfg* cfg, svg_server_t* out)
{
    if (!g || !out) return svg_set_error(SVG_E_ARG, "null argument");
    int rc = svg_ensure_device();
    if (rc) return rc;
    svg_server_s* S = new (std::nothrow) svg_server_s();
    if (!S) return svg_set_error(SVG_E_MEM, "host allocation failed");
    S->g = g;
    int slots = cfg && cfg->max_inflight > 0 ? cfg->max_inflight : 64;
    const size_t perSlot = solver_bytes_per_slot(g->d.V);
    size_t freeB = 0, totB = 0;
    hipMemGetInfo(&freeB, &totB);
    size_t budget = cfg && cfg->mem_budget > 0 ? (size_t)cfg->mem_budget : freeB / 4;
    const int memSlots = (int)std::min<size_t>(4096, budget / perSlot);
    if (memSlots < 1) { delete S; return svg_set_error(SVG_E_MEM, "memory budget below one request slot"); }
    slots = std::min(slots, memSlots);
    S->slots = slots;
    S->maxBatch = cfg && cfg->max_batch > 0 ? cfg->max_batch : slots;
    S->queueCap = (size_t)std::max(1024, 64 * slots);
    const double scale = cfg && cfg->delta_scale > 0.0 ? cfg->delta_scale : 1.0;
    const double delta = graph_mean_arc(g->d) * scale;
    std::string e = solver_create(g->d, g->mesh->d, slots, delta, S->st);
    if (!e.empty()) { solver_destroy(S->st); delete S; return svg_set_error(SVG_E_MEM, e); }
    if (cfg && cfg->rounds_per_sync > 0) S->st.roundsPerSync = std::min(cfg->rounds_per_sync, 64);
    if (cfg && cfg->msad_delta_scale > 0.0) S->st.deltaMsad = delta / scale * cfg->msad_delta_scale;
    bool ok = true;
    for (int i = 0; i < 2 && ok; ++i) {
        ok = hipHostMalloc(&S->hs[i], sizeof(SlotParam) * (size_t)(slots + 1), hipHostMallocDefault) == hipSuccess &&
             hipEventCreateWithFlags(&S->batchEv[i], hipEventDisableTiming) == hipSuccess;
    }
    ok = ok && hipStreamCreateWithFlags(&S->fetchStream, hipStreamNonBlocking) == hipSuccess;
    S->fetchEv.assign((size_t)slots, nullptr);
    for (int q = 0; q < slots && ok; ++q) ok = hipEventCreateWithFlags(&S->fetchEv[q], hipEventDisableTiming) == hipSuccess;
    ok = ok && hipHostMalloc(&S->fDist, sizeof(double) * (size_t)slots * g->d.V, hipHostMallocDefault) == hipSuccess &&
         hipHostMalloc(&S->fSrc, sizeof(int32_t) * (size_t)slots * g->d.V, hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        server_free(S);
        return svg_set_error(SVG_E_MEM, "server staging allocation failed");
    }
    S->slotTicket.assign((size_t)slots, -1);
    S->worker = std::thread(worker_loop, S);
    *out = S;
    return SVG_OK;
}

extern "C" long long svg_server_submit(svg_server_t S, const svg_request* r)
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
        for (int i = 0; i < r->ns; ++i)
            if (r->src[i] < 0 || r->src[i] >= V) return svg_set_error(SVG_E_ARG, "source out of range");
        T->src.assign(r->src, r->src + r->ns);
        T->w.assign(r->w, r->w + r->ns);
    }
    T->distOut = r->dist_out;
    T->srcidOut = r->srcid_out;
    T->submit = Clock::now();
    std::lock_guard<std::mutex> lk(S->mu);
    if (!S->error.empty()) return svg_set_error(SVG_E_HIP, "server failed: " + S->error);
    if (S->queue.size() >= S->queueCap) { S->refused++; return svg_set_error(SVG_E_BUSY, "queue full"); }
    if (!S->anySubmit) { S->firstSubmit = T->submit; S->anySubmit = true; }
    const long long id = (long long)S->tickets.size();
    S->tickets.push_back(std::move(T));
    S->queue.push_back(id);
    S->cvWork.notify_one();
    return id;
}

extern "C" int svg_server_wait(svg_server_t S, long long ticket, double* latency_ms)
{
    if (!S) return svg_set_error(SVG_E_ARG, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    if (ticket < 0 || ticket >= (long long)S->tickets.size()) return svg_set_error(SVG_E_ARG, "unknown ticket");
    Ticket& T = *S->tickets[ticket];
    S->cvDone.wait(lk, [&] { return T.state >= 2; });
    if (T.state == 3) return svg_set_error(SVG_E_HIP, "request failed: " + S->error);
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(T.done - T.submit).count();
    return SVG_OK;
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
    st->launches = S->st.launches;
    for (int k = 0; k < 2; ++k) {
        st->n_kind[k] = S->nKind[k];
        st->rounds_kind[k] = S->roundsKind[k];
        st->settled_kind[k] = S->settledKind[k];
        st->relaxed_kind[k] = S->relaxedKind[k];
    }
    return SVG_OK;
}

extern "C" int svg_server_stop(svg_server_t S)
{
    if (!S) return SVG_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stopping = true;
    }
    S->cvWork.notify_all();
    if (S->worker.joinable()) S->worker.join();
    svg_ensure_device();
    server_free(S);
    return SVG_OK;
}

extern "C" int svg_server_replay(svg_server_t S, const svg_request* reqs, const double* arrival_s, int n,
                                 double* latency_ms)
{
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
        int rc = svg_server_wait(S, ids[i], nullptr);
        if (rc) return rc;
        if (latency_ms) {
            std::lock_guard<std::mutex> lk(S->mu);
            latency_ms[i] = std::chrono::duration<double, std::milli>(S->tickets[ids[i]]->done - sched[i]).count();
        }
    }
    return SVG_OK;
}
