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

__global__ void k_gather_targets(const unsigned long long* dist, const int* slots, const int* targets, int n,
                                 int V, double* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = __longlong_as_double((long long)dist[(size_t)slots[i] * V + targets[i]]);
}

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
    // gather staging
    int* gSlots = nullptr;
    int* gTargets = nullptr;
    double* gOut = nullptr;
    void* pinned = nullptr;
};

static void worker_loop(svg_server_s* S)
{
    svg_ensure_device();
    const DevMesh& m = S->g->mesh->d;
    std::vector<int> active;
    for (;;) {
        std::vector<Admission> adm;
        std::vector<long long> admT;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (active.empty() && S->queue.empty()) {
                if (S->stopping) break;
                S->cvWork.wait_for(lk, std::chrono::milliseconds(2));
                continue;
            }
            for (int q = 0; q < S->slots && !S->queue.empty() && (int)adm.size() < S->maxBatch; ++q) {
                if (S->slotTicket[q] != -1) continue;
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
                admT.push_back(id);
            }
        }
        if (!adm.empty()) {
            std::string e = solver_admit_many(S->st, m, adm);
            if (!e.empty()) {
                std::lock_guard<std::mutex> lk(S->mu);
                S->error = e;
                for (long long id : admT) S->tickets[id]->state = 3;
                for (const Admission& a : adm) S->slotTicket[a.slot] = -1;
                S->cvDone.notify_all();
                continue;
            }
            for (const Admission& a : adm) active.push_back(a.slot);
            S->maxBatchSeen = std::max(S->maxBatchSeen, (int)active.size());
        }
        if (active.empty()) continue;
        int running = 0;
        std::string e = solver_round(S->st, S->g->d, m, running);
        if (!e.empty()) {
            std::lock_guard<std::mutex> lk(S->mu);
            S->error = e;
            for (int q : active) { S->tickets[S->slotTicket[q]]->state = 3; S->slotTicket[q] = -1; }
            active.clear();
            S->cvDone.notify_all();
            continue;
        }
        // retire finished slots: SSSD answers are gathered with one launch
        std::vector<int> fin, finS, finT;
        for (int q : active)
            if (S->st.spHost[q].done) fin.push_back(q);
        if (fin.empty()) continue;
        for (int q : fin) {
            Ticket& T = *S->tickets[S->slotTicket[q]];
            if (T.kind == Q_SSSD) { finS.push_back(q); finT.push_back(T.t); }
        }
        std::vector<double> sssdOut(finS.size());
        if (!finS.empty()) {
            const int n = (int)finS.size();
            std::memcpy(S->pinned, finS.data(), sizeof(int) * (size_t)n);
            std::memcpy((int*)S->pinned + n, finT.data(), sizeof(int) * (size_t)n);
            hipMemcpyAsync(S->gSlots, S->pinned, sizeof(int) * 2 * (size_t)n, hipMemcpyHostToDevice, 0);
            hipLaunchKernelGGL(k_gather_targets, dim3((n + 63) / 64), dim3(64), 0, 0, S->st.dist, S->gSlots,
                               S->gSlots + n, n, S->st.V, S->gOut);
            hipMemcpy(sssdOut.data(), S->gOut, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost);
        }
        size_t k = 0;
        for (int q : fin) {
            const long long id = S->slotTicket[q];
            Ticket& T = *S->tickets[id];
            if (T.kind == Q_SSSD) {
                if (T.distOut) *T.distOut = sssdOut[k];
                ++k;
            } else {
                solver_fetch(S->st, q, T.distOut, nullptr, T.srcidOut, nullptr);
            }
            const SlotParam& sp = S->st.spHost[q];
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
        }
        active.erase(std::remove_if(active.begin(), active.end(),
                                    [&](int q) { return S->slotTicket[q] == -1; }), active.end());
        S->cvDone.notify_all();
    }
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
    if (hipMalloc(&S->gSlots, sizeof(int) * 2 * (size_t)slots) != hipSuccess ||
        hipMalloc(&S->gOut, sizeof(double) * (size_t)slots) != hipSuccess ||
        hipHostMalloc(&S->pinned, sizeof(int) * 2 * (size_t)slots, hipHostMallocDefault) != hipSuccess) {
        solver_destroy(S->st); delete S;
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
    solver_destroy(S->st);
    hipFree(S->gSlots); hipFree(S->gOut); hipHostFree(S->pinned);
    delete S;
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
