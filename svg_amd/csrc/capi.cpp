// C ABI (include/svg_amd.h): handles, status codes, thread-local errors.
#include "svg_amd.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cmath>
#include <cstring>
#include <fstream>
#include <algorithm>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ich_engine.h"
#include "svg_build.h"
#include "svg_capi_internal.h"

using namespace svg;

static thread_local std::string g_err;

int svg_set_error(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

extern "C" const char* svg_last_error(void) { return g_err.c_str(); }

static int g_device = -1;
static std::mutex g_init_mu;

int svg_ensure_device()
{
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_device >= 0) {
        hipSetDevice(g_device);
        return SVG_OK;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return svg_set_error(SVG_E_HIP, "no HIP device visible");
    if (hipSetDevice(0) != hipSuccess) return svg_set_error(SVG_E_HIP, "hipSetDevice(0) failed");
    g_device = 0;
    return SVG_OK;
}

extern "C" int svg_init(int device)
{
    return svg_guarded<int>([&]() -> int {
        std::lock_guard<std::mutex> lk(g_init_mu);
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
            return svg_set_error(SVG_E_HIP, "no HIP device visible");
        if (device < 0 || device >= n) return svg_set_error(SVG_E_ARG, "device index out of range");
        if (hipSetDevice(device) != hipSuccess) return svg_set_error(SVG_E_HIP, "hipSetDevice failed");
        g_device = device;
        return SVG_OK;
    });
}

extern "C" int svg_device_name(char* buf, int cap)
{
    return svg_guarded<int>([&]() -> int {
        if (!buf || cap <= 0) return svg_set_error(SVG_E_ARG, "null buffer");
        int rc = svg_ensure_device();
        if (rc) return rc;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, g_device) != hipSuccess) return svg_set_error(SVG_E_HIP, "hipGetDeviceProperties");
        snprintf(buf, (size_t)cap, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
        return SVG_OK;
    });
}

// ---------------------------------------------------------------- mesh
extern "C" int svg_mesh_from_arrays(const double* xyz, int nv, const int32_t* tri, int nf, svg_mesh_t* out)
{
    return svg_guarded<int>([&]() -> int {
        if (!xyz || !tri || !out || nv <= 0 || nf <= 0) return svg_set_error(SVG_E_ARG, "bad mesh arrays");
        svg_mesh_s* m = new (std::nothrow) svg_mesh_s();
        if (!m) return svg_set_error(SVG_E_MEM, "host allocation failed");
        std::string e = build_host_mesh(xyz, nv, tri, nf, m->h);
        if (!e.empty()) { delete m; return svg_set_error(SVG_E_ARG, e); }
        *out = m;
        return SVG_OK;
    });
}

extern "C" int svg_mesh_load_obj(const char* path, svg_mesh_t* out)
{
    return svg_guarded<int>([&]() -> int {
        if (!path || !out) return svg_set_error(SVG_E_ARG, "null argument");
        std::vector<double> xyz;
        std::vector<int32_t> tri;
        std::string e = load_obj(path, xyz, tri);
        if (!e.empty()) return svg_set_error(SVG_E_IO, e);
        return svg_mesh_from_arrays(xyz.data(), (int)(xyz.size() / 3), tri.data(), (int)(tri.size() / 3), out);
    });
}

int svg_mesh_upload_impl(svg_mesh_s* m, bool aos)
{
    int rc = svg_ensure_device();
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(m->upMu);
    std::string e = upload_mesh(m->h, m->d, aos);
    if (!e.empty()) return svg_set_error(SVG_E_HIP, e);
    return SVG_OK;
}

extern "C" int svg_mesh_upload(svg_mesh_t m)
{
    return svg_guarded<int>([&]() -> int {
        if (!m) return svg_set_error(SVG_E_ARG, "null mesh");
        return svg_mesh_upload_impl(m, false);
    });
}

extern "C" int svg_mesh_info(svg_mesh_t m, int* nv, int* nf, int* ne)
{
    return svg_guarded<int>([&]() -> int {
        if (!m) return svg_set_error(SVG_E_ARG, "null mesh");
        if (nv) *nv = m->h.nv;
        if (nf) *nf = m->h.nf;
        if (ne) *ne = m->h.ne;
        return SVG_OK;
    });
}

extern "C" int svg_mesh_export(svg_mesh_t m, int32_t* twin, double* len, double* angle, int32_t* firstEdge,
                               uint8_t* boundary)
{
    return svg_guarded<int>([&]() -> int {
        if (!m) return svg_set_error(SVG_E_ARG, "null mesh");
        const HostMesh& h = m->h;
        if (twin) std::memcpy(twin, h.twin.data(), sizeof(int32_t) * (size_t)h.ne);
        if (len) std::memcpy(len, h.len.data(), sizeof(double) * (size_t)h.ne);
        if (angle) std::memcpy(angle, h.angle.data(), sizeof(double) * (size_t)h.nv);
        if (firstEdge) std::memcpy(firstEdge, h.firstEdge.data(), sizeof(int32_t) * (size_t)h.nv);
        if (boundary)
            for (int v = 0; v < h.nv; ++v) boundary[v] = (h.flags[v] & VF_BOUNDARY) ? 1 : 0;
        return SVG_OK;
    });
}

extern "C" int svg_mesh_free(svg_mesh_t m)
{
    return svg_guarded<int>([&]() -> int {
        if (!m) return SVG_OK;
        if (m->d.face || m->d.edges) {
            svg_ensure_device();
            free_dev_mesh(m->d);
        }
        delete m;
        return SVG_OK;
    });
}

// ---------------------------------------------------------------- build
extern "C" void svg_build_params_default(svg_build_params* p, int K)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->K = K;
    p->split_coef = 10;
    p->vert_coef = 2;
    p->mode = SVG_MODE_BATCHED;
    p->plain_blocks = 64;
    p->plain_threads = 128;
}

extern "C" int svg_build(svg_mesh_t m, const svg_build_params* p, svg_graph_t* out)
{
    return svg_guarded<int>([&]() -> int {
        if (!m || !p || !out) return svg_set_error(SVG_E_ARG, "null argument");
        if (p->K <= 0) return svg_set_error(SVG_E_ARG, "K must be positive");
        if (p->mode != SVG_MODE_PLAIN && p->mode != SVG_MODE_BATCHED) return svg_set_error(SVG_E_ARG, "unknown mode");
        if (p->split_coef == 0 || p->vert_coef == 0) return svg_set_error(SVG_E_ARG, "table coefficients must be positive");
        if (p->source_order < 0 || p->source_order > 2) return svg_set_error(SVG_E_ARG, "source_order must be 0, 1 or 2");
        // the reference-layout arrays only for the plain path; a batched
        // build uploads them if it has sources to recompute there
        int rc = svg_mesh_upload_impl(m, p->mode == SVG_MODE_PLAIN);
        if (rc) return rc;
        BuildConfig c;
        c.ensurePlainMesh = [m]() -> std::string {
            const int urc = svg_mesh_upload_impl(m, true);
            return urc ? std::string("reference-layout mesh upload: ") + svg_last_error() : std::string();
        };
        c.K = p->K; c.splitCoef = p->split_coef; c.vertCoef = p->vert_coef; c.mode = p->mode;
        c.plainBlocks = p->plain_blocks; c.plainThreads = p->plain_threads;
        c.lanes = p->lanes; c.propBlock = p->prop_block; c.chunk = p->chunk;
        c.vertCap = p->vert_cap; c.splitCap = p->split_cap; c.autotune = p->autotune;
        c.lanesPerWave = p->lanes_per_wave;
        c.queueCap = p->queue_cap;
        c.sourceOrder = p->source_order;
        if (c.plainThreads > 1024 || (c.plainThreads % 64) != 0 && c.plainThreads > 0)
            return svg_set_error(SVG_E_ARG, "plain_threads must be a multiple of 64 up to 1024");
        svg_graph_s* g = new (std::nothrow) svg_graph_s();
        if (!g) return svg_set_error(SVG_E_MEM, "host allocation failed");
        g->mesh = m;
        std::string e = alloc_graph(m->h.nv, p->K, g->d);
        if (!e.empty()) { free_graph(g->d); delete g; return svg_set_error(SVG_E_MEM, e); }
        BuildStats st;
        if (c.mode == SVG_MODE_BATCHED && c.autotune) {
            std::lock_guard<std::mutex> lk(m->mu);
            TuneRecord* r = nullptr;
            for (TuneRecord& t : m->tunes)
                if (t.K == c.K && t.splitCoef == c.splitCoef && t.vertCoef == c.vertCoef) r = &t;
            if (!r) {
                m->tunes.emplace_back();
                r = &m->tunes.back();
                r->K = c.K; r->splitCoef = c.splitCoef; r->vertCoef = c.vertCoef;
            }
            c.tune = r;
            e = build_batched(m->h, m->d, c, g->d, st);
        } else if (c.mode == SVG_MODE_PLAIN) {
            e = build_plain(m->h, m->d, c, g->d, st, nullptr, 0);
        } else {
            e = build_batched(m->h, m->d, c, g->d, st);
        }
        if (!e.empty()) { free_graph(g->d); delete g; return svg_set_error(SVG_E_HIP, e); }
        std::vector<int32_t> tails((size_t)m->h.nv);
        const hipError_t te = hipMemcpy(tails.data(), g->d.tail, sizeof(int32_t) * tails.size(), hipMemcpyDeviceToHost);
        if (te != hipSuccess) {
            free_graph(g->d); delete g;
            return svg_set_error(SVG_E_HIP, std::string("row lengths to host: ") + hipGetErrorString(te));
        }
        long long arcs = 0;
        for (int32_t t : tails) arcs += t;
        st.arcs = arcs;
        g->stats = st;
        g->arcs = arcs;
        *out = g;
        return SVG_OK;
    });
}

extern "C" int svg_build_stats_get(svg_graph_t g, svg_build_stats* st)
{
    return svg_guarded<int>([&]() -> int {
        if (!g || !st) return svg_set_error(SVG_E_ARG, "null argument");
        std::memset(st, 0, sizeof(*st));
        const BuildStats& s = g->stats;
        st->seconds = s.seconds; st->prop_seconds = s.propSeconds; st->extract_seconds = s.extractSeconds;
        st->redo_seconds = s.redoSeconds; st->alloc_seconds = s.allocSeconds;
        st->arcs = s.arcs; st->redone = s.redone;
        st->flagged_drop = s.flaggedDrop; st->flagged_overflow = s.flaggedOverflow; st->flagged_error = s.flaggedError;
        st->lanes = s.lanes; st->prop_block = s.propBlock; st->chunk = s.chunk;
        st->vert_cap = s.vertCap; st->split_cap = s.splitCap; st->tuned = s.tuned;
        for (size_t i = 0; i < s.counters.size() && i < 16; ++i) st->counters[i] = s.counters[i];
        st->lanes_per_wave = s.lanesPerWave; st->fallback_plain = s.fallbackPlain;
        st->queue_cap = s.queueCap; st->pseudo_queue_cap = s.pseudoCap;
        st->source_order = s.sourceOrder;
        return SVG_OK;
    });
}

extern "C" int svg_build_status(svg_graph_t g, uint32_t* status)
{
    return svg_guarded<int>([&]() -> int {
        if (!g || !status) return svg_set_error(SVG_E_ARG, "null argument");
        const auto& s = g->stats.status;
        for (int v = 0; v < g->d.V; ++v) status[v] = (size_t)v < s.size() ? s[v] : 0u;
        return SVG_OK;
    });
}

// ---------------------------------------------------------------- graph
extern "C" int svg_graph_info(svg_graph_t g, int* V, int* K, long long* arcs)
{
    return svg_guarded<int>([&]() -> int {
        if (!g) return svg_set_error(SVG_E_ARG, "null graph");
        if (V) *V = g->d.V;
        if (K) *K = g->d.K;
        if (arcs) *arcs = g->arcs;
        return SVG_OK;
    });
}

namespace {

// Device -> pageable host copies of the graph (about 18 B per row slot: 1.3 GB
// for V = 72K, K = 500) go through two pinned 64 MiB buffers: the DMA of
// chunk i+1 runs while four threads copy chunk i out, instead of the
// runtime's pageable path (about 7 GB/s).  The buffers live until the
// process ends.
struct Bounce {
    std::mutex mu;
    void* buf[2] = {nullptr, nullptr};
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
};
Bounce g_bounce;
constexpr size_t kBounceBytes = (size_t)64 << 20;

void par_memcpy(char* d, const char* s, size_t n)
{
    constexpr int kThreads = 4;
    if (n < ((size_t)4 << 20)) { std::memcpy(d, s, n); return; }
    const size_t part = (n + kThreads - 1) / kThreads;
    std::thread th[kThreads - 1];
    for (int t = 1; t < kThreads; ++t) {
        const size_t a = std::min(n, (size_t)t * part), b = std::min(n, a + part);
        th[t - 1] = std::thread([=] { if (b > a) std::memcpy(d + a, s + a, b - a); });
    }
    std::memcpy(d, s, std::min(n, part));
    for (auto& x : th) x.join();
}

hipError_t d2h(void* dst, const void* src, size_t n)
{
    if (n < ((size_t)8 << 20)) return hipMemcpy(dst, src, n, hipMemcpyDeviceToHost);
    std::lock_guard<std::mutex> lk(g_bounce.mu);
    hipError_t e = hipSuccess;
    if (!g_bounce.buf[0]) {
        for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipHostMalloc(&g_bounce.buf[i], kBounceBytes, hipHostMallocDefault);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&g_bounce.stream, hipStreamNonBlocking);
        for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&g_bounce.ev[i], hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    // the rows were written by launches on the null stream
    e = hipStreamSynchronize(0);
    const size_t chunks = (n + kBounceBytes - 1) / kBounceBytes;
    for (size_t i = 0; i <= chunks && e == hipSuccess; ++i) {
        if (i < chunks) {
            const size_t off = i * kBounceBytes, len = std::min(kBounceBytes, n - off);
            e = hipMemcpyAsync(g_bounce.buf[i & 1], (const char*)src + off, len, hipMemcpyDeviceToHost, g_bounce.stream);
            if (e == hipSuccess) e = hipEventRecord(g_bounce.ev[i & 1], g_bounce.stream);
        }
        if (i > 0 && e == hipSuccess) {
            const size_t j = i - 1, off = j * kBounceBytes, len = std::min(kBounceBytes, n - off);
            e = hipEventSynchronize(g_bounce.ev[j & 1]);
            if (e == hipSuccess) par_memcpy((char*)dst + off, (const char*)g_bounce.buf[j & 1], len);
        }
    }
    return e;
}

}  // namespace

extern "C" int svg_graph_to_host(svg_graph_t g, int32_t* tails, int32_t* adj, double* dist, int32_t* src_edge,
                                 int32_t* dst_edge, double* src_x, double* dst_x)
{
    return svg_guarded<int>([&]() -> int {
        if (!g) return svg_set_error(SVG_E_ARG, "null graph");
        int rc = svg_ensure_device();
        if (rc) return rc;
        const size_t n = (size_t)g->d.V * (size_t)g->d.K;
        hipError_t e = hipSuccess;
        if (tails && e == hipSuccess) e = hipMemcpy(tails, g->d.tail, sizeof(int32_t) * (size_t)g->d.V, hipMemcpyDeviceToHost);
        if (adj && e == hipSuccess) e = d2h(adj, g->d.adj, sizeof(int32_t) * n);
        if (dist && e == hipSuccess) e = d2h(dist, g->d.w, sizeof(double) * n);
        if (src_edge && e == hipSuccess) e = d2h(src_edge, g->d.srcEdge, sizeof(int32_t) * n);
        if (dst_edge && e == hipSuccess) e = d2h(dst_edge, g->d.dstEdge, sizeof(int32_t) * n);
        if (src_x && e == hipSuccess) e = d2h(src_x, g->d.srcX, sizeof(double) * n);
        if (dst_x && e == hipSuccess) e = d2h(dst_x, g->d.dstX, sizeof(double) * n);
        if (e != hipSuccess) return svg_set_error(SVG_E_HIP, std::string("device->host copy: ") + hipGetErrorString(e));
        return SVG_OK;
    });
}

static int copy_points(const std::vector<double>& pts, double* xyz, int cap, int* n)
{
    const int np = (int)(pts.size() / 3);
    if (n) *n = np;
    if (np > cap) return svg_set_error(SVG_E_ARG, "output holds " + std::to_string(cap) + " points, " +
                                                     std::to_string(np) + " needed");
    if (xyz) std::memcpy(xyz, pts.data(), sizeof(double) * pts.size());
    return SVG_OK;
}

extern "C" int svg_trace_arc(svg_mesh_t m, int u, int v, int32_t src_edge, double src_x, int32_t dst_edge,
                             double* xyz, int cap, int* n, double* dst_x_traced)
{
    return svg_guarded<int>([&]() -> int {
        if (!m) return svg_set_error(SVG_E_ARG, "null mesh");
        const svg::HostMesh& h = m->h;
        if (u < 0 || u >= h.nv) return svg_set_error(SVG_E_ARG, "vertex out of range");
        std::vector<double> pts(h.xyz.begin() + 3 * u, h.xyz.begin() + 3 * u + 3);
        std::string e = svg::trace_arc(h, u, v, src_edge, src_x, dst_edge, h.nf, pts, dst_x_traced);
        if (!e.empty()) return svg_set_error(SVG_E_ARG, "arc " + std::to_string(u) + "->" + std::to_string(v) + ": " + e);
        return copy_points(pts, xyz, cap, n);
    });
}

extern "C" int svg_geodesic_path(svg_graph_t g, const int32_t* verts, int nverts, double* xyz, int cap, int* n)
{
    return svg_guarded<int>([&]() -> int {
        if (!g || !verts || nverts < 1) return svg_set_error(SVG_E_ARG, "null graph or empty path");
        int rc = svg_ensure_device();
        if (rc) return rc;
        const svg::HostMesh& h = g->mesh->h;
        const int K = g->d.K;
        for (int i = 0; i < nverts; ++i)
            if (verts[i] < 0 || verts[i] >= h.nv) return svg_set_error(SVG_E_ARG, "vertex out of range");
        std::vector<double> pts(h.xyz.begin() + 3 * verts[0], h.xyz.begin() + 3 * verts[0] + 3);
        std::vector<int32_t> adj((size_t)K), se((size_t)K), de((size_t)K);
        std::vector<double> sx((size_t)K);
        for (int i = 0; i + 1 < nverts; ++i) {
            const int u = verts[i], v = verts[i + 1];
            int32_t tl = 0;
            const size_t o = (size_t)u * (size_t)K;
            hipError_t he = hipMemcpy(&tl, g->d.tail + u, sizeof(int32_t), hipMemcpyDeviceToHost);
            if (he == hipSuccess && tl > 0) he = hipMemcpy(adj.data(), g->d.adj + o, sizeof(int32_t) * (size_t)tl, hipMemcpyDeviceToHost);
            if (he != hipSuccess) return svg_set_error(SVG_E_HIP, std::string("row read: ") + hipGetErrorString(he));
            int j = -1;
            for (int k = 0; k < tl; ++k) if (adj[(size_t)k] == v) { j = k; break; }
            if (j < 0) return svg_set_error(SVG_E_ARG, "no arc " + std::to_string(u) + "->" + std::to_string(v));
            int32_t es = -1, ed = -1;
            double xs = 0.0;
            he = hipMemcpy(&es, g->d.srcEdge + o + j, sizeof(int32_t), hipMemcpyDeviceToHost);
            if (he == hipSuccess) he = hipMemcpy(&ed, g->d.dstEdge + o + j, sizeof(int32_t), hipMemcpyDeviceToHost);
            if (he == hipSuccess) he = hipMemcpy(&xs, g->d.srcX + o + j, sizeof(double), hipMemcpyDeviceToHost);
            if (he != hipSuccess) return svg_set_error(SVG_E_HIP, std::string("row read: ") + hipGetErrorString(he));
            std::string e = svg::trace_arc(h, u, v, es, xs, ed, h.nf, pts, nullptr);
            if (!e.empty()) return svg_set_error(SVG_E_ARG, "arc " + std::to_string(u) + "->" + std::to_string(v) + ": " + e);
        }
        return copy_points(pts, xyz, cap, n);
    });
}

// SVGNode file record (SVG.cuh:21-27, 40 bytes with 4 bytes of padding).
struct FileNode {
    int32_t adjNode;
    int32_t pad;
    double geodDist;
    uint32_t nextToSrcEdge, nextToDstEdge;
    double nextToSrcX, nextToDstX;
};
static_assert(sizeof(FileNode) == 40, "SVGNode file record is 40 bytes");

extern "C" int svg_graph_save(svg_graph_t g, const char* path)
{
    return svg_guarded<int>([&]() -> int {
        if (!g || !path) return svg_set_error(SVG_E_ARG, "null argument");
        const int V = g->d.V, K = g->d.K;
        const size_t n = (size_t)V * (size_t)K;
        std::vector<int32_t> tails(V), adj(n), se(n), de(n);
        std::vector<double> w(n), sx(n), dx(n);
        int rc = svg_graph_to_host(g, tails.data(), adj.data(), w.data(), se.data(), de.data(), sx.data(), dx.data());
        if (rc) return rc;
        std::ofstream out(path, std::ios::binary);
        if (!out) return svg_set_error(SVG_E_IO, std::string("cannot write ") + path);
        std::vector<FileNode> row((size_t)K);
        for (int v = 0; v < V; ++v) {
            const int t = tails[v];
            out.write(reinterpret_cast<const char*>(&tails[v]), sizeof(int32_t));
            for (int i = 0; i < t; ++i) {
                const size_t k = (size_t)v * K + i;
                FileNode& f = row[i];
                f.adjNode = adj[k]; f.pad = 0; f.geodDist = w[k];
                f.nextToSrcEdge = (uint32_t)se[k]; f.nextToDstEdge = (uint32_t)de[k];
                f.nextToSrcX = sx[k]; f.nextToDstX = dx[k];
            }
            out.write(reinterpret_cast<const char*>(row.data()), sizeof(FileNode) * (size_t)t);
        }
        if (!out) return svg_set_error(SVG_E_IO, "write failed");
        return SVG_OK;
    });
}

extern "C" int svg_graph_load(svg_mesh_t m, int K, const char* path, svg_graph_t* out)
{
    return svg_guarded<int>([&]() -> int {
        if (!m || !path || !out || K <= 0) return svg_set_error(SVG_E_ARG, "bad argument");
        std::ifstream in(path, std::ios::binary);
        if (!in) return svg_set_error(SVG_E_IO, std::string("cannot open ") + path);
        const int V = m->h.nv;
        const size_t n = (size_t)V * (size_t)K;
        std::vector<int32_t> tails(V, 0), adj(n, 0), se(n, -1), de(n, -1);
        std::vector<double> w(n, 0.0), sx(n, 0.0), dx(n, 0.0);
        std::vector<FileNode> row;
        long long arcs = 0;
        for (int v = 0; v < V; ++v) {
            int32_t t = 0;
            in.read(reinterpret_cast<char*>(&t), sizeof(int32_t));
            if (!in || t < 0 || t > K) return svg_set_error(SVG_E_IO, "bad or truncated .svg file");
            row.resize((size_t)t);
            in.read(reinterpret_cast<char*>(row.data()), sizeof(FileNode) * (size_t)t);
            if (!in) return svg_set_error(SVG_E_IO, "truncated .svg file");
            tails[v] = t;
            arcs += t;
            for (int i = 0; i < t; ++i) {
                const size_t k = (size_t)v * K + i;
                if (row[i].adjNode < 0 || row[i].adjNode >= V) return svg_set_error(SVG_E_IO, "arc target out of range");
                // the solvers bucket distances by the arc weights and the
                // polyline tracer walks from the key-point edges: a file
                // that is not the output of a build is checked here
                if (!(row[i].geodDist >= 0.0) || !std::isfinite(row[i].geodDist))
                    return svg_set_error(SVG_E_IO, "arc length is not a finite non-negative number");
                const int32_t es = (int32_t)row[i].nextToSrcEdge, ed = (int32_t)row[i].nextToDstEdge;
                if (es < -1 || es >= m->h.ne || ed < -1 || ed >= m->h.ne)
                    return svg_set_error(SVG_E_IO, "key-point edge out of range");
                if (!std::isfinite(row[i].nextToSrcX) || !std::isfinite(row[i].nextToDstX))
                    return svg_set_error(SVG_E_IO, "key-point position is not finite");
                adj[k] = row[i].adjNode; w[k] = row[i].geodDist;
                se[k] = (int32_t)row[i].nextToSrcEdge; de[k] = (int32_t)row[i].nextToDstEdge;
                sx[k] = row[i].nextToSrcX; dx[k] = row[i].nextToDstX;
            }
        }
        int rc = svg_mesh_upload_impl(m, false);
        if (rc) return rc;
        svg_graph_s* g = new (std::nothrow) svg_graph_s();
        if (!g) return svg_set_error(SVG_E_MEM, "host allocation failed");
        g->mesh = m;
        std::string e = alloc_graph(V, K, g->d);
        if (!e.empty()) { free_graph(g->d); delete g; return svg_set_error(SVG_E_MEM, e); }
        // every copy is checked: a handle over partly uploaded rows is never returned
        hipError_t he = hipMemcpy(g->d.tail, tails.data(), sizeof(int32_t) * (size_t)V, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(g->d.adj, adj.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(g->d.w, w.data(), sizeof(double) * n, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(g->d.srcEdge, se.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(g->d.dstEdge, de.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(g->d.srcX, sx.data(), sizeof(double) * n, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(g->d.dstX, dx.data(), sizeof(double) * n, hipMemcpyHostToDevice);
        if (he != hipSuccess) {
            free_graph(g->d); delete g;
            return svg_set_error(SVG_E_HIP, std::string(".svg upload: ") + hipGetErrorString(he));
        }
        g->arcs = arcs;
        g->stats.arcs = arcs;
        *out = g;
        return SVG_OK;
    });
}

extern "C" int svg_graph_free(svg_graph_t g)
{
    return svg_guarded<int>([&]() -> int {
        if (!g) return SVG_OK;
        svg_ensure_device();
        svg_solver_release(g);
        svg_point_release(g);
        free_graph(g->d);
        delete g;
        return SVG_OK;
    });
}
