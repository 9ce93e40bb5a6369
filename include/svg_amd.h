/*
 * svg_amd.h -- C ABI of the MI355X Saddle Vertex Graph engine.
 *
 * Plain C types only (pointers, sizes, ints, doubles); every entry point
 * returns 0 on success or a negative SVG_E* code, with a message available
 * from svg_last_error() (per thread).  Nothing calls exit().
 *
 * Each entry point replaces the reference interface cited beside it
 * (files of Narusaki/SVG).  Handles own all device memory; output arrays are
 * owned by the caller.
 */
#ifndef SVG_AMD_H
#define SVG_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVG_OK 0
#define SVG_E_ARG -1        /* bad argument */
#define SVG_E_HIP -2        /* HIP runtime error */
#define SVG_E_MEM -3        /* out of device or host memory */
#define SVG_E_IO -4         /* file error */
#define SVG_E_STATE -5      /* handle in the wrong state */
#define SVG_E_BUSY -6       /* serving: admission refused (memory / queue full) */
#define SVG_E_INTERNAL -7   /* host-side failure inside the library (message in svg_last_error) */

#define SVG_MODE_PLAIN 0    /* one source per thread, reference layout (SVG.cu:9-96) */
#define SVG_MODE_BATCHED 1  /* persistent work-stealing lanes + wave extraction */

typedef struct svg_mesh_s *svg_mesh_t;
typedef struct svg_graph_s *svg_graph_t;
typedef struct svg_server_s *svg_server_t;

/* Library / device.  Replaces cudaSetDevice(0) of kernel.cu:25. */
int svg_init(int device);
const char *svg_last_error(void);
int svg_device_name(char *buf, int cap);

/* ---------------------------------------------------------------- mesh */
/* Mesh::LoadFromFile (Mesh.cu:24-126) from arrays: xyz[3*nv] (f64),
 * tri[3*nf] (0-based, counter-clockwise).  Builds half-edges, twins,
 * angle sums and firstEdge with the reference rules; on open meshes a
 * boundary half-edge has twin -1, a boundary vertex's firstEdge is its
 * outgoing half-edge whose prev has no twin, and it is never a saddle. */
int svg_mesh_from_arrays(const double *xyz, int nv, const int32_t *tri, int nf, svg_mesh_t *out);
/* Mesh::LoadFromFile (Mesh.cu:24-126): 'v x y z' and 'f a b c' lines as
 * there; also 'f' corners written i/t/n, i//n or as negative indices, and
 * polygons (split into fans).  Indices are checked by svg_mesh_from_arrays. */
int svg_mesh_load_obj(const char *path, svg_mesh_t *out);
/* Mesh::copyToGPU (Mesh.cu:128-142).  Called implicitly by svg_build. */
int svg_mesh_upload(svg_mesh_t m);
int svg_mesh_info(svg_mesh_t m, int *nv, int *nf, int *ne);
/* Host half-edge arrays: twin[ne], len[ne], angle[nv], firstEdge[nv],
 * boundary[nv] (any may be NULL). */
int svg_mesh_export(svg_mesh_t m, int32_t *twin, double *len, double *angle, int32_t *firstEdge,
                    uint8_t *boundary);
/* Mesh::clear + Mesh::clearGPU (Mesh.cu:144-159). */
int svg_mesh_free(svg_mesh_t m);

/* ---------------------------------------------------------------- build */
typedef struct {
    int K;                  /* SVG::SetParameters K (SVG.cu:114-117) */
    uint32_t split_coef;    /* splitInfoCoef, reference value 10 (kernel.cu:422) */
    uint32_t vert_coef;     /* vertInfoCoef, reference value 2 */
    int mode;               /* SVG_MODE_PLAIN or SVG_MODE_BATCHED */
    int plain_blocks;       /* plain launch shape; 0 -> 64 x 128 (SVG.cuh:10-11) */
    int plain_threads;
    int lanes;              /* batched: propagation lanes in flight (0 = auto) */
    int prop_block;         /* batched: threads per propagation block (0 = auto) */
    int chunk;              /* batched: sources with resident tables (0 = auto) */
    int vert_cap;           /* batched: vertex-table slots per source (0 = auto) */
    int split_cap;          /* batched: split-table slots per source (0 = auto) */
    int autotune;           /* batched: learn from each build of this mesh (per K and
                               coefficients) the source order, table sizes and block shape */
    int lanes_per_wave;     /* batched: sources stepped by one 64-wide wave (1..64; 0 = auto) */
    int queue_cap;          /* batched: window / pseudo-window queue storage per source lane
                               (1..1023; 0 = the reference's 1023, or with autotune the peak
                               recorded by earlier builds).  A source that outgrows a storage
                               below 1023 is recomputed exactly on the plain path, so the
                               rows never depend on it. */
    int source_order;       /* batched: order in which lanes take sources (rows never depend
                               on it): 0 = auto (with autotune, the costliest first by the
                               event counts recorded by an earlier build; otherwise, when
                               there are more sources than lanes, the order predicted from
                               the share of saddle vertices around each source), 1 = vertex
                               id order, 2 = the predicted order */
} svg_build_params;

typedef struct {
    double seconds;         /* device time of the build kernels (HIP events) */
    double prop_seconds;
    double extract_seconds;
    double redo_seconds;
    double alloc_seconds;
    long long arcs;
    int redone;             /* sources recomputed on the exact plain path */
    int flagged_drop;       /* sources whose heap dropped a push (as the reference does) */
    int flagged_overflow;   /* sources whose reference tables would overflow */
    int flagged_error;
    int lanes, prop_block, chunk, vert_cap, split_cap;
    int tuned;              /* 1: source order, table sizes and block shape came from the
                               autotuner's record of earlier builds of this mesh */
    long long counters[16]; /* event counters (win push/pop, pseudo push/pop, propagate,
                               children, split/vert writes, fan half-edges, pseudo
                               neighbours, -, -, backtrace hops, reached, -, -) */
    int lanes_per_wave;     /* sources per wave of the propagation kernel */
    int queue_cap;          /* window queue storage per lane used (1023: the reference's) */
    int pseudo_queue_cap;   /* pseudo-window queue storage per lane used */
    int fallback_plain;     /* 1: a batched build ran every source on the plain path (the
                               mesh has 2^24-1 or more half-edges) */
    int source_order;       /* order the sources ran in: 1 vertex id, 2 predicted from the
                               mesh's saddle vertices, 3 costliest first by recorded event
                               counts (0: plain build) */
} svg_build_stats;

void svg_build_params_default(svg_build_params *p, int K);
/* SVG::SetParameters + Allocation + ConstructSVG + Free (SVG.cu:109-182). */
int svg_build(svg_mesh_t m, const svg_build_params *p, svg_graph_t *out);
int svg_build_stats_get(svg_graph_t g, svg_build_stats *st);
/* Per-vertex status bits of the build (1 drop, 2 pseudo drop, 4 reference
 * table overflow, 8 batched capacity, 16 error, 32 recomputed exactly). */
int svg_build_status(svg_graph_t g, uint32_t *status);

/* ---------------------------------------------------------------- graph */
int svg_graph_info(svg_graph_t g, int *V, int *K, long long *arcs);
/* SVG::CopySVGToHost (SVG.cu:184-206): tails[V]; rows in K-stride layout
 * (any array may be NULL): adj[V*K], dist[V*K], src_edge[V*K],
 * dst_edge[V*K], src_x[V*K], dst_x[V*K] -- the SVGNode fields (SVG.cuh:21-27). */
int svg_graph_to_host(svg_graph_t g, int32_t *tails, int32_t *adj, double *dist, int32_t *src_edge,
                      int32_t *dst_edge, double *src_x, double *dst_x);
/* SVG::SaveSVGToFile / LoadSVGFromFile (SVG.cu:208-239): per vertex an int32
 * tail then tail 40-byte SVGNode records. */
int svg_graph_save(svg_graph_t g, const char *path);
int svg_graph_load(svg_mesh_t m, int K, const char *path, svg_graph_t *out);
/* SVG::FreeSVGStructure (SVG.cu:152-158). */
int svg_graph_free(svg_graph_t g);

/* ---------------------------------------------------------------- queries */
/* SVG::SolveSSSD(s, t) (SVG.cu:241-251): A* distance s -> t on the graph;
 * path (optional, cap entries) receives t, parent(t), ..., s. */
int svg_sssd(svg_graph_t g, int s, int t, double *dist_t, int32_t *path, int path_cap, int *path_len);
/* SVG::SolveMSAD (SVG.cu:354-367): multi-source Dijkstra with initial
 * weights; dist[V], src_id[V], parent[V] (src_id / parent may be NULL). */
int svg_msad(svg_graph_t g, const int32_t *src, const double *w, int ns, double *dist,
             int32_t *src_id, int32_t *parent);
/* svg_msad without device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and answers as svg_msad; its working memory is a fixed 64 MiB
 * device array that the library holds on each GPU it runs on until the
 * process ends, which limits it to graphs of about 87,000 vertices (larger
 * ones return SVG_E_ARG).  It calls no hipMalloc / hipFree and no pinned
 * allocation.  Calls share the array: the library serialises them within a
 * process, so calls overlapping on one GPU wait for each other. */
int svg_msad_noalloc(svg_graph_t g, const int32_t *src, const double *w, int ns, double *dist,
                     int32_t *src_id, int32_t *parent);
/* SVG::SolveSSSD(f0, p0, f1, p1) (SVG.cu:253-352): geodesic distance
 * between point p0[3] on face f0 and point p1[3] on face f1 (coordinates in
 * the mesh's frame).  Bounded ICH runs from both points (K of the graph,
 * tables K*10 / K*2 as kernel.cu:188-191) answer directly through f1's
 * windows when that path records a source-side key point; otherwise the
 * graph joins them at the vertex minimising graph distance + p1 distance.
 * key_edge[2] / key_x[2] (may be NULL) receive nextToSrcEdge/X and
 * nextToDstEdge/X; *last_vert the last graph vertex (saddleVertNearDst, -1
 * when direct); *route 0 direct, 1 through the graph, -1 nothing joins the
 * points (*dist = DBL_MAX).  path (optional, cap entries) receives the graph
 * vertices from last_vert back to the first.  A point on a vertex of its
 * face makes the propagation's keys NaN (the reference loops there): the
 * call returns SVG_E_ARG. */
int svg_sssd_point(svg_graph_t g, int f0, const double *p0, int f1, const double *p1, double *dist,
                   int32_t *key_edge, double *key_x, int32_t *last_vert, int *route, int32_t *path,
                   int path_cap, int *path_len);
/* Many surface-point queries at once (no reference counterpart; the
 * reference answers one per call): faces[2*nq] = (f0, f1), points[6*nq] =
 * (p0, p1) per query; dist[nq]; key_edge[2*nq], key_x[2*nq], last_vert[nq],
 * route[nq] (may be NULL) as svg_sssd_point; a query whose propagation
 * stalls (a point on a vertex of its face) gets route -2 and a NaN distance.
 * Each ICH run is one lane (4 per wave, up to 8,192 queries per launch); the
 * graph stages share the solver's slots. */
int svg_sssd_point_batch(svg_graph_t g, int nq, const int32_t *faces, const double *points, double *dist,
                         int32_t *key_edge, double *key_x, int32_t *last_vert, int32_t *route);
/* Many SSSD queries in shared launches: st[2*nq] pairs -> out[nq]
 * (SVG::SolveSSSD, SVG.cu:241-251, once per pair). */
int svg_sssd_batch(svg_graph_t g, const int32_t *st, int nq, double *out);
/* Many MSAD queries in shared launches (SVG::SolveMSAD, SVG.cu:354-367, once
 * per query): query i has sources src[offsets[i] .. offsets[i+1]) with
 * weights w[same range] (offsets[nq+1], increasing); dist[nq*V] and
 * src_id[nq*V] (may be NULL) receive each query's V values in turn. */
int svg_msad_batch(svg_graph_t g, const int32_t *src, const double *w, const int64_t *offsets, int nq,
                   double *dist, int32_t *src_id);

/* ------------------------------------------------------- geodesic paths */
/* The surface polyline of an SVG arc u -> v from the key points its row
 * stores (SVGNode::nextToSrcEdge/X, nextToDstEdge, SVG.cuh:21-27, written by
 * BuildGeodesicPathTo ICHDevice.cu:336-480): the straight line through the
 * unfolded faces, traced as the reference's unused InitialValueGeodesic.cu:
 * 54-256 sketches.  xyz[3*cap] receives u, every edge crossing and v; *n the
 * point count (set, with SVG_E_ARG, also when cap is too small); the traced
 * position on the last key point's edge goes to *dst_x_traced (may be NULL).
 * Host-side; needs no device. */
int svg_trace_arc(svg_mesh_t m, int u, int v, int32_t src_edge, double src_x, int32_t dst_edge,
                  double *xyz, int cap, int *n, double *dst_x_traced);
/* The surface polyline of a vertex path whose consecutive vertices are SVG
 * arcs (such as the path svg_sssd returns, read from its end): where the
 * reference prints the vertex sequence (kernel.cu:75-81). */
int svg_geodesic_path(svg_graph_t g, const int32_t *verts, int nverts, double *xyz, int cap, int *n);

/* ---------------------------------------------------------------- serving */
typedef struct {
    int max_inflight;       /* request slots resident on the device */
    long long mem_budget;   /* bytes of device memory for request state (0 = auto) */
    int max_batch;          /* requests merged into one frontier launch */
    double delta_scale;     /* bucket width = delta_scale x mean arc length (0 = 1) */
    int rounds_per_sync;    /* frontier rounds queued between two checks for finished
                               requests and new admissions (0 = 4) */
    double msad_delta_scale; /* bucket width of MSAD requests / mean arc length
                               (0 = delta_scale) */
    int streams;            /* slot groups, each with its own HIP stream and worker, whose
                               rounds run concurrently (0 = 4) */
    int chain_rows;         /* same-bucket rows one wave expands itself within a round,
                               up to 64 (0 = 4, negative = none) */
    long long queue_cap;    /* requests that may wait for a slot; more are refused with
                               SVG_E_BUSY (0 = 64 per slot, at least 1024) */
} svg_server_cfg;

typedef struct {
    int kind;               /* 0 = SSSD (A*), 1 = MSAD */
    int s, t;               /* SSSD endpoints */
    const int32_t *src;     /* MSAD sources (copied at submit) */
    const double *w;        /* MSAD weights (copied at submit) */
    int ns;
    double *dist_out;       /* SSSD: 1 value; MSAD: V values (caller-owned) */
    int32_t *srcid_out;     /* MSAD: V values or NULL */
} svg_request;

typedef struct {
    long long completed, refused;
    double p50_ms, p99_ms, mean_ms, max_ms;
    double qps;
    int max_batch_seen;
    long long launches;
    /* work of the completed queries, per kind (0 = SSSD, 1 = MSAD): queries,
     * shared rounds they took part in, frontier entries expanded, arcs scanned */
    long long n_kind[2], rounds_kind[2], settled_kind[2], relaxed_kind[2];
    int slots;              /* request slots admitted by the memory budget */
    long long slot_bytes;   /* device memory of one slot */
} svg_server_stats;

int svg_server_start(svg_graph_t g, const svg_server_cfg *cfg, svg_server_t *out);
/* Enqueue a request; returns a ticket >= 0 or SVG_E_BUSY.  Replaces the
 * reference's one-query-at-a-time driver loops (kernel.cu:63-92, 336-389). */
long long svg_server_submit(svg_server_t s, const svg_request *req);
/* Completion callback: called once per request from a server worker thread
 * (never while a server lock is held) with the ticket, SVG_OK or an error
 * code, and the latency from submit in milliseconds (-1 on error); the
 * request's outputs are complete when it runs.  It may submit further
 * requests but must not wait for tickets or stop the server. */
typedef void (*svg_done_fn)(long long ticket, int status, double latency_ms, void *user);
/* svg_server_submit with a completion callback; such a ticket is not
 * waited for (svg_server_wait rejects it). */
long long svg_server_submit_cb(svg_server_t s, const svg_request *req, svg_done_fn cb, void *user);
/* Block until ticket completes; latency in milliseconds from submit.  The
 * ticket is consumed: a second wait on it is an error. */
int svg_server_wait(svg_server_t s, long long ticket, double *latency_ms);
int svg_server_stats_get(svg_server_t s, svg_server_stats *st);
int svg_server_stop(svg_server_t s);
/* Load driver: submits reqs[i] at arrival_s[i] seconds after the call
 * starts and waits for all; latency_ms[i] is completion minus the scheduled
 * arrival (-1 when the request was refused). */
int svg_server_replay(svg_server_t s, const svg_request *reqs, const double *arrival_s, int n,
                      double *latency_ms);

#ifdef __cplusplus
}
#endif

#endif /* SVG_AMD_H */
