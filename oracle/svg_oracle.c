/*
 * svg_oracle.c -- CPU restatement of the reference Saddle Vertex Graph path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * entry point and the benchmark's cpu_baseline leg load it; the product
 * library never links or calls it.
 *
 * It restates, in plain C, the reference's algorithm for
 *   - the half-edge mesh (Mesh.cu:24-126),
 *   - single-source bounded ICH window propagation (ICHDevice.cu:67-132,
 *     336-480, 525-581, 618-979),
 *   - its fixed-capacity containers with their exact tie / capacity rules
 *     (PriorityQueue.cuh:51-123, MyHashTable.cuh:51-117,
 *     PriorityQueueWithHandle.cuh:53-135),
 *   - SVG row extraction (SVG.cu:34-76),
 *   - the A* / Dijkstra graph solvers (SVG.cu:241-251, 354-413) and
 *   - surface-point sources and the two-point query (ICHDevice.cu:56-65,
 *     93-107, 143-334, 583-615; SVG.cu:253-352).
 *
 * Decisions that the reference leaves open (it cannot build open meshes and
 * leaves some fields undefined):
 *   D1  a boundary half-edge has twin -1 (the reference leaves 0, Mesh.cu:62).
 *   D2  a boundary vertex's firstEdge is its outgoing half-edge whose prev has
 *       no twin, so that the e -> next(twin(e)) fan covers every face.
 *   D4  the window-dependent acos (ICHDevice.cu:819) uses o_acos below, an
 *       implementation of the fdlibm method (Sun Microsystems, e_acos.c)
 *       that the device code implements with the same operations; all other
 *       acos calls use libm, as in the reference.
 *   D8  a boundary vertex is never a pseudo-source (its saddle tests report
 *       "not a saddle"), so fan sweeps never leave the mesh.
 *   D9  a boundary source also reaches the 1-ring vertex across its incoming
 *       boundary edge, and a 1-ring vertex reached over an outgoing boundary
 *       half-edge e records e itself as its enter edge; the backtrace treats
 *       verts[1](enterEdge) == vertex as a step along e.  Only open meshes are
 *       affected.
 *   X   row fields nextToSrcX / nextToDstX whose edge is -1 are 0.0 (the
 *       reference leaves them uninitialised, SVG.cu:49-50).
 * Closed meshes never meet D1, D2, D8 or D9.
 *
 * Parity pins (tests/test_oracle.py): mesh arrays by closed forms
 * (tetrahedron, flat grid); ICH distances by the planar closed form
 * (Euclidean) and the saddle-cone closed form; neighbour sets by the planar
 * and cone closed forms; heaps and hash tables by brute force; the solvers
 * by Bellman-Ford; o_acos against libm within 1 ulp.  Point sources
 * (tests/test_oracle_point.py): Euclidean on a flat mesh, the cone closed
 * form, the K-stop count; the two-point query: Euclidean direct answers with
 * key points on the segment, the through-a-vertex closed form on a flat mesh
 * with a complete graph, and the graph route as the pinned MSAD joined with
 * the second run.  The K-stop point and heap tie order on curved meshes (and
 * which stale pseudo-sources the point-source guard lets through) are pinned
 * only by following the cited lines (the reference cannot be built here: it
 * needs the CUDA headers).
 *
 * Build: oracle/Makefile, gcc -O2 -ffp-contract=off -fopenmp.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define O_PI 3.141592653589793238          /* Geometry.cuh:18 */
#define O_RELATIVE_ERROR 1e-8              /* ICHDevice.cuh:10 */
#define O_WINPQ_SIZE 1024                  /* SVG.cuh:13 */
#define O_PSEUDOWINPQ_SIZE 1024            /* SVG.cuh:14 */
#define O_NONE 0xFFFFFFFFu
#define O_STORED_WIN 100                   /* STORED_WIN_BUF_SIZE (SVG.cuh:15); the reference never sets
                                              storedWindowsSize (ICHDevice.cuh:150) */

/* ------------------------------------------------------------------------ */
/* acos, fdlibm method (D4)                                                  */
/* The coefficients and reductions follow fdlibm's e_acos.c, whose notice   */
/* is preserved as its licence asks:                                        */
/*   Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.      */
/*   Developed at SunPro, a Sun Microsystems, Inc. business.                */
/*   Permission to use, copy, modify, and distribute this software is       */
/*   freely granted, provided that this notice is preserved.                */
/* ------------------------------------------------------------------------ */
static const double o_pi = 3.14159265358979311600e+00;
static const double o_pio2_hi = 1.57079632679489655800e+00;
static const double o_pio2_lo = 6.12323399573676603587e-17;
static const double o_pS0 = 1.66666666666666657415e-01;
static const double o_pS1 = -3.25565818622400915405e-01;
static const double o_pS2 = 2.01212532134862925881e-01;
static const double o_pS3 = -4.00555345006794114027e-02;
static const double o_pS4 = 7.91534994289814532176e-04;
static const double o_pS5 = 3.47933107596021167570e-05;
static const double o_qS1 = -2.40339491173441421878e+00;
static const double o_qS2 = 2.02094576023350569471e+00;
static const double o_qS3 = -6.88283971605453293030e-01;
static const double o_qS4 = 7.70381505559019352791e-02;

static double o_ratR(double z)
{
    double p = z * (o_pS0 + z * (o_pS1 + z * (o_pS2 + z * (o_pS3 + z * (o_pS4 + z * o_pS5)))));
    double q = 1.0 + z * (o_qS1 + z * (o_qS2 + z * (o_qS3 + z * o_qS4)));
    return p / q;
}

double o_acos(double x)
{
    uint64_t bits;
    memcpy(&bits, &x, 8);
    uint32_t ix = (uint32_t)(bits >> 32) & 0x7fffffffu;
    if (ix >= 0x3ff00000u) {                       /* |x| >= 1 */
        uint32_t lo = (uint32_t)bits;
        if (((ix - 0x3ff00000u) | lo) == 0) {
            if ((int64_t)bits > 0) return 0.0;
            return o_pi + 2.0 * o_pio2_lo;
        }
        return (x - x) / (x - x);
    }
    if (ix < 0x3fe00000u) {                        /* |x| < 0.5 */
        if (ix <= 0x3c600000u) return o_pio2_hi + o_pio2_lo;
        double z = x * x;
        double r = o_ratR(z);
        return o_pio2_hi - (x - (o_pio2_lo - x * r));
    }
    if ((int64_t)bits < 0) {                       /* x <= -0.5 */
        double z = (1.0 + x) * 0.5;
        double s = sqrt(z);
        double r = o_ratR(z);
        double w = r * s - o_pio2_lo;
        return o_pi - 2.0 * (s + w);
    }
    {                                              /* x >= 0.5 */
        double z = (1.0 - x) * 0.5;
        double s = sqrt(z);
        uint64_t sb;
        memcpy(&sb, &s, 8);
        sb &= 0xFFFFFFFF00000000ull;
        double df;
        memcpy(&df, &sb, 8);
        double c = (z - df * df) / (s + df);
        double r = o_ratR(z);
        double w = r * s + c;
        return 2.0 * (df + w);
    }
}

/* ------------------------------------------------------------------------ */
/* Mesh (Mesh.cuh:6-52, Mesh.cu:24-126)                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
    int v0, v1;          /* verts[0], verts[1] */
    int twin, next, prev, face;
    double len;
} OEdge;

typedef struct {
    double x, y, z;
    double angle;
    int firstEdge;
    unsigned char boundary;
    unsigned char not_saddle;   /* angle < 2*PI  (ICHDevice.cu:558, 766), or boundary (D8) */
    unsigned char saddle;       /* angle > 2*PI  (ICHDevice.cu:688), and not boundary (D8) */
} OVert;

typedef struct {
    int nv, nf, ne;
    OEdge *e;
    OVert *v;
    int *fverts;   /* 3 per face */
} OMesh;

static double o_len3(double ax, double ay, double az, double bx, double by, double bz)
{
    double x = ax - bx, y = ay - by, z = az - bz;
    return sqrt(x * x + y * y + z * z);
}

typedef struct { uint64_t key; int val; } OPair;

/* twin map: the reference keeps the first half-edge of each sorted vertex
 * pair in a std::map and pairs every later one with it (Mesh.cu:79-89). */
static int o_build_twins(OMesh *m)
{
    int ne = m->ne;
    size_t cap = 1;
    while (cap < (size_t)ne * 2) cap <<= 1;
    OPair *tab = (OPair *)malloc(cap * sizeof(OPair));
    if (!tab) return -1;
    for (size_t i = 0; i < cap; ++i) tab[i].val = -1;
    for (int i = 0; i < ne; ++i) m->e[i].twin = -1;                /* D1 */
    for (int i = 0; i < ne; ++i) {
        uint64_t a = (uint64_t)(uint32_t)m->e[i].v0, b = (uint64_t)(uint32_t)m->e[i].v1;
        if (a > b) { uint64_t t = a; a = b; b = t; }
        uint64_t key = (a << 32) | b;
        uint64_t h = key * 0x9E3779B97F4A7C15ull;
        size_t s = (size_t)(h >> 20) & (cap - 1);
        for (;;) {
            if (tab[s].val == -1) { tab[s].key = key; tab[s].val = i; break; }
            if (tab[s].key == key) {
                int j = tab[s].val;
                m->e[i].twin = j;
                m->e[j].twin = i;
                break;
            }
            s = (s + 1) & (cap - 1);
        }
    }
    free(tab);
    return 0;
}

void oracle_mesh_free(OMesh *m)
{
    if (!m) return;
    free(m->e);
    free(m->v);
    free(m->fverts);
    free(m);
}

OMesh *oracle_mesh_new(const double *xyz, int nv, const int *tri, int nf)
{
    OMesh *m = (OMesh *)calloc(1, sizeof(OMesh));
    if (!m) return NULL;
    m->nv = nv; m->nf = nf; m->ne = 3 * nf;
    m->e = (OEdge *)calloc((size_t)m->ne, sizeof(OEdge));
    m->v = (OVert *)calloc((size_t)nv, sizeof(OVert));
    m->fverts = (int *)malloc(sizeof(int) * 3 * (size_t)nf);
    if (!m->e || !m->v || !m->fverts) { oracle_mesh_free(m); return NULL; }
    for (int i = 0; i < nv; ++i) {
        m->v[i].x = xyz[3 * i]; m->v[i].y = xyz[3 * i + 1]; m->v[i].z = xyz[3 * i + 2];
        m->v[i].angle = 0.0; m->v[i].firstEdge = -1;
    }
    for (int i = 0; i < nf; ++i) {
        for (int j = 0; j < 3; ++j) {
            int v0 = tri[3 * i + j], v1 = tri[3 * i + (j + 1) % 3];
            OEdge *e = &m->e[3 * i + j];
            e->v0 = v0; e->v1 = v1;
            e->len = o_len3(m->v[v0].x, m->v[v0].y, m->v[v0].z, m->v[v1].x, m->v[v1].y, m->v[v1].z);
            e->next = i * 3 + (j + 1) % 3;
            e->prev = i * 3 + (j + 2) % 3;
            e->face = i;
            m->fverts[3 * i + j] = tri[3 * i + j];
        }
    }
    if (o_build_twins(m) != 0) { oracle_mesh_free(m); return NULL; }
    /* angle sums and firstEdge = largest outgoing index (Mesh.cu:106-123) */
    for (int i = 0; i < m->ne; ++i) {
        double l0 = m->e[i].len, l1 = m->e[m->e[i].next].len, l2 = m->e[m->e[i].prev].len;
        double c = (l0 * l0 + l2 * l2 - l1 * l1) / (2.0 * l0 * l2);
        if (c > 1.0) c = 1.0; else if (c < -1.0) c = -1.0;
        c = acos(c);
        m->v[m->e[i].v0].angle += c;
        m->v[m->e[i].v0].firstEdge = i;
    }
    /* boundary vertices (D1) and their firstEdge (D2) */
    for (int i = 0; i < m->ne; ++i) {
        if (m->e[i].twin == -1) {
            m->v[m->e[i].v0].boundary = 1;
            m->v[m->e[i].v1].boundary = 1;
        }
    }
    for (int i = 0; i < m->ne; ++i) {
        int v = m->e[i].v0;
        if (m->v[v].boundary && m->e[m->e[i].prev].twin == -1) m->v[v].firstEdge = i;
    }
    for (int i = 0; i < nv; ++i) {
        OVert *v = &m->v[i];
        v->not_saddle = (unsigned char)(v->boundary || v->angle < 2.0 * O_PI);
        v->saddle = (unsigned char)(!v->boundary && v->angle > 2.0 * O_PI);
    }
    return m;
}

int oracle_mesh_counts(const OMesh *m, int *out3)
{
    out3[0] = m->nv; out3[1] = m->nf; out3[2] = m->ne;
    return 0;
}

/* Export the half-edge arrays, for comparison with the product's builder. */
int oracle_mesh_export(const OMesh *m, int *twin, double *len, double *angle,
                       int *firstEdge, unsigned char *boundary)
{
    for (int i = 0; i < m->ne; ++i) { twin[i] = m->e[i].twin; len[i] = m->e[i].len; }
    for (int i = 0; i < m->nv; ++i) {
        angle[i] = m->v[i].angle; firstEdge[i] = m->v[i].firstEdge; boundary[i] = m->v[i].boundary;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* ICH records (ICHDevice.cuh:16-86)                                         */
/* ------------------------------------------------------------------------ */
typedef struct {
    unsigned edgeID;
    double b0, b1, d0, d1;
    double pseudoSrcDist, minDist;
    unsigned srcID, pseudoSrcId;
    int pseudoSrcBirthTime;
    int level;
} OWindow;

typedef struct {
    unsigned vertID;
    double dist;
    unsigned srcId, pseudoSrcId;
    unsigned pseudoBirthTime;
    unsigned level;
} OPseudo;

typedef struct { double dist, x; unsigned pseudoSrcId; } OSplit;
typedef struct { signed char birthTime; double dist; unsigned char isSource; int enterEdge;
                 unsigned pseudoSrcId, srcId; } OVInfo;

typedef struct { OSplit item; int index; } OSplitItem;
typedef struct { OVInfo item; int index; } OVItem;

/* ------------------------------------------------------------------------ */
/* MyHashTable (MyHashTable.cuh:51-117): linear probing, overflow slot.      */
/* ------------------------------------------------------------------------ */
static unsigned o_hash(unsigned index, unsigned size)
{
    double p = index * 0.618033989;
    return (unsigned)(int)((p - (int)p) * size);
}

#define O_DEFINE_TABLE(NAME, ITEM)                                              \
    typedef struct { ITEM *data; unsigned size; int overflowed; } NAME;         \
    static unsigned NAME##_getKey(NAME *t, unsigned index)                      \
    {                                                                           \
        unsigned key = o_hash(index, t->size);                                  \
        unsigned i = 0;                                                         \
        for (; i < t->size; ++i) {                                              \
            if ((unsigned)t->data[key].index == index || t->data[key].index == -1) break; \
            key = (key + 1) % t->size;                                          \
        }                                                                       \
        if (i == t->size) { key = t->size; t->overflowed = 1; }                 \
        return key;                                                             \
    }

O_DEFINE_TABLE(OSplitTab, OSplitItem)
O_DEFINE_TABLE(OVertTab, OVItem)

static OSplit *split_w(OSplitTab *t, unsigned idx)          /* operator[] */
{
    unsigned k = OSplitTab_getKey(t, idx);
    t->data[k].index = (int)idx;
    return &t->data[k].item;
}
static const OSplit *split_r(OSplitTab *t, unsigned idx)    /* get() */
{
    return &t->data[OSplitTab_getKey(t, idx)].item;
}
static OVInfo *vert_w(OVertTab *t, unsigned idx)
{
    unsigned k = OVertTab_getKey(t, idx);
    t->data[k].index = (int)idx;
    return &t->data[k].item;
}
static const OVInfo *vert_r(OVertTab *t, unsigned idx)
{
    return &t->data[OVertTab_getKey(t, idx)].item;
}

/* ------------------------------------------------------------------------ */
/* PriorityQueues (PriorityQueue.cuh:51-123): 1-based binary min-heap.       */
/* ------------------------------------------------------------------------ */
#define O_DEFINE_HEAP(NAME, T)                                                  \
    typedef struct { double key; T item; } NAME##Item;                          \
    typedef struct { NAME##Item *a; int tail, maxSize; long drops; int maxLen; } NAME; \
    static void NAME##_push(NAME *q, const T *item, double key)                 \
    {                                                                           \
        if (q->tail - 1 >= q->maxSize) { q->drops++; return; }                  \
        ++q->tail;                                                              \
        int cur = q->tail - 1, p = cur / 2;                                     \
        while (p > 0 && q->a[p].key > key) { q->a[cur] = q->a[p]; cur = p; p = cur / 2; } \
        q->a[cur].item = *item;                                                 \
        q->a[cur].key = key;                                                    \
        if (q->tail - 1 > q->maxLen) q->maxLen = q->tail - 1;                   \
    }                                                                           \
    static void NAME##_pop(NAME *q)                                             \
    {                                                                           \
        --q->tail;                                                              \
        NAME##Item top = q->a[q->tail];                                         \
        int cur = 1, l = 2, r = 3, mc = -1;                                     \
        if (r < q->tail) mc = q->a[l].key < q->a[r].key ? l : r;                \
        else if (l < q->tail) mc = l;                                           \
        while (mc != -1 && q->a[mc].key < top.key) {                            \
            q->a[cur] = q->a[mc];                                               \
            cur = mc; l = cur * 2; r = cur * 2 + 1;                             \
            if (r < q->tail) mc = q->a[l].key < q->a[r].key ? l : r;            \
            else if (l < q->tail) mc = l;                                       \
            else mc = -1;                                                       \
        }                                                                       \
        q->a[cur] = top;                                                        \
    }

O_DEFINE_HEAP(OWinQ, OWindow)
O_DEFINE_HEAP(OPseudoQ, OPseudo)

/* ------------------------------------------------------------------------ */
/* Event counters (one set per source).                                      */
/* ------------------------------------------------------------------------ */
enum {
    C_WIN_PUSH = 0, C_WIN_POP, C_WIN_STALE, C_PSEUDO_PUSH, C_PSEUDO_POP, C_PSEUDO_STALE,
    C_PROPAGATE, C_CHILD_BUILT, C_CHILD_VALID, C_SPLIT_WRITE, C_VERT_WRITE,
    C_FAN_HALFEDGE, C_PSEUDO_NEIGH, C_WIN_DROP, C_PSEUDO_DROP, C_MAX_WINQ, C_MAX_PSEUDOQ,
    C_REACHED, C_TOTAL_CALC, C_BT_HOPS, C_ARCS, C_VERT_OVERFLOW, C_SPLIT_OVERFLOW,
    C_VERT_KEYS, C_SPLIT_KEYS, C_LEVEL_SKIP, C_ERRORS, C_ITER,
    C_WINQ_LEFT, C_PSEUDOQ_LEFT,      /* queue lengths when Execute returned */
    C_TOTAL_PREV,                     /* totalCalcVertNum before the last main pop */
    C_NCOUNTERS
};

int oracle_num_counters(void) { return C_NCOUNTERS; }

/* ------------------------------------------------------------------------ */
/* ICH engine                                                                */
/* ------------------------------------------------------------------------ */
typedef struct {
    const OMesh *mesh;
    OSplitTab split;
    OVertTab vert;
    OWinQ winQ;
    OPseudoQ pseudoQ;
    unsigned sourceVert;
    int totalCalcVertNum;
    int pathPassVert;
    long long *cnt;
    long hopCap;
    /* surface-point source (ICHDevice.cuh:148-151) */
    int srcFace;                      /* sourcePointFace, -1: vertex source */
    double spos[3];                   /* sourcePointPos */
    int keptFace;                     /* keptFaces[0], -1: none */
    OWindow stored[O_STORED_WIN];     /* storedWindows */
    int nStored;
    int stalled;                      /* the queue tops were unordered (NaN keys) */
    /* optional record of Execute's choices (ICHDevice.cu:88, 110): per main
     * pop the kind taken (0 window, 1 pseudo-window) and both queue tops'
     * keys (DBL_MAX for an empty queue) */
    int *trKind; double *trW, *trP; int trCap, trN;
} OIch;

/* The queues exactly as the build sizes them (SVG.cu:19-22): arrays of
 * WINPQ_SIZE / PSEUDOWINPQ_SIZE items, one of them the unused slot 0, so
 * push drops once WINPQ_SIZE - 1 items are queued (PriorityQueue.cuh:53). */
static int winq_init(OWinQ *q)
{
    q->a = (OWinQItem *)malloc(sizeof(OWinQItem) * O_WINPQ_SIZE);
    q->maxSize = O_WINPQ_SIZE - 1;
    q->tail = 1; q->drops = 0; q->maxLen = 0;
    return q->a != NULL;
}
static int pseudoq_init(OPseudoQ *q)
{
    q->a = (OPseudoQItem *)malloc(sizeof(OPseudoQItem) * O_PSEUDOWINPQ_SIZE);
    q->maxSize = O_PSEUDOWINPQ_SIZE - 1;
    q->tail = 1; q->drops = 0; q->maxLen = 0;
    return q->a != NULL;
}

static void ich_init(OIch *s, const OMesh *m, int K, unsigned splitCoef, unsigned vertCoef)
{
    memset(s, 0, sizeof(*s));
    s->mesh = m;
    s->split.size = (unsigned)K * splitCoef;
    s->vert.size = (unsigned)K * vertCoef;
    s->split.data = (OSplitItem *)malloc(sizeof(OSplitItem) * (s->split.size + 1));
    s->vert.data = (OVItem *)malloc(sizeof(OVItem) * (s->vert.size + 1));
    winq_init(&s->winQ);
    pseudoq_init(&s->pseudoQ);
    s->hopCap = 4L * ((long)m->nf + (long)m->nv) + 16;
}

static void ich_destroy(OIch *s)
{
    free(s->split.data); free(s->vert.data); free(s->winQ.a); free(s->pseudoQ.a);
}

/* ICHDevice::Clear (ICHDevice.cu:487-523) */
static void ich_clear(OIch *s)
{
    s->winQ.tail = 1; s->pseudoQ.tail = 1;
    s->winQ.drops = 0; s->pseudoQ.drops = 0; s->winQ.maxLen = 0; s->pseudoQ.maxLen = 0;
    for (unsigned i = 0; i < s->split.size + 1; ++i) {
        s->split.data[i].item.dist = DBL_MAX;
        s->split.data[i].item.x = DBL_MAX;
        s->split.data[i].item.pseudoSrcId = O_NONE;
        s->split.data[i].index = -1;
    }
    for (unsigned i = 0; i < s->vert.size + 1; ++i) {
        s->vert.data[i].item.birthTime = -1;
        s->vert.data[i].item.dist = DBL_MAX;
        s->vert.data[i].item.enterEdge = -1;
        s->vert.data[i].item.isSource = 0;
        s->vert.data[i].item.pseudoSrcId = O_NONE;
        s->vert.data[i].item.srcId = O_NONE;
        s->vert.data[i].index = -1;
    }
    s->split.overflowed = 0; s->vert.overflowed = 0;
    s->sourceVert = O_NONE;
    s->srcFace = -1;
    s->keptFace = -1;
    s->nStored = 0;
    s->stalled = 0;
    s->totalCalcVertNum = 0;
}

/* Window::calcMinDist (ICHDevice.cuh:30-39) */
static void win_calc_min(OWindow *w)
{
    double wLen = w->b1 - w->b0;
    double xProj = (w->d0 * w->d0 + wLen * wLen - w->d1 * w->d1) / (2 * wLen);
    if (xProj < 0.0) w->minDist = w->d0 + w->pseudoSrcDist;
    else if (xProj > wLen) w->minDist = w->d1 + w->pseudoSrcDist;
    else w->minDist = sqrt(fabs(w->d0 * w->d0 - xProj * xProj)) + w->pseudoSrcDist;
}

/* Window::FlatenedSrc (ICHDevice.cuh:41-49) */
static void win_flat_src(const OWindow *w, double *sx, double *sy)
{
    double wLen = w->b1 - w->b0;
    double x = (w->d0 * w->d0 + wLen * wLen - w->d1 * w->d1) / (2.0 * wLen);
    double y = sqrt(fabs(w->d0 * w->d0 - x * x));
    x += w->b0;
    *sx = x; *sy = y;
}

static double o_len2(double x, double y) { return sqrt(x * x + y * y); }

/* ICHDevice::Intersect (ICHDevice.cu:975-979) */
static double o_intersect(double v0x, double v0y, double v1x, double v1y,
                          double p0x, double p0y, double p1x, double p1y)
{
    double a00 = p0x - p1x, a01 = v1x - v0x, b0 = v1x - p1x;
    double a10 = p0y - p1y, a11 = v1y - v0y, b1 = v1y - p1y;
    return (b0 * a11 - b1 * a01) / (a00 * a11 - a10 * a01);
}

/* ICHDevice::BuildWindow (ICHDevice.cu:957-973) */
static void build_window(OIch *s, const OWindow *fa, unsigned edge, double t0, double t1,
                         double v0x, double v0y, double v1x, double v1y, OWindow *w)
{
    double sx, sy;
    win_flat_src(fa, &sx, &sy);
    double L = s->mesh->e[edge].len;
    w->edgeID = edge;
    w->b0 = (1 - t0) * L;
    w->b1 = (1 - t1) * L;
    {
        double px = t0 * v0x, py = t0 * v0y;
        double qx = (1 - t0) * v1x, qy = (1 - t0) * v1y;
        double mx = px + qx, my = py + qy;
        w->d0 = o_len2(sx - mx, sy - my);
    }
    {
        double px = t1 * v0x, py = t1 * v0y;
        double qx = (1 - t1) * v1x, qy = (1 - t1) * v1y;
        double mx = px + qx, my = py + qy;
        w->d1 = o_len2(sx - mx, sy - my);
    }
    w->pseudoSrcDist = fa->pseudoSrcDist;
    win_calc_min(w);
    w->srcID = fa->srcID;
    w->pseudoSrcId = fa->pseudoSrcId;
    w->pseudoSrcBirthTime = fa->pseudoSrcBirthTime;
    w->level = fa->level + 1;
}

/* ratio test a > b && a / b - 1.0 > tol */
static int o_beaten(double a, double b, double tol) { return a > b && a / b - 1.0 > tol; }

/* ICHDevice::IsValidWindow (ICHDevice.cu:906-955) */
static int is_valid_window(OIch *s, const OWindow *w, int isLeft)
{
    const OMesh *m = s->mesh;
    if (w->b1 <= w->b0) return 0;
    unsigned e = w->edgeID;
    unsigned v1 = (unsigned)m->e[e].v0;
    unsigned v2 = (unsigned)m->e[e].v1;
    unsigned v3 = (unsigned)m->e[m->e[e].next].v1;
    double l0 = m->e[e].len;
    double l1 = m->e[m->e[e].next].len;
    double l2 = m->e[m->e[m->e[e].next].next].len;
    double p3x = (l2 * l2 + l0 * l0 - l1 * l1) / (2.0 * l0);
    double p3y = sqrt(fabs(l2 * l2 - p3x * p3x));
    double ax = w->b0, bx = w->b1;
    double sx, sy;
    win_flat_src(w, &sx, &sy);
    double dv1 = vert_r(&s->vert, v1)->dist;
    double dv2 = vert_r(&s->vert, v2)->dist;
    double dv3 = vert_r(&s->vert, v3)->dist;
    double srcB = w->pseudoSrcDist + o_len2(sx - bx, sy - 0.0);
    double srcA = w->pseudoSrcDist + o_len2(sx - ax, sy - 0.0);
    if (o_beaten(srcB, dv1 + w->b1, 0.0)) return 0;
    if (o_beaten(srcA, dv2 + l0 - w->b0, 0.0)) return 0;
    if (isLeft) {
        if (o_beaten(srcA, dv3 + o_len2(p3x - ax, p3y - 0.0), 0.0)) return 0;
    } else {
        if (o_beaten(srcB, dv3 + o_len2(p3x - bx, p3y - 0.0), O_RELATIVE_ERROR)) return 0;
    }
    return 1;
}

static void push_win(OIch *s, const OWindow *w)
{
    s->cnt[C_WIN_PUSH]++;
    if (s->winQ.tail - 1 >= s->winQ.maxSize) s->cnt[C_WIN_DROP]++;
    OWinQ_push(&s->winQ, w, w->minDist);
}

static void push_pseudo(OIch *s, const OPseudo *p)
{
    s->cnt[C_PSEUDO_PUSH]++;
    if (s->pseudoQ.tail - 1 >= s->pseudoQ.maxSize) s->cnt[C_PSEUDO_DROP]++;
    OPseudoQ_push(&s->pseudoQ, p, p->dist);
}

/* ICHDevice::Initialize, vertex-source branch (ICHDevice.cu:527-581) + D9 */
static void ich_initialize(OIch *s)
{
    const OMesh *m = s->mesh;
    unsigned src = s->sourceVert;
    int startEdge = m->v[src].firstEdge;
    /* decision I: a vertex that no face uses keeps firstEdge -1 (Mesh.cuh:22)
     * and starts no window (the reference would read edges[-1]) */
    if (startEdge == -1) return;
    int cur = startEdge;
    do {
        unsigned opEdge = (unsigned)m->e[cur].next;
        OWindow w;
        memset(&w, 0, sizeof(w));
        w.edgeID = opEdge;
        w.b0 = 0.0; w.b1 = m->e[opEdge].len;
        w.d0 = m->e[cur].len;
        w.d1 = m->e[m->e[cur].prev].len;
        w.pseudoSrcDist = 0.0; win_calc_min(&w);
        w.srcID = src; w.pseudoSrcId = src;
        w.pseudoSrcBirthTime = 0;
        w.level = 0;
        push_win(s, &w);

        unsigned opVert = (unsigned)m->e[cur].v1;
        if (m->e[cur].len > vert_w(&s->vert, opVert)->dist) {
            cur = m->e[cur].twin;
            if (cur != -1) cur = m->e[cur].next;
            continue;
        }
        OVInfo *vi = vert_w(&s->vert, opVert);
        vi->birthTime = 0;
        vi->dist = m->e[cur].len;
        vi->enterEdge = m->e[cur].twin != -1 ? m->e[cur].twin : cur;      /* D9 */
        vi->srcId = src; vi->pseudoSrcId = src;
        if (m->v[opVert].not_saddle) {
            cur = m->e[cur].twin;
            if (cur != -1) cur = m->e[cur].next;
            continue;
        }
        OPseudo p;
        p.vertID = opVert; p.dist = m->e[cur].len;
        p.srcId = src; p.pseudoSrcId = src;
        p.pseudoBirthTime = (unsigned)(int)vert_r(&s->vert, opVert)->birthTime;
        p.level = 0;
        push_pseudo(s, &p);
        cur = m->e[cur].twin;
        if (cur != -1) cur = m->e[cur].next;
    } while (cur != startEdge && cur != -1);
    if (cur == -1) {                                   /* D9: boundary source */
        int inE = m->e[startEdge].prev;
        unsigned w = (unsigned)m->e[inE].v0;
        if (!(m->e[inE].len > vert_w(&s->vert, w)->dist)) {
            OVInfo *vi = vert_w(&s->vert, w);
            vi->birthTime = 0;
            vi->dist = m->e[inE].len;
            vi->enterEdge = inE;
            vi->srcId = src; vi->pseudoSrcId = src;
        }
    }
    OVInfo *sv = vert_w(&s->vert, src);
    sv->birthTime = 0;
    sv->dist = 0.0;
    sv->enterEdge = -1;
    sv->isSource = 1;
    sv->srcId = src; sv->pseudoSrcId = src;
}

/* ICHDevice::PropagateWindow (ICHDevice.cu:618-726) */
static void propagate_window(OIch *s, const OWindow *win)
{
    const OMesh *m = s->mesh;
    int e0i = m->e[win->edgeID].twin;
    if (e0i == -1) return;
    s->cnt[C_PROPAGATE]++;
    unsigned e0 = (unsigned)e0i;
    unsigned e1 = (unsigned)m->e[e0].next;
    unsigned e2 = (unsigned)m->e[e1].next;
    unsigned opVert = (unsigned)m->e[e1].v1;
    double sx, sy;
    win_flat_src(win, &sx, &sy);
    double leftx = win->b0, rightx = win->b1;
    double l0 = m->e[e0].len, l1 = m->e[e1].len, l2 = m->e[e2].len;
    double v0x = 0.0, v0y = 0.0, v1x = l0, v1y = 0.0;
    double v2x = (l1 * l1 + l0 * l0 - l2 * l2) / (2.0 * l0);
    double v2y = -sqrt(fabs(l1 * l1 - v2x * v2x));
    double interX = v2x - v2y * (v2x - sx) / (v2y - sy);
    OWindow lw, rw;
    int hasL = 1, hasR = 1;
    if (interX <= leftx) {
        hasL = 0;
        double t0 = o_intersect(sx, sy, leftx, 0.0, v2x, v2y, v1x, v1y);
        double t1 = o_intersect(sx, sy, rightx, 0.0, v2x, v2y, v1x, v1y);
        build_window(s, win, e2, t0, t1, v2x, v2y, v1x, v1y, &rw);
        s->cnt[C_CHILD_BUILT]++;
        if (!is_valid_window(s, &rw, 0)) hasR = 0;
    } else if (interX >= rightx) {
        hasR = 0;
        double t0 = o_intersect(sx, sy, leftx, 0.0, v0x, v0y, v2x, v2y);
        double t1 = o_intersect(sx, sy, rightx, 0.0, v0x, v0y, v2x, v2y);
        build_window(s, win, e1, t0, t1, v0x, v0y, v2x, v2y, &lw);
        s->cnt[C_CHILD_BUILT]++;
        if (!is_valid_window(s, &lw, 1)) hasL = 0;
    } else {
        double directDist = o_len2(v2x - sx, v2y - sy);
        double cand = directDist + win->pseudoSrcDist;
        const OSplit *sp = split_r(&s->split, e0);
        if (cand > sp->dist && cand / sp->dist - 1.0 > O_RELATIVE_ERROR) {
            hasL = sp->x < interX;
            hasR = !hasL;
        } else {
            if (cand < split_r(&s->split, e0)->dist) {
                OSplit *w = split_w(&s->split, e0);
                w->dist = cand;
                w->pseudoSrcId = win->pseudoSrcId;
                w->x = l0 - interX;
                s->cnt[C_SPLIT_WRITE]++;
            }
            if (cand < vert_r(&s->vert, opVert)->dist) {
                if (vert_r(&s->vert, opVert)->dist == DBL_MAX) ++s->totalCalcVertNum;
                OVInfo *vi = vert_w(&s->vert, opVert);
                ++vi->birthTime;
                vi->dist = cand;
                vi->enterEdge = (int)e0;
                vi->srcId = win->srcID; vi->pseudoSrcId = win->pseudoSrcId;
                s->cnt[C_VERT_WRITE]++;
                if (m->v[opVert].saddle) {
                    OPseudo p;
                    p.vertID = opVert; p.dist = vert_r(&s->vert, opVert)->dist;
                    p.srcId = win->srcID; p.pseudoSrcId = win->pseudoSrcId;
                    p.pseudoBirthTime = (unsigned)(int)vert_r(&s->vert, opVert)->birthTime;
                    p.level = (unsigned)(win->level + 1);
                    push_pseudo(s, &p);
                }
            }
        }
        if (hasL) {
            double t0 = o_intersect(sx, sy, leftx, 0.0, v0x, v0y, v2x, v2y);
            build_window(s, win, e1, t0, 0.0, v0x, v0y, v2x, v2y, &lw);
            s->cnt[C_CHILD_BUILT]++;
            if (!is_valid_window(s, &lw, 1)) hasL = 0;
        }
        if (hasR) {
            double t1 = o_intersect(sx, sy, rightx, 0.0, v2x, v2y, v1x, v1y);
            build_window(s, win, e2, 1.0, t1, v2x, v2y, v1x, v1y, &rw);
            s->cnt[C_CHILD_BUILT]++;
            if (!is_valid_window(s, &rw, 0)) hasR = 0;
        }
    }
    if (hasL) { s->cnt[C_CHILD_VALID]++; push_win(s, &lw); }
    if (hasR) { s->cnt[C_CHILD_VALID]++; push_win(s, &rw); }
}

static double corner_angle(const OMesh *m, double L0, double L1, double L2)
{
    double c = (L0 * L0 + L1 * L1 - L2 * L2) / (2.0 * L0 * L1);
    if (c > 1.0) c = 1.0; else if (c < -1.0) c = -1.0;
    (void)m;
    return acos(c);
}

/* ICHDevice::GenSubWinsForPseudoSrcFromWindow (ICHDevice.cu:795-858) */
static void sub_from_window(OIch *s, const OPseudo *pw, int *startEdge, int *endEdge)
{
    const OMesh *m = s->mesh;
    int e0 = vert_r(&s->vert, pw->vertID)->enterEdge;
    int e1 = m->e[e0].next;
    int e2 = m->e[e1].next;
    double l0 = m->e[e0].len, l1 = m->e[e1].len, l2 = m->e[e2].len;
    double epx = l0 - split_r(&s->split, (unsigned)e0)->x, epy = 0.0;
    double v0x = 0.0, v0y = 0.0, v1x = l0, v1y = 0.0;
    double v2x = (l1 * l1 + l0 * l0 - l2 * l2) / (2.0 * l0);
    double v2y = -sqrt(fabs(l1 * l1 - v2x * v2x));
    double dx = epx - v2x, dy = epy - v2y;
    double dlen = o_len2(dx, dy);
    double a0 = (dx * (v0x - v2x) + dy * (v0y - v2y)) / dlen / l1;
    double a1 = (dx * (v1x - v2x) + dy * (v1y - v2y)) / dlen / l2;
    if (a0 > 1.0) a0 = 1.0; else if (a0 < -1.0) a0 = -1.0;
    if (a1 > 1.0) a1 = 1.0; else if (a1 < -1.0) a1 = -1.0;
    double angle0 = o_acos(a0), angle1 = o_acos(a1);
    *startEdge = -1; *endEdge = -1;
    int cur = m->e[e1].twin;
    while (angle0 < O_PI && cur != -1) {
        int opE = m->e[cur].next, nxE = m->e[opE].next;
        angle0 += corner_angle(m, m->e[cur].len, m->e[nxE].len, m->e[opE].len);
        s->cnt[C_FAN_HALFEDGE]++;
        cur = m->e[nxE].twin;
    }
    if (cur != -1) *startEdge = m->e[m->e[cur].twin].next;
    cur = m->e[e2].twin;
    while (angle1 < O_PI && cur != -1) {
        int nxE = m->e[cur].next, opE = m->e[nxE].next;
        angle1 += corner_angle(m, m->e[cur].len, m->e[nxE].len, m->e[opE].len);
        s->cnt[C_FAN_HALFEDGE]++;
        cur = m->e[nxE].twin;
    }
    if (cur != -1) {
        int ee = m->e[m->e[m->e[cur].twin].next].next;
        *endEdge = m->e[ee].twin;
    }
}

/* ICHDevice::GenSubWinsForPseudoSrcFromPseudoSrc (ICHDevice.cu:860-904) */
static void sub_from_pseudo(OIch *s, const OPseudo *pw, int *startEdge, int *endEdge)
{
    const OMesh *m = s->mesh;
    double angle0 = 0.0, angle1 = 0.0;
    *startEdge = -1; *endEdge = -1;
    int enter = vert_r(&s->vert, pw->vertID)->enterEdge;
    int cur = enter;
    while (angle0 < O_PI && cur != -1) {
        int opE = m->e[cur].next, nxE = m->e[opE].next;
        angle0 += corner_angle(m, m->e[cur].len, m->e[nxE].len, m->e[opE].len);
        s->cnt[C_FAN_HALFEDGE]++;
        cur = m->e[nxE].twin;
    }
    if (cur != -1) *startEdge = m->e[m->e[cur].twin].next;
    cur = m->e[enter].twin;
    while (angle1 < O_PI && cur != -1) {
        int nxE = m->e[cur].next, opE = m->e[nxE].next;
        angle1 += corner_angle(m, m->e[cur].len, m->e[nxE].len, m->e[opE].len);
        s->cnt[C_FAN_HALFEDGE]++;
        cur = m->e[nxE].twin;
    }
    if (cur != -1) {
        int ee = m->e[m->e[m->e[cur].twin].next].next;
        *endEdge = m->e[ee].twin;
    }
}

/* ICHDevice::GenSubWinsForPseudoSrc (ICHDevice.cu:728-793) */
static void gen_sub_wins(OIch *s, const OPseudo *pw)
{
    const OMesh *m = s->mesh;
    int startEdge = -1, endEdge = -1;
    const OVInfo *vi = vert_r(&s->vert, pw->vertID);
    if (vi->enterEdge == -1 && vi->birthTime != -1) {
        startEdge = m->v[pw->vertID].firstEdge;
        endEdge = startEdge;
    } else if (vi->enterEdge >= 0 && (unsigned)m->e[vi->enterEdge].v0 == pw->vertID) {
        sub_from_pseudo(s, pw, &startEdge, &endEdge);
    } else if (vi->enterEdge >= 0 && (unsigned)m->e[m->e[vi->enterEdge].next].v1 == pw->vertID) {
        sub_from_window(s, pw, &startEdge, &endEdge);
    } else {
        s->cnt[C_ERRORS]++;          /* reference: assert(false); unreachable under D8 */
        return;
    }
    if (startEdge == -1 || endEdge == -1) { s->cnt[C_ERRORS]++; return; }
    long guard = 0;
    do {
        OWindow w;
        memset(&w, 0, sizeof(w));
        w.edgeID = (unsigned)m->e[startEdge].next;
        w.b0 = 0.0; w.b1 = m->e[w.edgeID].len;
        w.d0 = m->e[startEdge].len;
        w.d1 = m->e[m->e[w.edgeID].next].len;
        w.pseudoSrcDist = pw->dist; win_calc_min(&w);
        w.srcID = pw->srcId; w.pseudoSrcId = pw->vertID;
        w.pseudoSrcBirthTime = (int)pw->pseudoBirthTime;
        w.level = (int)pw->level + 1;
        push_win(s, &w);
        startEdge = m->e[m->e[m->e[startEdge].next].next].twin;
        if (startEdge == -1 || ++guard > s->hopCap) { s->cnt[C_ERRORS]++; break; }
    } while (startEdge != endEdge);

    int start = m->v[pw->vertID].firstEdge;
    int cur = start;
    do {
        unsigned opVert = (unsigned)m->e[cur].v1;
        s->cnt[C_PSEUDO_NEIGH]++;
        if (m->v[opVert].not_saddle ||
            vert_r(&s->vert, opVert)->dist < pw->dist + m->e[cur].len) {
            cur = m->e[cur].twin;
            if (cur != -1) cur = m->e[cur].next;
            continue;
        }
        if (vert_r(&s->vert, opVert)->dist == DBL_MAX) ++s->totalCalcVertNum;
        OVInfo *ov = vert_w(&s->vert, opVert);
        ov->dist = pw->dist + m->e[cur].len;
        ++ov->birthTime;
        ov->enterEdge = m->e[cur].twin;
        ov->srcId = pw->srcId; ov->pseudoSrcId = pw->pseudoSrcId;
        s->cnt[C_VERT_WRITE]++;
        OPseudo c;
        c.vertID = opVert; c.dist = vert_r(&s->vert, opVert)->dist;
        c.srcId = pw->srcId; c.pseudoSrcId = pw->vertID;
        c.pseudoBirthTime = (unsigned)(int)vert_r(&s->vert, opVert)->birthTime;
        c.level = pw->level;
        push_pseudo(s, &c);
        cur = m->e[cur].twin;
        if (cur != -1) cur = m->e[cur].next;
    } while (cur != start && cur != -1);
}

/* Records a main pop of Execute when a trace buffer is attached. */
static void ich_trace(OIch *s, int kind)
{
    if (!s->trKind || s->trN >= s->trCap) return;
    s->trKind[s->trN] = kind;
    s->trW[s->trN] = s->winQ.tail != 1 ? s->winQ.a[1].item.minDist : DBL_MAX;
    s->trP[s->trN] = s->pseudoQ.tail != 1 ? s->pseudoQ.a[1].item.dist : DBL_MAX;
    s->trN++;
}

/* ICHDevice::Execute (ICHDevice.cu:67-132) for a vertex source. */
static void ich_execute(OIch *s, int K)
{
    const OMesh *m = s->mesh;
    ich_initialize(s);
    while (s->winQ.tail != 1 || s->pseudoQ.tail != 1) {
        s->cnt[C_ITER]++;
        s->cnt[C_TOTAL_PREV] = s->totalCalcVertNum;
        while (s->winQ.tail != 1 &&
               s->winQ.a[1].item.pseudoSrcId < (unsigned)m->nv &&
               s->winQ.a[1].item.pseudoSrcBirthTime !=
                   (int)vert_r(&s->vert, s->winQ.a[1].item.pseudoSrcId)->birthTime) {
            OWinQ_pop(&s->winQ);
            s->cnt[C_WIN_POP]++; s->cnt[C_WIN_STALE]++;
        }
        /* the reference guards with winQ.top() here (ICHDevice.cu:84); for a
         * vertex source that guard is always true (D6). */
        while (s->pseudoQ.tail != 1 &&
               s->pseudoQ.a[1].item.pseudoBirthTime !=
                   (unsigned)(int)vert_r(&s->vert, s->pseudoQ.a[1].item.vertID)->birthTime) {
            OPseudoQ_pop(&s->pseudoQ);
            s->cnt[C_PSEUDO_POP]++; s->cnt[C_PSEUDO_STALE]++;
        }
        int wEmpty = s->winQ.tail == 1, pEmpty = s->pseudoQ.tail == 1;
        if (!wEmpty && (pEmpty || s->winQ.a[1].item.minDist < s->pseudoQ.a[1].item.dist)) {
            OWindow win = s->winQ.a[1].item;
            ich_trace(s, 0);
            OWinQ_pop(&s->winQ);
            s->cnt[C_WIN_POP]++;
            if (win.level > m->nf) { s->cnt[C_LEVEL_SKIP]++; continue; }
            propagate_window(s, &win);
        } else if (!pEmpty && (wEmpty || s->winQ.a[1].item.minDist >= s->pseudoQ.a[1].item.dist)) {
            OPseudo pw = s->pseudoQ.a[1].item;
            ich_trace(s, 1);
            OPseudoQ_pop(&s->pseudoQ);
            s->cnt[C_PSEUDO_POP]++;
            if (pw.level >= (unsigned)m->nf) { s->cnt[C_LEVEL_SKIP]++; continue; }
            gen_sub_wins(s, &pw);
        } else {
            /* decision S: the tops compare unordered (a NaN key); the
             * reference would spin on them (ICHDevice.cu:88, 110) */
            s->cnt[C_ERRORS]++;
            break;
        }
        if (K != -1 && s->totalCalcVertNum >= K) break;
    }
    s->cnt[C_WINQ_LEFT] = s->winQ.tail - 1;
    s->cnt[C_PSEUDOQ_LEFT] = s->pseudoQ.tail - 1;
}

/* ICHDevice::BuildGeodesicPathTo(vertId, ...) (ICHDevice.cu:336-480) + D9. */
typedef struct { int srcEdge, dstEdge; double srcX, dstX; unsigned srcId; } OPath;

static void unfold_step(const OMesh *m, int *gid, double *gpos, double *lpx, double *lpy,
                        double *cpx, double *cpy, int *err)
{
    int e0 = m->e[*gid].twin;
    if (e0 == -1) { *err = 1; return; }
    int e1 = m->e[e0].next, e2 = m->e[e1].next;
    double l0 = m->e[e0].len, l1 = m->e[e1].len, l2 = m->e[e2].len;
    double ox = (l0 * l0 + l2 * l2 - l1 * l1) / (2.0 * l0);
    double oy = sqrt(fabs(l2 * l2 - ox * ox));
    /* toLeft(opVert2D, lastPoint, curPoint) (Geometry.cu:306-311) */
    double ax = *cpx - *lpx, ay = *cpy - *lpy;
    double bx = ox - *lpx, by = oy - *lpy;
    double la = o_len2(ax, ay);
    if (!(fabs(la) < 1e-12)) { ax /= la; ay /= la; }
    double lb = o_len2(bx, by);
    if (!(fabs(lb) < 1e-12)) { bx /= lb; by /= lb; }
    int left = (ax * by - ay * bx) > 0.0;
    double f = *gpos / l0;
    if (left) {
        double p0x = (l2 * l2 + l1 * l1 - l0 * l0) / (2.0 * l1);
        double p0y = -sqrt(fabs(l2 * l2 - p0x * p0x));
        double p1x = l1, p1y = 0.0;
        double nlx = f * p0x + (1.0 - f) * p1x, nly = f * p0y + (1.0 - f) * p1y;
        double t = o_intersect(*lpx, *lpy, *cpx, *cpy, l0, 0.0, ox, oy);
        *gpos = (1.0 - t) * l1;
        *gid = e1;
        *cpx = l1 - *gpos; *cpy = 0.0;
        *lpx = nlx; *lpy = nly;
    } else {
        double p0x = 0.0, p0y = 0.0;
        double p1x = (l2 * l2 + l0 * l0 - l1 * l1) / (2.0 * l2);
        double p1y = -sqrt(fabs(l0 * l0 - p1x * p1x));
        double nlx = f * p0x + (1.0 - f) * p1x, nly = f * p0y + (1.0 - f) * p1y;
        double t = o_intersect(*lpx, *lpy, *cpx, *cpy, ox, oy, 0.0, 0.0);
        *gpos = (1.0 - t) * l2;
        *gid = e2;
        *cpx = l2 - *gpos; *cpy = 0.0;
        *lpx = nlx; *lpy = nly;
    }
}

static void build_path_to(OIch *s, unsigned vertId, OPath *out)
{
    const OMesh *m = s->mesh;
    int V = m->nv;
    s->pathPassVert = 0;
    out->srcEdge = -1; out->dstEdge = -1; out->srcX = 0.0; out->dstX = 0.0;
    unsigned curVert = vertId;
    long hops = 0;
    while (vert_r(&s->vert, curVert)->dist != 0.0) {
        if (++hops > s->hopCap) { s->pathPassVert = 1; s->cnt[C_ERRORS]++; return; }
        s->cnt[C_BT_HOPS]++;
        int enter = vert_r(&s->vert, curVert)->enterEdge;
        if (enter == -1) {
            out->srcId = (unsigned)V;
            if (out->dstEdge == -1) { out->dstEdge = m->v[curVert].firstEdge; out->dstX = 0.0; }
            else { out->srcEdge = m->v[curVert].firstEdge; out->srcX = 0.0; }
            s->pathPassVert = 1;
            return;
        } else if ((unsigned)m->e[enter].v0 == curVert) {
            unsigned nextVert = (unsigned)m->e[enter].v1;
            if (!vert_r(&s->vert, nextVert)->isSource) {
                if (out->dstEdge == -1) { out->dstEdge = m->e[enter].next; out->dstX = 0.0; }
                else { out->srcEdge = m->e[enter].next; out->srcX = 0.0; }
                s->pathPassVert = 1;
            }
            curVert = nextVert;
        } else if ((unsigned)m->e[enter].v1 == curVert) {          /* D9 */
            unsigned nextVert = (unsigned)m->e[enter].v0;
            if (!vert_r(&s->vert, nextVert)->isSource) {
                if (out->dstEdge == -1) { out->dstEdge = enter; out->dstX = 0.0; }
                else { out->srcEdge = enter; out->srcX = 0.0; }
                s->pathPassVert = 1;
            }
            curVert = nextVert;
        } else {
            int gid = enter;
            double gpos = split_r(&s->split, (unsigned)enter)->x;
            if (out->dstEdge == -1) { out->dstEdge = gid; out->dstX = gpos; }
            else { out->srcEdge = gid; out->srcX = gpos; }
            int tw = m->e[gid].twin;
            if (tw == -1) { s->pathPassVert = 1; s->cnt[C_ERRORS]++; return; }
            unsigned opVert = (unsigned)m->e[m->e[tw].next].v1;
            double l0 = m->e[gid].len;
            double l1 = m->e[m->e[gid].next].len;
            double l2 = m->e[m->e[m->e[gid].next].next].len;
            double lpx = (l1 * l1 + l0 * l0 - l2 * l2) / (2.0 * l0);
            double lpy = -sqrt(fabs(l1 * l1 - lpx * lpx));
            double cpx = l0 - gpos, cpy = 0.0;
            unsigned ps = split_r(&s->split, (unsigned)enter)->pseudoSrcId;
            if (ps >= (unsigned)V && s->srcFace < 0) {
                /* surface-point pseudo-source in a vertex-source run */
                s->pathPassVert = 1; s->cnt[C_ERRORS]++;
                return;
            }
            /* unfold towards the pseudo-source vertex, or (surface-point
             * source, ICHDevice.cu:415-418) until the face behind the key
             * point is the source's face */
            while (ps < (unsigned)V ? opVert != ps : m->e[tw].face != s->srcFace) {
                if (++hops > s->hopCap) { s->pathPassVert = 1; s->cnt[C_ERRORS]++; return; }
                s->cnt[C_BT_HOPS]++;
                int err = 0;
                unfold_step(m, &gid, &gpos, &lpx, &lpy, &cpx, &cpy, &err);
                if (err) { s->pathPassVert = 1; s->cnt[C_ERRORS]++; return; }
                out->srcEdge = gid; out->srcX = gpos;
                tw = m->e[gid].twin;
                if (tw == -1) { s->pathPassVert = 1; s->cnt[C_ERRORS]++; return; }
                opVert = (unsigned)m->e[m->e[tw].next].v1;
            }
            if (ps >= (unsigned)V) {
                /* reached the source point's face (ICHDevice.cu:465-468) */
                out->srcId = ps;
                return;
            }
            if (!vert_r(&s->vert, opVert)->isSource) {
                out->srcEdge = m->v[opVert].firstEdge; out->srcX = 0.0;
                s->pathPassVert = 1;
            }
            curVert = opVert;
        }
    }
    out->srcId = curVert;
}

/* ------------------------------------------------------------------------ */
/* Surface-point sources (ICHDevice.cu:56-65, 67-132, 143-334, 583-615).     */
/* ------------------------------------------------------------------------ */
static double o_vdist(const OMesh *m, const double *p, unsigned v)
{
    return o_len3(p[0], p[1], p[2], m->v[v].x, m->v[v].y, m->v[v].z);
}

/* ICHDevice::Initialize, surface-point branch (ICHDevice.cu:583-615): one
 * window over each edge of the source face, pseudo-source id V (never stale),
 * and the face's vertices at their straight-line distances. */
static void ich_initialize_point(OIch *s)
{
    const OMesh *m = s->mesh;
    const unsigned V = (unsigned)m->nv;
    const double *p = s->spos;
    for (int j = 0; j < 3; ++j) {
        unsigned opEdge = (unsigned)(3 * s->srcFace + j);      /* faces[f].edges[j] */
        OWindow w;
        memset(&w, 0, sizeof(w));
        w.edgeID = opEdge;
        w.b0 = 0.0; w.b1 = m->e[opEdge].len;
        w.d0 = o_vdist(m, p, (unsigned)m->e[opEdge].v0);
        w.d1 = o_vdist(m, p, (unsigned)m->e[opEdge].v1);
        w.pseudoSrcDist = 0.0; win_calc_min(&w);
        w.srcID = V; w.pseudoSrcId = V;
        w.pseudoSrcBirthTime = 0; w.level = 0;
        push_win(s, &w);

        unsigned opVert = (unsigned)m->e[opEdge].v0;
        if (o_vdist(m, p, opVert) > vert_w(&s->vert, opVert)->dist) continue;
        OVInfo *vi = vert_w(&s->vert, opVert);
        vi->birthTime = 0;
        vi->dist = o_vdist(m, p, opVert);
        vi->enterEdge = -1;
        vi->srcId = V; vi->pseudoSrcId = V;
        if (m->v[opVert].not_saddle) continue;
        OPseudo q;
        q.vertID = opVert;
        q.dist = o_len3(m->v[opVert].x, m->v[opVert].y, m->v[opVert].z, p[0], p[1], p[2]);
        q.srcId = V; q.pseudoSrcId = V;
        q.pseudoBirthTime = (unsigned)(int)vert_r(&s->vert, opVert)->birthTime;
        q.level = 0;
        push_pseudo(s, &q);
    }
}

/* ICHDevice::Execute (ICHDevice.cu:67-132) for a surface-point source, with
 * the windows that cross into the kept face stored as they are popped
 * (ICHDevice.cu:93-107). */
static void ich_execute_point(OIch *s, int K)
{
    const OMesh *m = s->mesh;
    const unsigned V = (unsigned)m->nv;
    ich_initialize_point(s);
    while (s->winQ.tail != 1 || s->pseudoQ.tail != 1) {
        s->cnt[C_ITER]++;
        while (s->winQ.tail != 1 &&
               s->winQ.a[1].item.pseudoSrcId < V &&
               s->winQ.a[1].item.pseudoSrcBirthTime !=
                   (int)vert_r(&s->vert, s->winQ.a[1].item.pseudoSrcId)->birthTime) {
            OWinQ_pop(&s->winQ);
            s->cnt[C_WIN_POP]++; s->cnt[C_WIN_STALE]++;
        }
        /* The stale test of the pseudo-source queue is guarded by the
         * window queue's top (ICHDevice.cu:84), i.e. slot 1 of its array,
         * which keeps the last window popped once the queue is empty.  While
         * that window descends from the point (pseudo-source id V) stale
         * pseudo-sources are not skipped and are expanded. */
        while (s->pseudoQ.tail != 1 &&
               s->winQ.a[1].item.pseudoSrcId < V &&
               s->pseudoQ.a[1].item.pseudoBirthTime !=
                   (unsigned)(int)vert_r(&s->vert, s->pseudoQ.a[1].item.vertID)->birthTime) {
            OPseudoQ_pop(&s->pseudoQ);
            s->cnt[C_PSEUDO_POP]++; s->cnt[C_PSEUDO_STALE]++;
        }
        int wEmpty = s->winQ.tail == 1, pEmpty = s->pseudoQ.tail == 1;
        if (!wEmpty && (pEmpty || s->winQ.a[1].item.minDist < s->pseudoQ.a[1].item.dist)) {
            OWindow win = s->winQ.a[1].item;
            OWinQ_pop(&s->winQ);
            s->cnt[C_WIN_POP]++;
            if (win.level > m->nf) { s->cnt[C_LEVEL_SKIP]++; continue; }
            if (win.minDist != win.minDist) {          /* NaN: see the stall below */
                s->cnt[C_ERRORS]++;
                s->stalled = 1;
                break;
            }
            int tw = m->e[win.edgeID].twin;
            if (tw != -1 && s->keptFace >= 0 && m->e[tw].face == s->keptFace &&
                s->nStored < O_STORED_WIN)
                s->stored[s->nStored++] = win;
            propagate_window(s, &win);
        } else if (!pEmpty && (wEmpty || s->winQ.a[1].item.minDist >= s->pseudoQ.a[1].item.dist)) {
            OPseudo pw = s->pseudoQ.a[1].item;
            OPseudoQ_pop(&s->pseudoQ);
            s->cnt[C_PSEUDO_POP]++;
            if (pw.level >= (unsigned)m->nf) { s->cnt[C_LEVEL_SKIP]++; continue; }
            gen_sub_wins(s, &pw);
        } else {
            /* neither branch: a NaN key.  A point on a vertex of its face
             * puts the source on a window's line and its windows' keys become
             * NaN; the reference then loops or expands NaN windows until
             * their level passes F.  The run stops at the first NaN key. */
            s->cnt[C_ERRORS]++;
            s->stalled = 1;
            break;
        }
        if (K != -1 && s->totalCalcVertNum >= K) break;
    }
}

/* ICHDevice::BuildGeodesicPathTo(faceId, pos, ...) (ICHDevice.cu:143-334):
 * the distance to a point of the kept face through a stored window or one of
 * the face's vertices, and the key points of that path.  out's X fields keep
 * what the caller put there unless a key point is written. */
static double face_path_to(OIch *s, int f, const double *pos, OPath *out)
{
    const OMesh *m = s->mesh;
    const unsigned V = (unsigned)m->nv;
    s->pathPassVert = 0;
    out->srcEdge = -1; out->dstEdge = -1;
    double minDist = DBL_MAX, xInter = 0.0, posx = 0.0, posy = 0.0;
    OWindow minWin;
    memset(&minWin, 0, sizeof(minWin));
    int dstVert = -1, throughWin = 1;
    for (int i = 0; i < s->nStored; ++i) {
        const OWindow *w = &s->stored[i];
        int tw = m->e[w->edgeID].twin;
        if (tw == -1) continue;
        if (m->e[tw].face != f) continue;
        int e0 = tw, e1 = m->e[e0].next, e2 = m->e[e1].next;
        double l0 = m->e[e0].len, l1 = m->e[e1].len, l2 = m->e[e2].len;
        unsigned v0 = (unsigned)m->e[e0].v1, v1 = (unsigned)m->e[e0].v0, v2 = (unsigned)m->e[e1].v1;
        double p0x = 0.0, p0y = 0.0, p1x = l0, p1y = 0.0;
        double p2x = (l1 * l1 + l0 * l0 - l2 * l2) / (2.0 * l0);
        double p2y = -sqrt(fabs(l1 * l1 - p2x * p2x));
        double sx, sy;
        win_flat_src(w, &sx, &sy);
        /* the point's barycentric weights from Heron's formula */
        double a = o_vdist(m, pos, v0), b = o_vdist(m, pos, v1), c = o_vdist(m, pos, v2);
        double s0 = (b + c + l2) / 2.0;
        double s1 = (a + c + l1) / 2.0;
        double s2 = (a + b + l0) / 2.0;
        s0 = sqrt(fabs(s0 * (s0 - b) * (s0 - c) * (s0 - l2)));
        s1 = sqrt(fabs(s1 * (s1 - a) * (s1 - c) * (s1 - l1)));
        s2 = sqrt(fabs(s2 * (s2 - a) * (s2 - b) * (s2 - l0)));
        double w0 = s0 / (s0 + s1 + s2);
        double w1 = s1 / (s0 + s1 + s2);
        double w2 = s2 / (s0 + s1 + s2);
        double cx = w0 * p0x + w1 * p1x + w2 * p2x;
        double cy = w0 * p0y + w1 * p1y + w2 * p2y;
        double xi = sx - (cx - sx) / (cy - sy) * sy;
        if (xi <= 0.0 || xi >= l0) continue;
        double cur;
        if (xi > w->b0 && xi < w->b1)
            cur = o_len2(cx - sx, cy - sy) + w->pseudoSrcDist;
        else if (xi <= w->b0)
            cur = o_len2(cx - w->b0, cy - 0.0) + w->d0 + w->pseudoSrcDist;
        else
            cur = o_len2(cx - w->b1, cy - 0.0) + w->d1 + w->pseudoSrcDist;
        if (cur < minDist) { minDist = cur; minWin = *w; xInter = xi; posx = cx; posy = cy; }
    }
    for (int i = 0; i < 3; ++i) {
        unsigned ov = (unsigned)m->e[3 * f + i].v0;
        double cd = o_vdist(m, pos, ov) + vert_r(&s->vert, ov)->dist;
        if (cd < minDist) { throughWin = 0; dstVert = (int)ov; minDist = cd; }
    }
    if (minDist == DBL_MAX) return minDist;
    if (!throughWin) {
        build_path_to(s, (unsigned)dstVert, out);
        s->pathPassVert = 1;
        out->dstEdge = m->v[dstVert].firstEdge; out->dstX = 0.0;
        return minDist;
    }
    int gid = m->e[minWin.edgeID].twin;
    double gpos = m->e[gid].len - xInter;
    out->dstEdge = gid; out->dstX = gpos;
    int tw = m->e[gid].twin;
    unsigned opVert = (unsigned)m->e[m->e[tw].next].v1;
    double lpx = posx, lpy = posy, cpx = m->e[gid].len - gpos, cpy = 0.0;
    unsigned ps = minWin.pseudoSrcId;
    long hops = 0;
    while (ps < V ? opVert != ps : m->e[tw].face != s->srcFace) {
        if (++hops > s->hopCap) { s->pathPassVert = 1; s->cnt[C_ERRORS]++; return minDist; }
        s->cnt[C_BT_HOPS]++;
        int err = 0;
        unfold_step(m, &gid, &gpos, &lpx, &lpy, &cpx, &cpy, &err);
        if (err) { s->pathPassVert = 1; s->cnt[C_ERRORS]++; return minDist; }
        out->srcEdge = gid; out->srcX = gpos;
        tw = m->e[gid].twin;
        if (tw == -1) { s->pathPassVert = 1; s->cnt[C_ERRORS]++; return minDist; }
        opVert = (unsigned)m->e[m->e[tw].next].v1;
    }
    if (ps < V && !vert_r(&s->vert, opVert)->isSource) {
        /* the path bends at the pseudo-source: the source-side key point is
         * the one next to the source on that vertex's own path */
        s->pathPassVert = 1;
        OPath sub;
        build_path_to(s, opVert, &sub);
        out->srcEdge = sub.srcEdge; out->srcX = sub.srcX;
    }
    out->srcId = minWin.srcID;
    return minDist;
}

/* ------------------------------------------------------------------------ */
/* SVG build (SVG.cu:9-96): rows in the reference's K-stride layout.         */
/* ------------------------------------------------------------------------ */
typedef struct {
    int adj;
    int nextToSrcEdge;
    int nextToDstEdge;
    int pad;
    double geodDist;
    double nextToSrcX;
    double nextToDstX;
} ORow;   /* field set of SVGNode (SVG.cuh:21-27) */

static void count_table_keys(OIch *s)
{
    long long nv = 0, ns = 0;
    for (unsigned i = 0; i < s->vert.size; ++i) nv += s->vert.data[i].index != -1;
    for (unsigned i = 0; i < s->split.size; ++i) ns += s->split.data[i].index != -1;
    s->cnt[C_VERT_KEYS] = nv;
    s->cnt[C_SPLIT_KEYS] = ns;
}

static void svg_one_source(OIch *s, int i, int K, int rowCap, ORow *row, int *tail)
{
    const OMesh *m = s->mesh;
    ich_clear(s);
    s->sourceVert = (unsigned)i;
    ich_execute(s, K);
    s->cnt[C_TOTAL_CALC] = s->totalCalcVertNum;
    s->cnt[C_MAX_WINQ] = s->winQ.maxLen;
    s->cnt[C_MAX_PSEUDOQ] = s->pseudoQ.maxLen;
    int t = 0;
    long long reached = 0;
    for (int j = 0; j < m->nv; ++j) {
        double dj = vert_r(&s->vert, (unsigned)j)->dist;
        if (dj == DBL_MAX) continue;
        ++reached;
        if (j == i) continue;
        if (t == rowCap) continue;            /* keep counting reached vertices */
        OPath p;
        build_path_to(s, (unsigned)j, &p);
        if (s->pathPassVert) continue;
        if (p.srcEdge == -1) { p.srcEdge = p.dstEdge; p.srcX = p.dstX; }
        ORow *r = &row[t];
        r->adj = j;
        r->geodDist = dj;
        r->nextToSrcEdge = p.srcEdge; r->nextToSrcX = p.srcX;
        r->nextToDstEdge = p.dstEdge; r->nextToDstX = p.dstX;
        r->pad = 0;
        ++t;
    }
    *tail = t;
    s->cnt[C_REACHED] = reached;
    s->cnt[C_ARCS] = t;
    s->cnt[C_VERT_OVERFLOW] = s->vert.overflowed;
    s->cnt[C_SPLIT_OVERFLOW] = s->split.overflowed;
    count_table_keys(s);
}

/*
 * Build rows for sources [s0, s1).  K is the stop count and row capacity of
 * the reference (SVG.cu:39, 75); row_cap may exceed K (closed-form tests use
 * it to see every neighbour).  rows: (s1-s0)*row_cap ORow; tails: s1-s0;
 * counters: (s1-s0)*C_NCOUNTERS (may be NULL).  Returns wall seconds.
 */
double oracle_svg_build(const OMesh *m, int K, int row_cap, unsigned splitCoef, unsigned vertCoef,
                        int s0, int s1, int nthreads, ORow *rows, int *tails, long long *counters)
{
    if (row_cap <= 0) row_cap = K;
    if (s0 < 0 || s1 > m->nv || s0 > s1) return -1.0;
    int tabK = K > 0 ? K : m->nv;
    struct timespec ta, tb;
    clock_gettime(CLOCK_MONOTONIC, &ta);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        OIch s;
        ich_init(&s, m, tabK, splitCoef, vertCoef);
        long long local[C_NCOUNTERS];
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int i = s0; i < s1; ++i) {
            memset(local, 0, sizeof(local));
            s.cnt = counters ? counters + (size_t)(i - s0) * C_NCOUNTERS : local;
            memset(s.cnt, 0, sizeof(long long) * C_NCOUNTERS);
            svg_one_source(&s, i, K, row_cap, rows + (size_t)(i - s0) * row_cap, tails + (i - s0));
        }
        ich_destroy(&s);
    }
    clock_gettime(CLOCK_MONOTONIC, &tb);
    return (double)(tb.tv_sec - ta.tv_sec) + 1e-9 * (double)(tb.tv_nsec - ta.tv_nsec);
}

/* Single-source ICH: distances (DBL_MAX when unreached), enter edges and
 * birth times for every vertex, after Execute(K) (K = -1: run to the end). */
int oracle_ich_single(const OMesh *m, int src, int K, unsigned splitCoef, unsigned vertCoef,
                      double *dist, int *enterEdge, long long *counters)
{
    if (src < 0 || src >= m->nv) return -1;
    OIch s;
    int tabK = K > 0 ? K : m->nv;
    ich_init(&s, m, tabK, splitCoef, vertCoef);
    long long local[C_NCOUNTERS];
    s.cnt = counters ? counters : local;
    memset(s.cnt, 0, sizeof(long long) * C_NCOUNTERS);
    ich_clear(&s);
    s.sourceVert = (unsigned)src;
    ich_execute(&s, K);
    s.cnt[C_TOTAL_CALC] = s.totalCalcVertNum;
    s.cnt[C_MAX_WINQ] = s.winQ.maxLen;
    s.cnt[C_MAX_PSEUDOQ] = s.pseudoQ.maxLen;
    s.cnt[C_VERT_OVERFLOW] = s.vert.overflowed;
    s.cnt[C_SPLIT_OVERFLOW] = s.split.overflowed;
    long long reached = 0;
    for (int j = 0; j < m->nv; ++j) {
        const OVInfo *vi = vert_r(&s.vert, (unsigned)j);
        dist[j] = vi->dist;
        if (enterEdge) enterEdge[j] = vi->enterEdge;
        reached += vi->dist != DBL_MAX;
    }
    s.cnt[C_REACHED] = reached;
    count_table_keys(&s);
    ich_destroy(&s);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Graph solvers (SVG.cu:241-251, 354-413) with PriorityQueuesWithHandle.    */
/* ------------------------------------------------------------------------ */
typedef struct { double key; int item; int *bh; } OHItem;
typedef struct { OHItem *a; int tail, maxSize; } OHeap;

static void oh_push(OHeap *q, int item, int *bh, double key)
{
    if (q->tail - 1 >= q->maxSize) return;
    ++q->tail;
    int cur = q->tail - 1, p = cur / 2;
    while (p > 0 && q->a[p].key > key) {
        q->a[cur] = q->a[p];
        *(q->a[cur].bh) = cur;
        cur = p; p = cur / 2;
    }
    q->a[cur].item = item; q->a[cur].key = key; q->a[cur].bh = bh;
    *bh = cur;
}

static int oh_pop(OHeap *q)
{
    int ret = q->a[1].item;
    *(q->a[1].bh) = -1;
    --q->tail;
    OHItem top = q->a[q->tail];
    int cur = 1, l = 2, r = 3, mc = -1;
    if (r < q->tail) mc = q->a[l].key < q->a[r].key ? l : r;
    else if (l < q->tail) mc = l;
    while (mc != -1 && q->a[mc].key < top.key) {
        q->a[cur] = q->a[mc];
        *(q->a[cur].bh) = cur;
        cur = mc; l = cur * 2; r = cur * 2 + 1;
        if (r < q->tail) mc = q->a[l].key < q->a[r].key ? l : r;
        else if (l < q->tail) mc = l;
        else mc = -1;
    }
    q->a[cur] = top;
    *(q->a[cur].bh) = cur;
    return ret;
}

static void oh_decrease(OHeap *q, int handle, double key)
{
    int cur = handle, p = cur / 2;
    OHItem it = q->a[cur];
    while (p > 0 && q->a[p].key > key) {
        q->a[cur] = q->a[p];
        *(q->a[cur].bh) = cur;
        cur = p; p = cur / 2;
    }
    q->a[cur] = it;
    q->a[cur].key = key;
    *(q->a[cur].bh) = cur;
}

/* Graph in the reference layout: rows K-stride + tails, positions xyz[3V]. */
typedef struct {
    int V, K;
    const int *adj;        /* V*K */
    const double *w;       /* V*K */
    const int *tails;      /* V */
    const double *xyz;     /* 3V */
} OGraph;

static double o_pdist(const double *xyz, int a, int b)
{
    return o_len3(xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2], xyz[3 * b], xyz[3 * b + 1], xyz[3 * b + 2]);
}

/* SVG::Astar (SVG.cu:369-413); astar != 0 is SEARCHTYPE ASTAR. */
static void o_astar(const OGraph *g, int t, int astar, double *dist, int *parent, int *srcId,
                    int *idx, OHeap *q, long long *stat)
{
    while (q->tail != 1) {
        int u = oh_pop(q);
        if (stat) stat[0]++;
        if (astar && u == t) { q->tail = 1; break; }
        const int *adj = g->adj + (size_t)g->K * u;
        const double *w = g->w + (size_t)g->K * u;
        for (int i = 0; i < g->tails[u]; ++i) {
            int v = adj[i];
            double nd = dist[u] + w[i];
            if (stat) stat[1]++;
            if (nd >= dist[v]) continue;
            if (stat) stat[2]++;
            dist[v] = nd;
            parent[v] = u;
            srcId[v] = srcId[u];
            double pr = astar ? nd + o_pdist(g->xyz, v, t) : nd;
            if (idx[v] == -1) oh_push(q, v, &idx[v], pr);
            else oh_decrease(q, idx[v], pr);
        }
    }
}

static int o_solver_alloc(int V, OHeap *q, int **idx)
{
    q->a = (OHItem *)malloc(sizeof(OHItem) * ((size_t)V + 1));
    q->tail = 1; q->maxSize = V;
    *idx = (int *)malloc(sizeof(int) * (size_t)V);
    return (q->a && *idx) ? 0 : -1;
}

/* SVG::SolveSSSD(s, t) (SVG.cu:241-251). stat: [settled, arcs scanned, improving]. */
double oracle_sssd(int V, int K, const int *adj, const double *w, const int *tails,
                   const double *xyz, int s, int t, double *dist, int *parent, int *srcId,
                   long long *stat)
{
    OGraph g = {V, K, adj, w, tails, xyz};
    OHeap q; int *idx;
    if (o_solver_alloc(V, &q, &idx)) return -1.0;
    for (int i = 0; i < V; ++i) { dist[i] = DBL_MAX; parent[i] = -1; srcId[i] = -1; idx[i] = -1; }
    dist[s] = 0.0; parent[s] = -1; srcId[s] = s;
    oh_push(&q, s, &idx[s], 0.0 + o_pdist(xyz, s, t));
    o_astar(&g, t, 1, dist, parent, srcId, idx, &q, stat);
    free(q.a); free(idx);
    return dist[t];
}

/* SVG::SolveMSAD (SVG.cu:354-367). */
int oracle_msad(int V, int K, const int *adj, const double *w, const int *tails,
                const double *xyz, const int *sources, const double *weights, int ns,
                double *dist, int *parent, int *srcId, long long *stat)
{
    OGraph g = {V, K, adj, w, tails, xyz};
    OHeap q; int *idx;
    if (o_solver_alloc(V, &q, &idx)) return -1;
    for (int i = 0; i < V; ++i) { dist[i] = DBL_MAX; parent[i] = -1; srcId[i] = -1; idx[i] = -1; }
    for (int i = 0; i < ns; ++i) {
        dist[sources[i]] = weights[i];
        parent[sources[i]] = -1;
        srcId[sources[i]] = sources[i];
        oh_push(&q, sources[i], &idx[sources[i]], weights[i]);
    }
    o_astar(&g, -1, 0, dist, parent, srcId, idx, &q, stat);
    free(q.a); free(idx);
    return 0;
}

/* Single ICH run from a point of face f (K = -1: to the end): distances
 * (DBL_MAX when unreached) and enter edges of every vertex. */
int oracle_ich_point(const OMesh *m, int f, const double *p, int K, unsigned splitCoef,
                     unsigned vertCoef, double *dist, int *enterEdge, long long *counters)
{
    if (f < 0 || f >= m->nf) return -1;
    OIch *s = (OIch *)malloc(sizeof(OIch));
    if (!s) return -1;
    int tabK = K > 0 ? K : m->nv;
    ich_init(s, m, tabK, splitCoef, vertCoef);
    long long local[C_NCOUNTERS];
    s->cnt = counters ? counters : local;
    memset(s->cnt, 0, sizeof(long long) * C_NCOUNTERS);
    ich_clear(s);
    s->srcFace = f;
    memcpy(s->spos, p, sizeof(s->spos));
    ich_execute_point(s, K);
    s->cnt[C_TOTAL_CALC] = s->totalCalcVertNum;
    s->cnt[C_MAX_WINQ] = s->winQ.maxLen;
    s->cnt[C_MAX_PSEUDOQ] = s->pseudoQ.maxLen;
    for (int j = 0; j < m->nv; ++j) {
        const OVInfo *vi = vert_r(&s->vert, (unsigned)j);
        dist[j] = vi->dist;
        enterEdge[j] = vi->enterEdge;
    }
    ich_destroy(s);
    free(s);
    return 0;
}

/*
 * SVG::SolveSSSD(f0, p0, f1, p1) (SVG.cu:253-352): geodesic distance between
 * two surface points.  An ICH from p0 keeps the windows crossing into f1; if
 * the path through them (or through a vertex of f1) records a source-side
 * key point, that distance is the answer (route 0).  Otherwise a Dijkstra on
 * the graph seeded with every vertex the p0 ICH reached, at its ICH distance,
 * is joined with an ICH from p1 at the vertex minimising the sum (route 1).
 *
 * dout[3] = {geodDist, nextToSrcX, nextToDstX};
 * iout[4] = {nextToSrcEdge, nextToDstEdge, saddleVertNearDst, route};
 * gdist / gparent (V, may be NULL) receive the graph stage's distances and
 * parents (route 1).  route -1: no vertex joins the two (distance DBL_MAX).
 *
 * Decisions where the reference is undefined: SVGNode X fields start at 0.0;
 * a key-point edge without a twin is kept as it is when the reference would
 * take its twin; no join vertex gives route -1 (the reference backtraces
 * vertex 0xFFFFFFFF).  Returns -2 when a run stalls on unordered keys (a
 * point on a vertex of its face), where the reference loops.
 */
int oracle_sssd_point(const OMesh *m, int K, unsigned splitCoef, unsigned vertCoef,
                      int Kg, const int *adj, const double *w, const int *tails,
                      int f0, const double *p0, int f1, const double *p1,
                      double *dout, int *iout, double *gdist, int *gparent, long long *stat)
{
    const int V = m->nv;
    if (f0 < 0 || f0 >= m->nf || f1 < 0 || f1 >= m->nf || K <= 0) return -1;
    OIch *a = (OIch *)malloc(sizeof(OIch));
    OIch *b = (OIch *)malloc(sizeof(OIch));
    double *gd = gdist ? gdist : (double *)malloc(sizeof(double) * (size_t)V);
    int *gp = gparent ? gparent : (int *)malloc(sizeof(int) * (size_t)V);
    int *sid = (int *)malloc(sizeof(int) * (size_t)V);
    OHeap q; int *idx = NULL;
    if (!a || !b || !gd || !gp || !sid || o_solver_alloc(V, &q, &idx)) return -1;
    long long cnt[C_NCOUNTERS];
    ich_init(a, m, K, splitCoef, vertCoef);
    ich_init(b, m, K, splitCoef, vertCoef);
    a->cnt = cnt; b->cnt = cnt;
    memset(cnt, 0, sizeof(cnt));

    double resD = 0.0, resSX = 0.0, resDX = 0.0;
    int resSE = -1, resDE = -1, last = -1, route = 0, rcode = 0;

    ich_clear(a);
    a->srcFace = f0; memcpy(a->spos, p0, sizeof(a->spos));
    a->keptFace = f1;
    ich_execute_point(a, K);
    if (a->stalled) { rcode = -2; goto done; }
    OPath fp;
    fp.srcX = resSX; fp.dstX = resDX; fp.srcEdge = fp.dstEdge = -1; fp.srcId = 0;
    resD = face_path_to(a, f1, p1, &fp);
    resSE = fp.srcEdge; resSX = fp.srcX; resDE = fp.dstEdge; resDX = fp.dstX;

    for (int i = 0; i < V; ++i) { gd[i] = DBL_MAX; gp[i] = -1; sid[i] = -1; idx[i] = -1; }
    if (resSE == -1) {
        route = 1;
        /* seeds in the p0 table's slot order (SVG.cu:283-291) */
        for (unsigned i = 0; i < a->vert.size; ++i) {
            int vi = a->vert.data[i].index;
            if (a->vert.data[i].item.dist == DBL_MAX) continue;
            gd[vi] = a->vert.data[i].item.dist;
            gp[vi] = -1;
            sid[vi] = vi;
            oh_push(&q, vi, &idx[vi], a->vert.data[i].item.dist);
        }
        OGraph g = {V, Kg, adj, w, tails, NULL};
        o_astar(&g, -1, 0, gd, gp, sid, idx, &q, stat);

        ich_clear(b);
        b->srcFace = f1; memcpy(b->spos, p1, sizeof(b->spos));
        ich_execute_point(b, K);
        if (b->stalled) { rcode = -2; goto done; }
        double minDist = DBL_MAX;
        for (unsigned i = 0; i < b->vert.size; ++i) {
            int vi = b->vert.data[i].index;
            if (b->vert.data[i].item.dist == DBL_MAX) continue;
            if (gd[vi] + b->vert.data[i].item.dist >= minDist) continue;
            minDist = gd[vi] + b->vert.data[i].item.dist;
            last = vi;
        }
        resD = minDist;
        if (last < 0) {
            route = -1;
            resSE = resDE = -1; resSX = resDX = 0.0;
        } else {
            /* the p1 side: its source-side key point becomes the result's
             * destination-side one (SVG.cu:309-318) */
            OPath pb;
            build_path_to(b, (unsigned)last, &pb);
            int nDE = pb.srcEdge;
            double nDX = pb.srcX;
            resSE = pb.dstEdge; resSX = pb.dstX;
            if (nDE == -1) { nDE = resSE; nDX = resDX; }   /* reads nextToDstX, as SVG.cu:314 does */
            if (nDE != -1 && m->e[nDE].twin != -1) {
                nDE = m->e[nDE].twin;
                nDX = m->e[nDE].len - nDX;
            }
            /* the p0 side from the first vertex of the graph path (SVG.cu:320-351);
             * the p0 run is the one above (a second run is identical) */
            int first = last;
            while (gp[first] != -1) first = gp[first];
            if (vert_r(&a->vert, (unsigned)first)->dist != DBL_MAX) {
                OPath pa;
                build_path_to(a, (unsigned)first, &pa);
                int nSE = pa.srcEdge;
                double nSX = pa.srcX;
                resDE = pa.dstEdge; resDX = pa.dstX;
                if (nSE == -1) { nSE = resDE; nSX = resDX; }
                resSE = nSE; resSX = nSX;
                resDE = nDE; resDX = nDX;
            }
        }
    }
    dout[0] = resD; dout[1] = resSX; dout[2] = resDX;
    iout[0] = resSE; iout[1] = resDE; iout[2] = last; iout[3] = route;
done:
    ich_destroy(a); ich_destroy(b); free(a); free(b);
    if (!gdist) free(gd);
    if (!gparent) free(gp);
    free(sid); free(q.a); free(idx);
    return rcode;
}

/* Many SSSD queries, timed, for the CPU baseline: one query per thread. */
double oracle_sssd_many(int V, int K, const int *adj, const double *w, const int *tails,
                        const double *xyz, const int *st, int nq, int nthreads, double *out)
{
    struct timespec ta, tb;
    clock_gettime(CLOCK_MONOTONIC, &ta);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        double *dist = (double *)malloc(sizeof(double) * (size_t)V);
        int *parent = (int *)malloc(sizeof(int) * (size_t)V);
        int *sid = (int *)malloc(sizeof(int) * (size_t)V);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int i = 0; i < nq; ++i)
            out[i] = oracle_sssd(V, K, adj, w, tails, xyz, st[2 * i], st[2 * i + 1], dist, parent, sid, NULL);
        free(dist); free(parent); free(sid);
    }
    clock_gettime(CLOCK_MONOTONIC, &tb);
    return (double)(tb.tv_sec - ta.tv_sec) + 1e-9 * (double)(tb.tv_nsec - ta.tv_nsec);
}

/* Self-tests of the containers against brute force; returns 0 on success. */
int oracle_selftest_heap(const double *keys, int n, int *order_out)
{
    OWinQ q;
    q.a = (OWinQItem *)malloc(sizeof(OWinQItem) * ((size_t)n + 1));
    q.tail = 1; q.maxSize = n; q.drops = 0; q.maxLen = 0;
    for (int i = 0; i < n; ++i) {
        OWindow w; memset(&w, 0, sizeof(w));
        w.edgeID = (unsigned)i; w.minDist = keys[i];
        OWinQ_push(&q, &w, keys[i]);
    }
    for (int i = 0; i < n; ++i) {
        order_out[i] = (int)q.a[1].item.edgeID;
        OWinQ_pop(&q);
    }
    free(q.a);
    return 0;
}

int oracle_selftest_hash(unsigned size, const unsigned *keys, int n, int *slot_out)
{
    OVertTab t;
    t.size = size; t.overflowed = 0;
    t.data = (OVItem *)malloc(sizeof(OVItem) * (size + 1));
    for (unsigned i = 0; i < size + 1; ++i) { t.data[i].index = -1; t.data[i].item.dist = DBL_MAX; }
    for (int i = 0; i < n; ++i) {
        unsigned k = OVertTab_getKey(&t, keys[i]);
        t.data[k].index = (int)keys[i];
        slot_out[i] = (int)k;
    }
    int ov = t.overflowed;
    free(t.data);
    return ov;
}

/* The window queue as the build configures it (winq_init): pushes n keys
 * (item i carries edge id i), then pops everything; order_out receives the
 * ids in pop order, *kept how many were queued, *dropped how many pushes
 * were refused. */
int oracle_selftest_winq(const double *keys, int n, int *order_out, int *kept, int *dropped)
{
    OWinQ q;
    if (!winq_init(&q)) return -1;
    for (int i = 0; i < n; ++i) {
        OWindow w; memset(&w, 0, sizeof(w));
        w.edgeID = (unsigned)i; w.minDist = keys[i];
        OWinQ_push(&q, &w, keys[i]);
    }
    *kept = q.tail - 1;
    *dropped = (int)q.drops;
    for (int i = 0; q.tail != 1; ++i) {
        order_out[i] = (int)q.a[1].item.edgeID;
        OWinQ_pop(&q);
    }
    free(q.a);
    return 0;
}

/* Same for the pseudo-window queue. */
int oracle_selftest_pseudoq(const double *keys, int n, int *order_out, int *kept, int *dropped)
{
    OPseudoQ q;
    if (!pseudoq_init(&q)) return -1;
    for (int i = 0; i < n; ++i) {
        OPseudo p; memset(&p, 0, sizeof(p));
        p.vertID = (unsigned)i; p.dist = keys[i];
        OPseudoQ_push(&q, &p, keys[i]);
    }
    *kept = q.tail - 1;
    *dropped = (int)q.drops;
    for (int i = 0; q.tail != 1; ++i) {
        order_out[i] = (int)q.a[1].item.vertID;
        OPseudoQ_pop(&q);
    }
    free(q.a);
    return 0;
}

/* Vertex distances of the ICH state for the selftests below: vids[i] gets
 * dist vd[i] (DBL_MAX leaves it unreached). */
static void preset_verts(OIch *s, int nv, const int *vids, const double *vd)
{
    for (int i = 0; i < nv; ++i)
        if (vd[i] != DBL_MAX) vert_w(&s->vert, (unsigned)vids[i])->dist = vd[i];
}

/* IsValidWindow (ICHDevice.cu:906-955) on one window of edge `edge`:
 * (b0, b1, d0, d1, ps) as the Window fields, vertex distances preset for
 * the vertices vids[0..nv). Returns 1 valid, 0 rejected, -1 bad input. */
int oracle_selftest_valid(const OMesh *m, int edge, double b0, double b1, double d0, double d1,
                          double ps, int nv, const int *vids, const double *vd, int isLeft)
{
    if (edge < 0 || edge >= m->ne) return -1;
    OIch s;
    long long cnt[C_NCOUNTERS];
    ich_init(&s, m, 8, 10, 2);
    s.cnt = cnt;
    memset(cnt, 0, sizeof(cnt));
    ich_clear(&s);
    preset_verts(&s, nv, vids, vd);
    OWindow w; memset(&w, 0, sizeof(w));
    w.edgeID = (unsigned)edge;
    w.b0 = b0; w.b1 = b1; w.d0 = d0; w.d1 = d1; w.pseudoSrcDist = ps;
    win_calc_min(&w);
    int r = is_valid_window(&s, &w, isLeft);
    ich_destroy(&s);
    return r;
}

/* PropagateWindow (ICHDevice.cu:618-726) of one window of edge `edge` from
 * source vertex `src`: vertex distances preset as above, the split record of
 * twin(edge) preset to (split_dist, split_x) unless split_dist is DBL_MAX.
 * Outputs: child_edge[2], child_b[4] (b0, b1 per child) of the windows it
 * queued in pop order, *nchild their number, counters[C_NCOUNTERS],
 * out_state[3] = the far vertex's distance and the split record (dist, x)
 * afterwards.  Returns 0, or -1 on bad input. */
int oracle_selftest_propagate(const OMesh *m, int src, int edge, double b0, double b1, double d0,
                              double d1, double ps, int nv, const int *vids, const double *vd,
                              double split_dist, double split_x, int *child_edge, double *child_b,
                              int *nchild, long long *counters, double *out_state)
{
    if (edge < 0 || edge >= m->ne || m->e[edge].twin == -1) return -1;
    OIch s;
    ich_init(&s, m, 8, 10, 2);
    s.cnt = counters;
    memset(counters, 0, sizeof(long long) * C_NCOUNTERS);
    ich_clear(&s);
    s.sourceVert = (unsigned)src;
    preset_verts(&s, nv, vids, vd);
    unsigned e0 = (unsigned)m->e[edge].twin;
    if (split_dist != DBL_MAX) {
        OSplit *sp = split_w(&s.split, e0);
        sp->dist = split_dist; sp->x = split_x; sp->pseudoSrcId = (unsigned)src;
    }
    OWindow w; memset(&w, 0, sizeof(w));
    w.edgeID = (unsigned)edge;
    w.b0 = b0; w.b1 = b1; w.d0 = d0; w.d1 = d1; w.pseudoSrcDist = ps;
    w.srcID = (unsigned)src; w.pseudoSrcId = (unsigned)src;
    win_calc_min(&w);
    propagate_window(&s, &w);
    *nchild = 0;
    while (s.winQ.tail != 1 && *nchild < 2) {
        child_edge[*nchild] = (int)s.winQ.a[1].item.edgeID;
        child_b[2 * *nchild] = s.winQ.a[1].item.b0;
        child_b[2 * *nchild + 1] = s.winQ.a[1].item.b1;
        ++*nchild;
        OWinQ_pop(&s.winQ);
    }
    out_state[0] = vert_r(&s.vert, (unsigned)m->e[m->e[e0].next].v1)->dist;
    out_state[1] = split_r(&s.split, e0)->dist;
    out_state[2] = split_r(&s.split, e0)->x;
    ich_destroy(&s);
    return 0;
}

/* Execute(K) from vertex src with Execute's choices recorded (ich_trace):
 * kind[i], wkey[i], pkey[i] for the first cap main pops; returns how many
 * were recorded (-1 bad input).  counters as oracle_ich_single. */
int oracle_ich_trace(const OMesh *m, int src, int K, int cap, int *kind, double *wkey, double *pkey,
                     long long *counters)
{
    if (src < 0 || src >= m->nv || cap < 0) return -1;
    OIch s;
    int tabK = K > 0 ? K : m->nv;
    ich_init(&s, m, tabK, 10, 2);
    s.cnt = counters;
    memset(counters, 0, sizeof(long long) * C_NCOUNTERS);
    ich_clear(&s);
    s.trKind = kind; s.trW = wkey; s.trP = pkey; s.trCap = cap; s.trN = 0;
    s.sourceVert = (unsigned)src;
    ich_execute(&s, K);
    counters[C_TOTAL_CALC] = s.totalCalcVertNum;
    int n = s.trN;
    ich_destroy(&s);
    return n;
}

/* Vectorised o_acos for the tests. */
int oracle_acos_many(const double *x, double *y, int n)
{
    for (int i = 0; i < n; ++i) y[i] = o_acos(x[i]);
    return 0;
}
