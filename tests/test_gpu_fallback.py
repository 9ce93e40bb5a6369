"""GPU tests of the build's fallback and failure paths, each against the
oracle at the same settings: batched table capacity overflow (recomputed on
the plain path), reference-table overflow with its aliasing overflow slot
(MyHashTable.cuh:104-109), chunked builds, large K beyond the batched
vertex-table cap, every lanes-per-wave shape, a mesh with a non-manifold
vertex (fans walked, not read by position), and the serving loop's memory
admission.  Bit-exact like the other build tests."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import svg_amd as S
from svg_amd import _lib as L
from svg_amd import meshgen as mg
from tests.helpers import gpu_rows

pytestmark = pytest.mark.gpu

SF_DROP, SF_REF_OVERFLOW, SF_CAPACITY, SF_ERROR, SF_REDONE = 1, 4, 8, 16, 32


def _build(xyz, tri, K, mode, split_coef=10, vert_coef=2, **knobs):
    m = S.Mesh.from_arrays(xyz, tri)
    g = S.SVG()
    g.AssignMesh(m)
    g.SetParameters(K, split_coef, vert_coef)
    g.ConstructSVG(mode=mode, **knobs)
    return m, g


def _oracle(xyz, tri, K, split_coef=10, vert_coef=2):
    om = O.Mesh(xyz, tri)
    rows, tails, cnt, _ = O.svg_build(om, K, split_coef=split_coef, vert_coef=vert_coef, nthreads=4)
    return rows, tails, {n: cnt[:, i] for i, n in enumerate(O.COUNTERS)}


def _assert_rows(g, rows, tails, K):
    t, h = gpu_rows(g)
    assert np.array_equal(t, tails), f"{int(np.sum(t != tails))} row lengths differ"
    mask = np.arange(K)[None, :] < tails[:, None]
    for gk, ok in (("adj", "adj"), ("dist", "geodDist"), ("src_edge", "nextToSrcEdge"),
                   ("dst_edge", "nextToDstEdge")):
        a, b = h[gk][mask], rows[ok][mask]
        assert np.array_equal(a, b), f"{gk}: {int(np.sum(a != b))} entries differ"
    for gk, ok, ek in (("src_x", "nextToSrcX", "nextToSrcEdge"), ("dst_x", "nextToDstX", "nextToDstEdge")):
        mm = mask & (rows[ek] != -1)
        assert np.array_equal(h[gk][mm], rows[ok][mm]), gk


TORUS = lambda: mg.bumpy_torus(nu=40, nv=24)           # noqa: E731  (960 vertices)
SPHERE = lambda: mg.noisy_sphere(seed=4, splits4=3, centroid_split=False)   # noqa: E731  (642)


@pytest.mark.parametrize("caps", [dict(vert_cap=64), dict(split_cap=64), dict(vert_cap=128, split_cap=128)])
def test_capacity_overflow_is_recomputed_exactly(caps):
    # Tables too small for some sources: they stop early, are flagged
    # (SF_CAPACITY) and rebuilt on the plain path, which keeps the
    # reference's tables; every row must still equal the oracle's.
    xyz, tri = TORUS()
    K = 80
    print(f"case capacity {caps} V={len(xyz)} K={K}", flush=True)
    rows, tails, _ = _oracle(xyz, tri, K)
    _, g = _build(xyz, tri, K, S.SVG_MODE_BATCHED, **caps)
    st, stt = g.stats(), g.status()
    assert st["redone"] > 0
    flagged = (stt & SF_CAPACITY) > 0
    assert flagged.sum() == st["redone"]
    assert np.all(stt[flagged] & SF_REDONE)
    assert not np.any(stt & SF_ERROR)
    _assert_rows(g, rows, tails, K)


@pytest.mark.parametrize("coefs", [(10, 1), (1, 2), (2, 1)])
@pytest.mark.parametrize("mode", [S.SVG_MODE_PLAIN, S.SVG_MODE_BATCHED])
def test_reference_table_overflow_aliases_like_the_reference(coefs, mode):
    # splitInfoCoef / vertInfoCoef below the reference's 10 / 2: the
    # reference tables (K*coef slots) fill up and every further key aliases
    # the overflow slot (MyHashTable.cuh:104-109).  The oracle restates that
    # aliasing; the plain path replicates the tables, the batched path flags
    # the sources (SF_REF_OVERFLOW) and recomputes them on the plain path.
    split_coef, vert_coef = coefs
    xyz, tri = SPHERE()
    K = 20
    print(f"case ref overflow coefs {coefs} mode {mode}", flush=True)
    rows, tails, c = _oracle(xyz, tri, K, split_coef, vert_coef)
    assert np.any(c["vert_overflow"] | c["split_overflow"]), "the case must overflow"
    _, g = _build(xyz, tri, K, mode, split_coef, vert_coef)
    _assert_rows(g, rows, tails, K)
    if mode == S.SVG_MODE_BATCHED:
        stt = g.status()
        over = (c["vert_overflow"] | c["split_overflow"]) > 0
        assert np.all(stt[over] & SF_REF_OVERFLOW) and np.all(stt[over] & SF_REDONE)
        assert g.stats()["redone"] >= int(over.sum())


@pytest.mark.parametrize("chunk", [100, 215, 641])
def test_chunked_build(chunk):
    # Fewer resident source tables than sources: the build runs in chunks
    # (the path taken when the tables do not fit in device memory at once).
    xyz, tri = SPHERE()
    K = 60
    print(f"case chunk {chunk}", flush=True)
    rows, tails, _ = _oracle(xyz, tri, K)
    _, g = _build(xyz, tri, K, S.SVG_MODE_BATCHED, chunk=chunk)
    assert g.stats()["chunk"] == chunk
    _assert_rows(g, rows, tails, K)


def test_chunked_build_with_redo():
    xyz, tri = TORUS()
    K = 80
    print("case chunk + capacity", flush=True)
    rows, tails, _ = _oracle(xyz, tri, K)
    _, g = _build(xyz, tri, K, S.SVG_MODE_BATCHED, chunk=300, vert_cap=64)
    assert g.stats()["redone"] > 0
    _assert_rows(g, rows, tails, K)


@pytest.mark.parametrize("order", [1, 2])
def test_chunked_build_in_source_order(order):
    # chunks take consecutive slices of the source order (here on fewer
    # lanes than sources, so the predicted order is a real permutation) and,
    # with a capacity that forces redos, still give the oracle's rows
    xyz, tri = TORUS()
    K = 80
    print(f"case chunk + order {order}", flush=True)
    rows, tails, _ = _oracle(xyz, tri, K)
    _, g = _build(xyz, tri, K, S.SVG_MODE_BATCHED, chunk=250, lanes=128, source_order=order, vert_cap=64)
    st = g.stats()
    assert st["chunk"] == 250 and st["source_order"] == order and st["redone"] > 0
    _assert_rows(g, rows, tails, K)


def test_large_k_caps_the_vertex_table():
    # K = 2,100 asks for 4K+2 > 8,192 vertex slots, more than the
    # extraction's LDS sort holds: the table is capped at 8,192 and the
    # build still matches the oracle (every vertex of the 642-vertex sphere
    # is reached, so rows hold every vertex whose path passes none).
    xyz, tri = SPHERE()
    K = 2100
    print("case K=2100", flush=True)
    rows, tails, _ = _oracle(xyz, tri, K)
    for mode in (S.SVG_MODE_BATCHED, S.SVG_MODE_PLAIN):
        _, g = _build(xyz, tri, K, mode)
        if mode == S.SVG_MODE_BATCHED:
            assert g.stats()["vert_cap"] == 8192
        _assert_rows(g, rows, tails, K)
        g.FreeSVGStructure()


@pytest.mark.parametrize("qcap", [31, 100])
def test_small_queue_storage_is_recomputed_exactly(qcap):
    # Queue storage below the reference's 1023 items: sources whose queues
    # outgrow it are flagged and recomputed on the plain path, so the rows
    # equal the oracle's, which queues up to 1023.
    xyz, tri = TORUS()
    K = 80
    print(f"case queue_cap {qcap}", flush=True)
    rows, tails, c = _oracle(xyz, tri, K)
    _, g = _build(xyz, tri, K, S.SVG_MODE_BATCHED, queue_cap=qcap)
    st, stt = g.stats(), g.status()
    assert st["queue_cap"] == qcap and st["pseudo_queue_cap"] == qcap
    over = (c["max_winq"] > qcap) | (c["max_pseudoq"] > qcap)
    assert over.any()
    assert np.all(stt[over] & SF_CAPACITY) and np.all(stt[over] & SF_REDONE)
    _assert_rows(g, rows, tails, K)


def test_autotuned_queue_storage_and_shape():
    # The autotuner sizes the queues to the recorded peak (with headroom)
    # and tries the launch shapes one per build; the rows stay the oracle's.
    xyz, tri = SPHERE()
    K = 60
    print("case autotuned queue storage", flush=True)
    rows, tails, c = _oracle(xyz, tri, K)
    m = S.Mesh.from_arrays(xyz, tri)
    seen = set()
    for it in range(7):
        g = S.SVG()
        g.AssignMesh(m)
        g.SetParameters(K, 10, 2)
        g.ConstructSVG(mode=S.SVG_MODE_BATCHED, autotune=1)
        st = g.stats()
        seen.add((st["prop_block"], st["lanes_per_wave"]))
        if it == 0:
            assert st["queue_cap"] == 1023 and st["tuned"] == 0
        else:
            assert st["tuned"] == 1
            assert max(c["max_winq"]) <= st["queue_cap"] < 1023
            assert max(c["max_pseudoq"]) <= st["pseudo_queue_cap"] < 1023
        _assert_rows(g, rows, tails, K)
        g.FreeSVGStructure()
    assert len(seen) == 5


@pytest.mark.parametrize("lpw", [1, 7, 16, 28, 64])
def test_lanes_per_wave_do_not_change_rows(lpw):
    xyz, tri = TORUS()
    K = 80
    print(f"case lanes_per_wave {lpw}", flush=True)
    rows, tails, _ = _oracle(xyz, tri, K)
    _, g = _build(xyz, tri, K, S.SVG_MODE_BATCHED, lanes_per_wave=lpw, prop_block=128)
    assert g.stats()["lanes_per_wave"] == lpw
    _assert_rows(g, rows, tails, K)


@pytest.mark.parametrize("mesh,order,lanes,used", [
    ("torus", 1, 256, 1), ("torus", 2, 256, 2), ("torus", 0, 256, 2), ("torus", 0, 0, 1),
    ("heightfield", 0, 1024, 2), ("convex", 2, 128, 1)])
def test_source_order_does_not_change_rows(mesh, order, lanes, used):
    # 1: vertex id order; 2: predicted from the saddle share around each
    # vertex; 0 with at most four sources per lane picks the prediction, and
    # with a lane per source keeps the id order.  The open heightfield (4,096
    # sources on 1,024 lanes) takes the prediction; an icosphere without
    # noise has no saddle vertex, so its prediction keeps the id order.
    # Every order runs each source once and builds the oracle's rows.
    xyz, tri = {"torus": TORUS, "heightfield": lambda: mg.heightfield64(),
                "convex": lambda: mg.noisy_sphere(splits4=3, centroid_split=False, sigma=0.0)}[mesh]()
    K = 50 if mesh == "heightfield" else 80
    print(f"case source_order {mesh} order {order} lanes {lanes}", flush=True)
    rows, tails, _ = _oracle(xyz, tri, K)
    _, g = _build(xyz, tri, K, S.SVG_MODE_BATCHED, source_order=order, lanes=lanes)
    assert g.stats()["source_order"] == used
    _assert_rows(g, rows, tails, K)


def test_server_memory_admission():
    # The slot count follows the memory budget; a budget below one slot is
    # refused at start (SVG_E_MEM), a budget of three slots gives three.
    xyz, tri = TORUS()
    _, g = _build(xyz, tri, 60, S.SVG_MODE_BATCHED)
    lib = L.lib()
    print("case server memory admission", flush=True)
    srv = C.c_void_p()
    L.check(lib.svg_server_start(g.g, C.byref(L.ServerCfg(64, 0, 64)), C.byref(srv)))
    st = L.ServerStats()
    L.check(lib.svg_server_stats_get(srv, C.byref(st)))
    L.check(lib.svg_server_stop(srv))
    per = st.slot_bytes
    assert per > 0 and st.slots == 64
    rc = lib.svg_server_start(g.g, C.byref(L.ServerCfg(64, per - 1, 64)), C.byref(srv))
    assert rc == -3
    L.check(lib.svg_server_start(g.g, C.byref(L.ServerCfg(64, 3 * per, 64)), C.byref(srv)))
    L.check(lib.svg_server_stats_get(srv, C.byref(st)))
    assert st.slots == 3
    # requests still complete through the three slots, with the answers alone
    V = len(xyz)
    outs, tickets, ref = [], [], []
    for i in range(40):
        d = np.empty(1)
        outs.append(d)
        s_, t_ = i % V, (i * 37 + 5) % V
        ref.append(g.SolveSSSD(s_, t_))
        r = L.Request(0, s_, t_, None, None, 0, d.ctypes.data_as(C.POINTER(C.c_double)), None)
        tickets.append(lib.svg_server_submit(srv, C.byref(r)))
    for t in tickets:
        L.check(lib.svg_server_wait(srv, t, None))
    L.check(lib.svg_server_stats_get(srv, C.byref(st)))
    assert st.max_batch_seen <= 3
    # a consumed ticket cannot be waited for again
    assert lib.svg_server_wait(srv, tickets[0], None) == -1
    L.check(lib.svg_server_stop(srv))
    assert [o[0] for o in outs] == ref


def test_server_completion_callback():
    # svg_server_submit_cb: the callback runs once per request, after its
    # outputs are written; the answers equal the queries alone.
    xyz, tri = TORUS()
    V = len(xyz)
    _, g = _build(xyz, tri, 60, S.SVG_MODE_BATCHED)
    lib = L.lib()
    print("case server callback", flush=True)
    srv = C.c_void_p()
    L.check(lib.svg_server_start(g.g, C.byref(L.ServerCfg(8, 0, 8)), C.byref(srv)))
    seen = {}
    outs = {}

    def done(ticket, status, ms, user):
        seen[ticket] = (status, ms, float(outs[ticket][0][0]))

    cb = L.DONE_FN(done)
    keep = []
    exp = {}
    rng = np.random.default_rng(5)
    for i in range(30):
        if i % 3 == 2:
            src = rng.choice(V, 4, replace=False).astype(np.int32)
            w = np.zeros(4)
            dout = np.empty(V)
            keep += [src, w]
            r = L.Request(1, 0, 0, src.ctypes.data_as(C.POINTER(C.c_int32)), w.ctypes.data_as(C.POINTER(C.c_double)),
                          4, dout.ctypes.data_as(C.POINTER(C.c_double)), None)
            want = g.SolveMSAD(src, w, want_parent=False)[0]
        else:
            s_, t_ = (int(x) for x in rng.choice(V, 2, replace=False))
            dout = np.empty(1)
            r = L.Request(0, s_, t_, None, None, 0, dout.ctypes.data_as(C.POINTER(C.c_double)), None)
            want = np.array([g.SolveSSSD(s_, t_)])
        keep.append(r)
        # the callback can fire before submit returns: register the output first
        nxt = len(exp)
        outs[nxt] = (dout,)
        t = lib.svg_server_submit_cb(srv, C.byref(r), cb, None)
        assert t == nxt
        exp[t] = (dout, want)
    # a callback ticket is never waited for
    assert lib.svg_server_wait(srv, 0, None) == -1
    L.check(lib.svg_server_stop(srv))       # drains the queue; every callback has run
    assert sorted(seen) == sorted(exp)
    for t, (dout, want) in exp.items():
        assert seen[t][0] == 0 and seen[t][1] >= 0
        assert np.array_equal(dout, want)
        assert seen[t][2] == want[0]


def _pinched_spheres():
    # Two spheres that share one vertex: its outgoing half-edges form two
    # fans, so the mesh is not manifold there and the batched build walks
    # the fans (ICHDevice.cu:728-904) instead of reading them by position.
    xyz, tri = SPHERE()
    V = len(xyz)
    p0 = xyz[0]
    n = p0 / np.linalg.norm(p0)
    xyz2 = xyz - 2.0 * np.outer((xyz - p0) @ n, n)     # mirror image through the tangent plane at vertex 0
    remap = np.concatenate([[0], np.arange(1, V) + V - 1])
    tri2 = remap[tri][:, ::-1]                         # the mirror flips the orientation
    return np.vstack([xyz, xyz2[1:]]), np.vstack([tri, tri2]).astype(np.int32)


@pytest.mark.parametrize("mode", [S.SVG_MODE_BATCHED, S.SVG_MODE_PLAIN])
def test_non_manifold_vertex_rows_are_the_oracles(mode):
    xyz, tri = _pinched_spheres()
    K = 40
    print(f"case pinched spheres mode {mode}", flush=True)
    rows, tails, _ = _oracle(xyz, tri, K)
    _, g = _build(xyz, tri, K, mode)
    _assert_rows(g, rows, tails, K)


def test_server_refuses_weights_outside_the_distance_order():
    # the server's MSAD requests follow the rule of svg_msad: a NaN,
    # infinite or negative weight (or -0.0) is refused at submission
    xyz, tri = TORUS()
    _, g = _build(xyz, tri, 60, S.SVG_MODE_BATCHED)
    lib = L.lib()
    print("case server weight refusal", flush=True)
    srv = C.c_void_p()
    L.check(lib.svg_server_start(g.g, C.byref(L.ServerCfg(8, 0, 8)), C.byref(srv)))
    V = len(xyz)
    d = np.empty(V)
    src = np.array([1, 2], np.int32)
    try:
        for bad in (np.nan, np.inf, -1.0, -0.0):
            w = np.array([0.0, bad])
            r = L.Request(1, 0, 0, src.ctypes.data_as(C.POINTER(C.c_int32)), w.ctypes.data_as(C.POINTER(C.c_double)),
                          2, d.ctypes.data_as(C.POINTER(C.c_double)), None)
            assert lib.svg_server_submit(srv, C.byref(r)) == -1
            assert b"non-negative" in lib.svg_last_error()
        w = np.array([0.0, 0.25])
        r = L.Request(1, 0, 0, src.ctypes.data_as(C.POINTER(C.c_int32)), w.ctypes.data_as(C.POINTER(C.c_double)),
                      2, d.ctypes.data_as(C.POINTER(C.c_double)), None)
        t = lib.svg_server_submit(srv, C.byref(r))
        assert t >= 0
        lat = C.c_double()
        L.check(lib.svg_server_wait(srv, t, C.byref(lat)))
        ref, _, _ = g.SolveMSAD(src, w)
        assert np.array_equal(d, ref)
    finally:
        L.check(lib.svg_server_stop(srv))
