"""Hand-derived pins of the reference rules that decide which vertices end up
in an SVG row: the queue's drop-on-full capacity, the exact K-stop, the
IsValidWindow and split-filter thresholds, the window-vs-pseudo tie and the
saddle tests at an angle sum of exactly 2*pi.

Each case is built so that the reference's rule fixes the answer by hand
(exact arithmetic on 3-4-5 triangles, ratio gaps of 5e-9 and 2e-8 around
the 1e-8 tolerance, a tie that Initialize creates exactly), so that the
plausible slips (a wrong capacity, >= for >, the two tolerances swapped,
<= for <) change the outcome.  No GPU needed.
"""
import math

import numpy as np
import pytest

import oracle as O
from svg_amd import meshgen as mg

BIG = float(np.finfo(np.float64).max)


# ------------------------------------------------------- queue capacity

@pytest.mark.parametrize("pseudo", [False, True])
def test_queue_drops_the_1024th_push(pseudo):
    # SVG.cu:19-22 sizes both queues at WINPQ_SIZE = PSEUDOWINPQ_SIZE = 1024
    # items with slot 0 unused, and PriorityQueue.cuh:53 returns from push
    # once tail - 1 >= maxSize = 1023.  Keys 1024, 1023, ..., 1 pushed in
    # that order: ids 0..1022 (keys 1024..2) are kept, id 1023 (key 1, the
    # smallest) is dropped, and the pops come out as ids 1022, 1021, ..., 0.
    keys = np.arange(1024, 0, -1, dtype=np.float64)
    order, kept, dropped = O.queue_order(keys, pseudo=pseudo)
    assert kept == 1023 and dropped == 1
    assert order.tolist() == list(range(1022, -1, -1))


@pytest.mark.parametrize("pseudo", [False, True])
def test_queue_keeps_1023_and_drops_every_later_push(pseudo):
    # Increasing keys: the first 1023 pushes fill the queue, the next 7 are
    # all refused; pops return ids 0..1022.
    keys = np.arange(1030, dtype=np.float64)
    order, kept, dropped = O.queue_order(keys, pseudo=pseudo)
    assert kept == 1023 and dropped == 7
    assert order.tolist() == list(range(1023))
    # exactly at capacity nothing is dropped
    order, kept, dropped = O.queue_order(np.arange(1023, dtype=np.float64), pseudo=pseudo)
    assert kept == 1023 and dropped == 0


def test_drop_rule_reaches_the_build():
    # A small bumpy torus at K=400 fills the window queue on many sources:
    # a source drops pushes exactly when its queue reached the 1023 cap.
    xyz, tri = mg.bumpy_torus(nu=90, nv=50, seed=11)
    m = O.Mesh(xyz, tri)
    _, _, cnt, _ = O.svg_build(m, 400, s0=0, s1=len(xyz), nthreads=4)
    c = {n: cnt[:, i] for i, n in enumerate(O.COUNTERS)}
    assert np.all(c["max_winq"] <= 1023)
    assert np.all(c["max_winq"][c["win_drop"] > 0] == 1023)
    assert np.any(c["win_drop"] > 0) and np.any(c["max_winq"] < 1023)


# ------------------------------------------------------------ K-stop

@pytest.mark.parametrize("K", [20, 40, 80])
def test_k_stop_is_exact_when_queues_remain(K):
    # ICHDevice.cu:117-118 breaks after the first main pop that leaves
    # totalCalcVertNum >= K.  A window pop reaches at most one new vertex
    # (ICHDevice.cu:680-682), a pseudo-window pop one per neighbour
    # (ICHDevice.cu:774-776), so a run whose queues were not empty when it
    # stopped had counted fewer than K before its last pop and at least K
    # after it, and exactly K when the last pop added a single vertex.
    # (K large enough that the 2K-slot vertex table never overflows: with
    # an aliased overflow slot "newly reached" is no longer well defined.)
    xyz, tri = mg.noisy_sphere(seed=4, splits4=3, centroid_split=False)
    m = O.Mesh(xyz, tri)
    _, _, cnt, _ = O.svg_build(m, K, nthreads=2)
    c = {n: cnt[:, i] for i, n in enumerate(O.COUNTERS)}
    assert np.all(c["vert_overflow"] == 0) and np.all(c["split_overflow"] == 0)
    left = c["winq_left"] + c["pseudoq_left"]
    assert np.all(left > 0)                     # 642 vertices: no run drains its queues
    assert np.all(c["total_prev"] < K) and np.all(c["total_calc"] >= K)
    one = c["total_calc"] - c["total_prev"] == 1
    assert np.all(c["total_calc"][one] == K) and one.mean() > 0.9


def test_k_stop_not_reached_drains_the_queues():
    # With K above what a source can reach the loop ends on empty queues
    # and counts every reached vertex outside the 1-ring (ICHDevice.cu:72).
    xyz, tri = mg.noisy_sphere(seed=4, splits4=1, centroid_split=False)
    m = O.Mesh(xyz, tri)
    V = len(xyz)
    _, _, cnt, _ = O.svg_build(m, 10 * V, nthreads=2)
    c = {n: cnt[:, i] for i, n in enumerate(O.COUNTERS)}
    assert np.all(c["winq_left"] + c["pseudoq_left"] == 0)
    assert np.all(c["total_calc"] < 10 * V)
    assert np.all(c["reached"] == V)


# ------------------------------------------------- IsValidWindow thresholds
#
# One triangle v0 = (0,0), v1 = (6,0), v2 = (3,4): edge 0 runs v0 -> v1
# (l0 = 6), then v1 -> v2 and v2 -> v0 (both 5), so p3 = (3, 4) exactly.
# A window on edge 0 with b0 = 0, b1 = 6, d0 = sqrt(20), d1 = sqrt(32) has
# its flattened source at xProj = (20 + 36 - 32) / 12 = 2, y = sqrt(20 - 4)
# = 4, i.e. at (2, 4): srcA = ps + |(2,4) - A| = ps + sqrt(20) and srcB =
# ps + |(2,4) - B| = ps + sqrt(32) differ, so the two filters that pair dv1
# with B and dv2 with A cannot be confused.

TRI_XYZ = np.array([[0.0, 0.0, 0.0], [6.0, 0.0, 0.0], [3.0, 4.0, 0.0]])
TRI = np.array([[0, 1, 2]], np.int32)
PS = 100.0
SRC_A = PS + math.sqrt(20.0)
SRC_B = PS + math.sqrt(32.0)


def _valid(vd, is_left, b0=0.0, b1=6.0):
    m = O.Mesh(TRI_XYZ, TRI)
    return O.window_valid(m, 0, b0, b1, math.sqrt(20.0), math.sqrt(32.0), PS, vd, is_left)


def test_window_empty_interval_rejected():
    # ICHDevice.cu:908: b1 <= b0 is rejected before any filter
    assert not _valid({}, True, b0=3.0, b1=3.0)
    assert _valid({}, True)


@pytest.mark.parametrize("is_left", [True, False])
def test_filter_through_v1_has_zero_tolerance(is_left):
    # ICHDevice.cu:924-929: srcB > dv1 + b1 and srcB / (dv1 + b1) - 1 > 0.
    # dv1 + 6 = srcB / (1 + 5e-9): rejected (a 1e-8 tolerance would keep it).
    assert not _valid({0: SRC_B / (1 + 5e-9) - 6.0}, is_left)
    # Pairing dv1 with srcA instead (the smaller one) would keep the window:
    # the same dv1 against A is far from beaten.
    assert SRC_A < SRC_B / (1 + 5e-9)
    # Far from the threshold: kept.
    assert _valid({0: SRC_B - 6.0 + 1e-3}, is_left)


@pytest.mark.parametrize("is_left", [True, False])
def test_filter_through_v2_has_zero_tolerance(is_left):
    # ICHDevice.cu:930-935: srcA against dv2 + l0 - b0.
    assert not _valid({1: SRC_A / (1 + 5e-9) - 6.0}, is_left)
    assert _valid({1: SRC_A - 6.0 + 1e-3}, is_left)
    # dv2 just under srcB - 6 does not beat srcA, which is smaller: kept
    # (a window that paired dv2 with srcB would be rejected here).
    assert _valid({1: SRC_B / (1 + 5e-9) - 6.0}, is_left)


def test_third_filter_tolerance_differs_by_side():
    # ICHDevice.cu:936-952: the left child tests srcA against dv3 + |p3 - A|
    # with tolerance 0, the right child srcB against dv3 + |p3 - B| with
    # RELATIVE_ERROR = 1e-8 (ICHDevice.cuh:10).  |p3 - A| = |p3 - B| = 5.
    # A ratio gap of 5e-9 rejects the left child only:
    assert not _valid({2: SRC_A / (1 + 5e-9) - 5.0}, True)
    assert _valid({2: SRC_B / (1 + 5e-9) - 5.0}, False)
    # a gap of 2e-8 rejects both:
    assert not _valid({2: SRC_A / (1 + 2e-8) - 5.0}, True)
    assert not _valid({2: SRC_B / (1 + 2e-8) - 5.0}, False)
    # and the left child is judged by A, the right one by B: dv3 tuned to
    # srcB's threshold (2e-8) is far from beating srcA, so the left child stays.
    assert _valid({2: SRC_B / (1 + 2e-8) - 5.0}, True)


def test_filters_are_strict_at_exact_ties():
    # Exact arithmetic: d0 = d1 = 5 on [0, 6] puts the source at (3, 4)
    # (xProj = 3, y = sqrt(25 - 9) = 4), so srcA = srcB = 100 + 5 = 105 and
    # dv + 5 = 105 or dv + 6 = 105 are exact ties: '>' keeps the window.
    m = O.Mesh(TRI_XYZ, TRI)
    for vd in ({0: 99.0}, {1: 99.0}, {2: 100.0}):
        for is_left in (True, False):
            assert O.window_valid(m, 0, 0.0, 6.0, 5.0, 5.0, 100.0, vd, is_left), (vd, is_left)
    # one ulp closer beats it (tolerance 0) for the first filter
    assert not O.window_valid(m, 0, 0.0, 6.0, 5.0, 5.0, 100.0, {0: np.nextafter(99.0, 0.0)}, True)


# ------------------------------------------------------- split filter
#
# Two faces: f0 = (v1, vA, v2) with the source vA = (3,-4) and f1 = (v1, v2,
# vOp) with vOp = (3, 4); v1 = (0,0), v2 = (6,0).  A window on f0's edge
# v2 -> v1 (edge 2) with b0 = 0, b1 = 6, d0 = d1 = 5 has its source at
# (3, 4) in its frame.  PropagateWindow unfolds f1 over e0 = v1 -> v2 with
# the far vertex at (3, -4): interX = 3 - (-4)(3 - 3)/(-4 - 4) = 3 lies in
# (0, 6), directDist = |(0, -8)| = 8 and cand = 8 + ps, all exact.

QUAD_XYZ = np.array([[0.0, 0.0, 0.0], [6.0, 0.0, 0.0], [3.0, -4.0, 0.0], [3.0, 4.0, 0.0]])
QUAD = np.array([[0, 2, 1], [0, 1, 3]], np.int32)


def _prop(split, ps=0.0):
    m = O.Mesh(QUAD_XYZ, QUAD)
    return O.propagate_one(m, 2, 2, 0.0, 6.0, 5.0, 5.0, ps, {2: 0.0}, split)


def test_split_filter_relative_tolerance():
    # ICHDevice.cu:662-667: cand > split.dist and cand / split.dist - 1 >
    # RELATIVE_ERROR keeps only the child on split.x's side.
    r = _prop((8.0 / (1 + 2e-8), 2.0))               # gap 2e-8: filtered
    assert r["counters"]["child_built"] == 1
    assert r["counters"]["split_write"] == 0 and r["counters"]["vert_write"] == 0
    assert r["far_dist"] == BIG                        # the far vertex is not touched
    r = _prop((8.0 / (1 + 5e-9), 2.0))               # gap 5e-9: not filtered
    assert r["counters"]["child_built"] == 2
    assert r["counters"]["split_write"] == 0           # cand is not below split.dist
    assert r["counters"]["vert_write"] == 1 and r["far_dist"] == 8.0


def test_split_filter_side_is_split_x_below_interx():
    # hasL = split.x < interX (ICHDevice.cu:664): x = 2 keeps the left child
    # (on f1's edge v2 -> vOp, edge 4), x = 4 the right one (vOp -> v1, edge
    # 5), and x = interX = 3 exactly is not below it: right child.
    for x, side in ((2.0, 4), (4.0, 5), (3.0, 5)):
        r = _prop((8.0 / (1 + 2e-8), x))
        assert r["counters"]["child_built"] == 1
        assert [c[0] for c in r["children"]] in ([side], []), (x, r["children"])
        assert r["children"] and r["children"][0][0] == side, (x, r["children"])


def test_split_record_written_only_when_improved():
    # No record: cand = 8 is written with x = l0 - interX = 3 (ICHDevice.cu:
    # 671-676), and the far vertex gets distance 8 (ICHDevice.cu:678-687).
    r = _prop(None)
    assert r["counters"]["split_write"] == 1 and r["split"] == (8.0, 3.0)
    assert r["far_dist"] == 8.0 and r["counters"]["child_built"] == 2
    # A record exactly at cand is neither filtered (8 > 8 is false) nor
    # improved (8 < 8 is false).
    r = _prop((8.0, 1.0))
    assert r["counters"]["split_write"] == 0 and r["split"] == (8.0, 1.0)
    assert r["counters"]["child_built"] == 2


# ------------------------------------------------ window-vs-pseudo tie

def _saddle_field(n=12, seed=3):
    rng = np.random.default_rng(seed)
    xs = np.linspace(-1, 1, n)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    X = X + rng.uniform(-0.02, 0.02, X.shape)
    Y = Y + rng.uniform(-0.02, 0.02, Y.shape)
    Z = 0.6 * (X ** 2 - Y ** 2)
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    tri = []
    for i in range(n - 1):
        for j in range(n - 1):
            a, b, c, d = i * n + j, (i + 1) * n + j, (i + 1) * n + j + 1, i * n + j + 1
            tri += [[a, b, c], [a, c, d]]
    return xyz, np.array(tri, np.int32)


def test_window_pseudo_tie_goes_to_the_pseudo_window():
    # ICHDevice.cu:88 takes the window only if its key is strictly smaller,
    # :110 the pseudo-window when window.minDist >= pseudo.dist.  Initialize
    # makes exact ties: the window on a fan face's far edge whose projection
    # falls before b0 has minDist = d0 = len(cur), the same double as the
    # pseudo-window pushed for that saddle neighbour (ICHDevice.cu:534-570).
    xyz, tri = _saddle_field()
    m = O.Mesh(xyz, tri)
    ties = 0
    for s in range(len(xyz)):
        kind, wk, pk, _ = O.ich_trace(m, s, K=30)
        both = (wk != BIG) & (pk != BIG)
        assert np.all((kind[both] == 0) == (wk[both] < pk[both]))
        t = both & (wk == pk)
        assert np.all(kind[t] == 1)
        ties += int(t.sum())
    assert ties > 0                 # the rule was exercised on exact ties


# ------------------------------------------- angle sum of exactly 2*pi
#
# c = (0,0) with a = (3,0), b = (0,4), d = (-3,0), e = (0,-4): every corner at
# c is acos((9 + 16 - 25) / 24) = acos(0), and four of them add up to the
# double 2*pi exactly.  f = (3,4) closes face (a, f, b).  Every vertex but c
# is on the boundary (never a saddle).

CROSS_XYZ = np.array([[0, 0, 0], [3, 0, 0], [0, 4, 0], [-3, 0, 0], [0, -4, 0], [3, 4, 0]], np.float64)
CROSS = np.array([[0, 1, 2], [0, 2, 3], [0, 3, 4], [0, 4, 1], [1, 5, 2]], np.int32)


def test_angle_exactly_two_pi_counts_as_saddle_in_initialize_only():
    m = O.Mesh(CROSS_XYZ, CROSS)
    ex = m.export()
    assert ex["angle"][0] == 2.0 * math.pi and not ex["boundary"][0]
    # Source a: Initialize pushes a pseudo-window for every 1-ring vertex
    # whose angle is not < 2*pi (ICHDevice.cu:558); c is the only one, and no
    # later step pushes another (c is not > 2*pi at ICHDevice.cu:688, the
    # other vertices are boundary vertices).
    _, _, c = O.ich_single(m, 1, K=-1)
    assert c["pseudo_push"] == 1
    assert c["pseudo_pop"] == 1
    # Source f reaches c only through PropagateWindow, whose test is
    # angle > 2*pi (ICHDevice.cu:688): no pseudo-window at all.
    dist, _, c = O.ich_single(m, 5, K=-1)
    assert c["pseudo_push"] == 0
    assert abs(dist[0] - 5.0) < 1e-12


def test_vertex_without_faces_starts_no_window():
    # decision I: a vertex that no face uses (firstEdge -1) has an empty row
    # and leaves every other row as it is without it
    xyz, tri = mg.noisy_sphere(seed=4, splits4=3, centroid_split=False)
    K = 60
    r0, t0, _, _ = O.svg_build(O.Mesh(xyz, tri), K)
    xyz2 = np.vstack([xyz, [[0.0, 0.0, 0.0]]])
    r1, t1, _, _ = O.svg_build(O.Mesh(xyz2, tri), K)
    assert t1[-1] == 0
    assert np.array_equal(t0, t1[:-1]) and np.array_equal(r0, r1[:-1])
