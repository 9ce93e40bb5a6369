"""Benchmark: SVG build throughput (batched vs plain) and served query latency.

One step = one batched SVG build of the workload (default: bumpy torus,
V=72,000, F=144,000, K=500) with the mesh already resident in HBM.  After the
timed steps the script also measures, once:
  * the plain one-source-per-thread build (best of a small grid sweep, timed
    in full),
  * a serving replay on the resident graph: Poisson arrivals at --qps,
    70% A* SSSD / 30% MSAD with log-uniform source-set sizes in [1, 64],
    latency from scheduled arrival to host completion,
  * the oracle's CPU build on a bounded source sample (cpu_baseline).
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic bytes per build event (unpadded payloads of the reference
# records: Window 68 B, PseudoWindow 28 B, SplitInfo 20 B, VertInfo 22 B,
# SVGNode 36 B; ICHDevice.cuh:16-72, SVG.cuh:21).
BYTES_PROP = {"win_push": 68, "win_pop": 68, "pseudo_push": 28, "pseudo_pop": 28,
              "propagate": 68, "child_built": 60, "split_write": 20, "vert_write": 22,
              "fan_halfedge": 32, "pseudo_neigh": 16}
BYTES_EXTRACT = {"reached": 22, "bt_hops": 32}
BYTES_ARC = 36
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(workload, K):
    """HBM bytes per launch of the propagation kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload (both counters
    scaled by a 1 GiB copy measured in the same process); None if absent."""
    import re
    names = {"bumpy_torus": f"torus72k_K{K}"}
    if workload not in names:
        return None
    p = os.path.join(ROOT, "profiles", f"batched_prop_kernel_{names[workload]}.md")
    if not os.path.exists(p):
        return None
    m = re.search(r"traffic_bytes_per_launch: (\d+)", open(p).read())
    return int(m.group(1)) if m else None


def prop_bytes(c):
    return sum(c[k] * b for k, b in BYTES_PROP.items())


def extract_bytes(c, arcs):
    return sum(c[k] * b for k, b in BYTES_EXTRACT.items()) + BYTES_ARC * arcs


def make_requests(V, n, qps, seed=21, msad_frac=0.3):
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, n)
    arrival = np.cumsum(gaps) - gaps[0]
    kinds = (rng.uniform(0, 1, n) < msad_frac).astype(np.int32)
    reqs = []
    for i in range(n):
        if kinds[i] == 0:
            s, t = rng.choice(V, 2, replace=False)
            reqs.append((0, int(s), int(t), None))
        else:
            k = int(math.floor(math.exp(rng.uniform(0.0, math.log(65.0)))))
            k = max(1, min(64, k))
            src = rng.choice(V, k, replace=False).astype(np.int32)
            reqs.append((1, 0, 0, src))
    return arrival, reqs


def serve_replay(g, V, n, qps, inflight, delta_scale=1.0, rounds_per_sync=4, msad_delta_scale=0.0, streams=0,
                 chain_rows=0, queue_cap=0, given=None):
    """Replays n requests (or `given` = (arrival, reqs)) through the server;
    queue_cap > 0 submits them all at once (throughput)."""
    import ctypes as C
    from svg_amd import _lib as L
    arrival, reqs = given if given is not None else make_requests(V, n, qps)
    n = len(reqs)
    if queue_cap:
        arrival = np.zeros(n)
        queue_cap = max(queue_cap, n)
    lib = L.lib()
    cfg = L.ServerCfg(inflight, 0, inflight, delta_scale, rounds_per_sync, msad_delta_scale, streams, chain_rows,
                      queue_cap)
    srv = C.c_void_p()
    L.check(lib.svg_server_start(g.g, C.byref(cfg), C.byref(srv)))
    arr = (L.Request * n)()
    keep = []
    n_msad = 0
    sssd_out = np.zeros(n, np.float64)
    msad_out = None
    for i, (kind, s, t, src) in enumerate(reqs):
        r = arr[i]
        r.kind = kind
        if kind == 0:
            r.s, r.t = s, t
            r.dist_out = sssd_out[i:].ctypes.data_as(C.POINTER(C.c_double))
        else:
            n_msad += 1
            w = np.zeros(len(src), np.float64)
            dout = np.empty(V, np.float64)
            sout = np.empty(V, np.int32)      # nearest-source ids, as SolveMSAD returns
            keep += [src, w, dout, sout]
            r.src = src.ctypes.data_as(C.POINTER(C.c_int32))
            r.w = w.ctypes.data_as(C.POINTER(C.c_double))
            r.ns = len(src)
            r.dist_out = dout.ctypes.data_as(C.POINTER(C.c_double))
            r.srcid_out = sout.ctypes.data_as(C.POINTER(C.c_int32))
    lat = np.zeros(n, np.float64)
    t0 = time.perf_counter()
    L.check(lib.svg_server_replay(srv, arr, np.ascontiguousarray(arrival), n, lat))
    wall = time.perf_counter() - t0
    st = L.ServerStats()
    L.check(lib.svg_server_stats_get(srv, C.byref(st)))
    L.check(lib.svg_server_stop(srv))
    ok = lat >= 0
    out = {
        "requests": n, "completed": int(ok.sum()), "refused": int((~ok).sum()),
        "offered_qps": None if queue_cap else qps, "achieved_qps": float(ok.sum() / wall),
        "arrivals": "all at once (batch)" if queue_cap else "Poisson",
        "p50_ms": float(np.percentile(lat[ok], 50)), "p99_ms": float(np.percentile(lat[ok], 99)),
        "mean_ms": float(lat[ok].mean()), "max_ms": float(lat[ok].max()),
        "msad_fraction": n_msad / n, "inflight_slots": inflight, "launches": int(st.launches),
        "max_concurrent": int(st.max_batch_seen), "delta_scale": delta_scale,
        "rounds_per_sync": rounds_per_sync, "msad_delta_scale": msad_delta_scale or delta_scale,
        "streams": streams or 4, "chain_rows": chain_rows or 4,
        "msad_returns": "distances and nearest-source ids",
        "latency_from": "scheduled arrival to host completion",
    }
    kinds = np.array([r[0] for r in reqs])
    for k, name in ((0, "sssd"), (1, "msad")):
        nk = max(1, st.n_kind[k])
        sel = ok & (kinds == k)
        out[name] = {"n": int(st.n_kind[k]), "rounds_per_query": st.rounds_kind[k] / nk,
                     "expanded_per_query": st.settled_kind[k] / nk,
                     "arcs_per_query": st.relaxed_kind[k] / nk,
                     "p50_ms": float(np.percentile(lat[sel], 50)) if sel.any() else None,
                     "p99_ms": float(np.percentile(lat[sel], 99)) if sel.any() else None}
    return out, (arrival, reqs, sssd_out)


def point_workload(xyz, tri, n, seed=5):
    """n surface-point pairs: half with f1 a face index within 40 of f0 (nearby
    on the generated grids), half uniform; points Dirichlet(2,2,2) in their faces."""
    rng = np.random.default_rng(seed)
    F = len(tri)
    f0 = rng.integers(F, size=n)
    f1 = rng.integers(F, size=n)
    near = rng.random(n) < 0.5
    f1[near] = (f0[near] + rng.integers(1, 40, size=int(near.sum()))) % F
    w0 = rng.dirichlet([2.0, 2.0, 2.0], size=n)
    w1 = rng.dirichlet([2.0, 2.0, 2.0], size=n)
    pts = np.stack([np.einsum("ij,ijk->ik", w0, xyz[tri[f0]]), np.einsum("ij,ijk->ik", w1, xyz[tri[f1]])], 1)
    return np.stack([f0, f1], 1).astype(np.int32), pts


def point_queries(g, xyz, tri, n, n_single=16):
    """Surface-point SSSD (SVG.cu:253-352): one batch of n queries, then
    n_single queries one at a time."""
    faces, pts = point_workload(xyz, tri, n)
    g.SolveSSSDPointsBatch(faces[:64], pts[:64])          # buffers and solver slots
    t0 = time.perf_counter()
    b = g.SolveSSSDPointsBatch(faces, pts)
    dt = time.perf_counter() - t0
    lat = {0: [], 1: []}
    for i in range(n_single):
        t1 = time.perf_counter()
        r = g.SolveSSSDPoints(int(faces[i, 0]), pts[i, 0], int(faces[i, 1]), pts[i, 1])
        lat[max(r.route, 0)].append((time.perf_counter() - t1) * 1e3)
    routes = np.bincount(b["route"] + 1, minlength=3)
    return {"n": n, "queries_per_s": n / dt, "seconds": dt,
            "routes": {"direct": int(routes[1]), "graph": int(routes[2]), "none": int(routes[0])},
            "single_ms_p50": {"direct": float(np.median(lat[0])) if lat[0] else None,
                              "graph": float(np.median(lat[1])) if lat[1] else None},
            "workload": "half near (face index within 40), half uniform; Dirichlet(2,2,2) points"}, faces, pts


def copy_bandwidth(mb=1024, reps=5):
    """Device-to-device copy of mb MiB timed with HIP events (bytes read +
    written per second): the on-box bandwidth beside the 8 TB/s spec."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mb << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    hip.hipMemset(a, 1, C.c_size_t(n))
    e0, e1 = C.c_void_p(), C.c_void_p()
    hip.hipEventCreate(C.byref(e0))
    hip.hipEventCreate(C.byref(e1))
    best = None
    for _ in range(reps):
        hip.hipEventRecord(e0, None)
        assert hip.hipMemcpyAsync(b, a, C.c_size_t(n), 3, None) == 0
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        best = ms.value if best is None else min(best, ms.value)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return 2.0 * n / (best * 1e-3) / 1e9


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# Algorithmic bytes of one query (model of the reference's solver,
# SVG.cu:369-413): per settled vertex its distance and row pointer (16 B),
# per scanned arc the target id, weight and target distance (20 B), per
# improvement distance + parent + source id (16 B); A* also reads the
# target-side position for the heuristic per improvement (24 B).
def query_bytes(kind, st):
    b = 16 * st[0] + 20 * st[1] + 16 * st[2]
    return b + (24 * st[2] if kind == 0 else 0)


def cpu_query_baselines(og, reqs, cores, single_s=6.0, all_s=10.0):
    """The oracle's A* / Dijkstra (the reference's host solvers) on the
    serving requests: one thread per query over a prefix of each kind, then
    all cores over a prefix of the mixed list without pacing.  Also returns
    the prefix used for all cores with its algorithmic bytes."""
    import concurrent.futures as cf
    out = {}
    for kind, name in ((0, "sssd"), (1, "msad")):
        t0 = time.perf_counter()
        n = 0
        for k, s, t, src in reqs:
            if k != kind:
                continue
            if kind == 0:
                og.sssd(s, t)
            else:
                og.msad(src, np.zeros(len(src)))
            n += 1
            if time.perf_counter() - t0 > single_s and n >= 4:
                break
        dt = time.perf_counter() - t0
        out[name] = {"value": n / dt, "unit": "queries/s", "ms_per_query": dt / n * 1e3, "cores": 1,
                     "kind": "port", "sample": f"oracle on the first {n} {name.upper()} requests ({dt:.1f} s, one thread)"}

    def one(r):
        k, s, t, src = r
        if k == 0:
            d, _, _, _, st = og.sssd(s, t)
        else:
            _, _, _, st = og.msad(src, np.zeros(len(src)))
        return query_bytes(k, st)

    # all cores: grow the prefix until the run takes about all_s seconds
    n = min(len(reqs), max(cores * 4, 64))
    while True:
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(cores) as ex:
            qb = list(ex.map(one, reqs[:n]))
        dt = time.perf_counter() - t0
        if dt >= all_s or n >= len(reqs):
            break
        n = min(len(reqs), int(n * max(1.5, all_s / max(dt, 1e-3))))
    out["all_cores"] = {"value": n / dt, "unit": "queries/s", "cores": cores, "kind": "port",
                        "sample": f"oracle on the first {n} requests of the mix without pacing ({dt:.1f} s, "
                                  f"{cores} threads)"}
    return out, n, float(np.sum(qb))


def calib_copy(mb):
    """Device-to-device copy of a known size, to scale hardware byte counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mb << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0      # device to device
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"calibration copy: {mb} MiB read + {mb} MiB written")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="bumpy_torus")
    ap.add_argument("--K", type=int, default=0)
    ap.add_argument("--plain", default="sweep", help="sweep | ref | none")
    ap.add_argument("--requests", type=int, default=100000)
    ap.add_argument("--qps", type=float, default=1000.0)
    ap.add_argument("--inflight", type=int, default=64)
    ap.add_argument("--delta-scale", default="1.0",
                    help="bucket width / mean arc; a comma list runs one replay per value")
    ap.add_argument("--msad-delta-scale", default="1",
                    help="MSAD bucket width / mean arc (0 = --delta-scale); a comma list runs one replay per value")
    ap.add_argument("--rounds-per-sync", default="4",
                    help="frontier rounds between host checks; a comma list runs one replay per value")
    ap.add_argument("--streams", default="4",
                    help="slot groups on concurrent HIP streams; a comma list runs one replay per value")
    ap.add_argument("--chain-rows", default="4",
                    help="same-bucket rows a wave expands itself per round (-1 = none); a comma list runs one "
                         "replay per value")
    ap.add_argument("--batch", action="store_true",
                    help="submit every request at once (queue sized to hold them): throughput, not latency")
    ap.add_argument("--calib-mb", type=int, default=0,
                    help="first copy this many MiB device-to-device (byte-counter calibration)")
    ap.add_argument("--points", type=int, default=8192,
                    help="surface-point queries in the batch (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU only; several GPUs would each "
                                   "build a cost-balanced share of the sources (not implemented)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import svg_amd as S
    from svg_amd import meshgen as mg

    w = mg.WORKLOADS[args.workload]
    K = args.K or w["K"]
    xyz, tri = mg.workload(args.workload)
    V = len(xyz)
    digest = mg.mesh_digest(xyz, tri)
    log(f"device: {S.device_name()}; workload {args.workload} V={V} F={len(tri)} K={K}")
    mesh = S.Mesh.from_arrays(xyz, tri)
    mesh.copyToGPU()
    if args.calib_mb > 0:
        calib_copy(args.calib_mb)

    def build(mode, **kn):
        g = S.SVG()
        g.AssignMesh(mesh)
        g.SetParameters(K, 10, 2)
        g.ConstructSVG(mode=mode, **kn)
        return g

    # the autotuner learns from the warm-up build(s) of this mesh: source
    # order, table sizes, block shape (rows do not depend on it)
    first_build = None
    for i in range(args.warmup):
        g = build(S.SVG_MODE_BATCHED, autotune=1)
        st0 = g.stats()
        if i == 0:
            first_build = {"seconds": st0["seconds"], "tuned": st0["tuned"], "prop_block": st0["prop_block"],
                           "source_order": st0["source_order"]}
        log(f"warmup {i}: build {st0['seconds']:.4f} s (tuned {st0['tuned']})")
        g.FreeSVGStructure()
    times, stats, walls = [], [], []
    wall0 = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        g = build(S.SVG_MODE_BATCHED, autotune=1)
        st = g.stats()
        if i < args.steps - 1:
            g.FreeSVGStructure()
        walls.append(time.perf_counter() - t0)
        times.append(st["seconds"])
        stats.append(st)
        log(f"step {i}: build {st['seconds']:.4f} s (prop {st['prop_seconds']:.4f}, "
            f"extract {st['extract_seconds']:.4f}, redo {st['redo_seconds']:.4f}), arcs {st['arcs']}, "
            f"wall {walls[-1]:.4f} s")
    wall = time.perf_counter() - wall0
    st = stats[-1]
    c = st["counters"]
    build_s = float(np.median(times))
    pb = prop_bytes(c)
    eb = extract_bytes(c, st["arcs"])
    prop_gbs = pb / st["prop_seconds"] / 1e9
    res = {
        "metric": "SVG build sec + queries/sec with p50/p99 latency, batched vs plain, 1 GPU",
        "value": V / build_s, "unit": "sources/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": build_s * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "V": V, "F": len(tri), "K": K,
                   "mesh_sha256": digest, "mode": "batched", "parallelism": "single GPU"},
        "build_s": build_s, "build_s_all": times, "wall_s_steps": wall,
        "wall_ms_per_step": wall / args.steps * 1e3, "wall_s_all": walls,
        "wall_ms_per_step_median": float(np.median(walls)) * 1e3,
        "wall_includes": "per-build hipMalloc / hipFree of the tables, their resets, the stats copies",
        "first_build_untuned": first_build,
        "batched": {k: st[k] for k in ("prop_seconds", "extract_seconds", "redo_seconds", "redone",
                                       "flagged_drop", "flagged_overflow", "flagged_error", "lanes",
                                       "prop_block", "lanes_per_wave", "queue_cap", "pseudo_queue_cap",
                                       "chunk", "vert_cap", "split_cap", "tuned", "source_order", "arcs")},
        "windows_per_s": c["win_pop"] / st["prop_seconds"],
        "counters": c,
        "roofline": {"bound": "hbm", "kernel": "batched_prop_kernel",
                     "achieved": prop_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": prop_gbs / HBM_PEAK_GBS, "traffic": measured_traffic(args.workload, K),
                     "traffic_source": "profiles/ rocprofv3 FETCH_SIZE + WRITE_SIZE summary, copy-calibrated",
                     "algorithmic_bytes": pb, "kernel_seconds": st["prop_seconds"],
                     "extract_achieved_gbs": eb / st["extract_seconds"] / 1e9},
    }
    g_keep = g
    try:
        cbw = copy_bandwidth()
        res["roofline"]["peak_measured"] = cbw
        res["roofline"]["frac_of_measured"] = prop_gbs / cbw
        res["roofline"]["peak_measured_source"] = "1 GiB device-to-device hipMemcpy, read + written bytes / event time"
        log(f"copy bandwidth {cbw:.0f} GB/s")
    except Exception as e:       # the bench line still prints without it
        log(f"copy bandwidth not measured: {e}")

    # ---- plain form
    if args.plain != "none":
        grids = [(64, 128)] if args.plain == "ref" else [(64, 128), (256, 128), (1125, 64), (576, 128)]
        best = None
        per_grid = {}
        for (b, t) in grids:
            gp = build(S.SVG_MODE_PLAIN, plain_blocks=b, plain_threads=t)
            s = gp.stats()["seconds"]
            per_grid[f"{b}x{t}"] = s
            log(f"plain grid {b}x{t}: {s:.3f} s")
            if best is None or s < best[0]:
                best = (s, b, t)
            gp.FreeSVGStructure()
        res["plain"] = {"best_seconds": best[0], "grid": [best[1], best[2]],
                        "grids_tried": [list(x) for x in grids], "seconds_per_grid": per_grid}
        res["speedup_batched_vs_plain"] = best[0] / build_s
        # the reference launches its plain kernel on 64 blocks of 128 threads (SVG.cuh:10-11, SVG.cu:163)
        res["speedup_batched_vs_plain_reference_grid"] = per_grid["64x128"] / build_s

    # ---- serving
    if args.requests > 0:
        runs = []
        for ds in [float(x) for x in args.delta_scale.split(",")]:
            for mds in [float(x) for x in args.msad_delta_scale.split(",")]:
                for rps in [int(x) for x in args.rounds_per_sync.split(",")]:
                    for nst in [int(x) for x in args.streams.split(",")]:
                        for ch in [int(x) for x in args.chain_rows.split(",")]:
                            sv, rq = serve_replay(g_keep, V, args.requests, 1e9 if args.batch else args.qps,
                                                  args.inflight, ds, rps, mds, nst, ch,
                                                  args.requests if args.batch else 0)
                            serve_reqs = rq
                            runs.append(sv)
                            log(f"serving (delta x{ds}, msad delta x{mds or ds}, {rps} rounds/sync, {nst} streams, "
                                f"chain {ch}): p50 {sv['p50_ms']:.3f} ms p99 {sv['p99_ms']:.3f} ms "
                                f"qps {sv['achieved_qps']:.1f}; sssd {sv['sssd']}; msad {sv['msad']}")
        res["serving"] = min(runs, key=lambda r: r["p99_ms"])
        if len(runs) > 1:
            res["serving_sweep"] = [{k: r[k] for k in ("delta_scale", "msad_delta_scale", "rounds_per_sync", "streams",
                                                       "chain_rows", "p50_ms", "p99_ms", "achieved_qps")}
                                    for r in runs]

    # ---- surface-point queries
    pq_faces = None
    if args.points > 0:
        pq, pq_faces, pq_pts = point_queries(g_keep, xyz, tri, args.points)
        res["point_queries"] = pq
        log(f"point queries: {pq['queries_per_s']:.0f}/s over {pq['n']} ({pq['routes']}); "
            f"single p50 {pq['single_ms_p50']} ms")

    # ---- CPU baseline: the oracle on a bounded sample of the same build
    if not args.no_cpu:
        import oracle as O
        om = O.Mesh(xyz, tri)
        cores = os.cpu_count() or 1
        try:
            cores = len(os.sched_getaffinity(0))
        except Exception:
            pass
        n_probe = max(cores, 64)
        _, _, _, sp = O.svg_build(om, K, s0=0, s1=n_probe, nthreads=cores, counters=False)
        per = sp / n_probe
        n = int(min(V, max(n_probe, args.cpu_seconds / max(per, 1e-6))))
        _, _, _, secs = O.svg_build(om, K, s0=0, s1=n, nthreads=cores, counters=False)
        res["cpu_baseline"] = {"value": n / secs, "unit": "sources/s", "cores": cores,
                               "kind": "port", "cpu": cpu_model(),
                               "sample": f"oracle build of sources [0,{n}) of {V} "
                                         f"({secs:.1f} s on {cores} threads)"}
        # one core, sources spread over the mesh (about a quarter of the time)
        n1 = max(8, int(args.cpu_seconds / 4 / max(per * cores, 1e-6)))
        t1 = 0.0
        for s0 in np.linspace(0, V - 1, n1).astype(int):
            _, _, _, d1 = O.svg_build(om, K, s0=int(s0), s1=int(s0) + 1, nthreads=1, counters=False)
            t1 += d1
        res["cpu_baseline_1core"] = {"value": n1 / t1, "unit": "sources/s", "cores": 1, "kind": "port",
                                     "sample": f"oracle build of {n1} sources spread over the mesh "
                                               f"({t1:.1f} s, one thread)"}
        log(f"cpu build: {n / secs:.0f} sources/s on {cores} threads, {n1 / t1:.0f} on one")
        og = None
        if pq_faces is not None or args.requests > 0:
            Vg, Kg, _ = g_keep.info()
            tails = np.empty(Vg, np.int32)
            adj = np.empty(Vg * Kg, np.int32)
            w = np.empty(Vg * Kg, np.float64)
            S.check(S.lib().svg_graph_to_host(g_keep.g, tails.ctypes.data, adj.ctypes.data, w.ctypes.data,
                                              None, None, None, None))
            og = O.Graph.from_arrays(xyz, adj, w, tails, Kg)
            del adj, w
        if args.requests > 0:
            # the reference's host solvers on the serving requests, then the
            # same prefix on the GPU all at once: query throughput and the
            # query path's algorithmic bytes per second
            arrival, reqs, _ = serve_reqs
            qb, nq, qbytes = cpu_query_baselines(og, reqs, cores)
            res["query_cpu_baseline"] = qb
            sv_b, _ = serve_replay(g_keep, V, nq, 1e9, args.inflight, 1.0, 4, 1.0, 4, 4, nq,
                                   given=(arrival[:nq], reqs[:nq]))
            gpu_s = nq / sv_b["achieved_qps"]
            res["query_roofline"] = {
                "bound": "hbm", "achieved": qbytes / gpu_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": qbytes / gpu_s / 1e9 / HBM_PEAK_GBS, "queries": nq, "seconds": gpu_s,
                "algorithmic_bytes": qbytes, "gpu_queries_per_s": sv_b["achieved_qps"],
                "vs_cpu_all_cores": sv_b["achieved_qps"] / qb["all_cores"]["value"],
                "bytes_model": "16 B per settled vertex + 20 B per scanned arc + 16 B per improvement "
                               "(+24 B per improvement for A*), counts of the oracle's heap solvers",
                "time": "wall time of the same requests submitted at once to the server"}
            log(f"query cpu baseline {qb}; gpu batch of {nq}: {sv_b['achieved_qps']:.0f} queries/s, "
                f"{qbytes / gpu_s / 1e9:.1f} GB/s algorithmic")
        if pq_faces is not None:
            # the oracle's point queries on the same graph, one core, ~3 s
            t0 = time.perf_counter()
            m_done = 0
            while m_done < len(pq_faces) and (time.perf_counter() - t0 < 3.0 or m_done < 4):
                f0, f1 = (int(x) for x in pq_faces[m_done])
                O.sssd_point(om, og, None, Kg, f0, pq_pts[m_done, 0], f1, pq_pts[m_done, 1])
                m_done += 1
            secs_p = time.perf_counter() - t0
            res["point_queries"]["cpu_baseline"] = {
                "value": m_done / secs_p, "unit": "queries/s", "cores": 1, "kind": "port",
                "sample": f"oracle on the first {m_done} queries of the batch ({secs_p:.1f} s, one thread)"}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
