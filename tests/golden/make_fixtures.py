"""Writes tests/golden/oracle_small.npz: frozen outputs of the oracle (the CPU
restatement of the reference path) on two small meshes.

The reference ships no tests, meshes or expected outputs, and it cannot run
here (it needs the CUDA headers), so these are not reference outputs: they
are snapshots of the oracle after its closed-form pins (tests/test_oracle.py,
tests/test_oracle_point.py) passed, kept so that any later change of the
oracle's results, or of the GPU path compared against them, is seen.

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402
from svg_amd import meshgen as mg  # noqa: E402

K_SPHERE, K_DISK = 24, 30


def point_cases(xyz, tri, seed, n):
    rng = np.random.default_rng(seed)
    f = rng.integers(0, len(tri), size=(n, 2)).astype(np.int32)
    w = rng.dirichlet([2.0, 2.0, 2.0], size=(n, 2))
    p = np.einsum("nsj,nsjk->nsk", w, xyz[tri[f]])
    return f, p


def compute(xyz, tri, K, faces, pts):
    m = O.Mesh(xyz, tri)
    rows, tails, _, _ = O.svg_build(m, K, counters=False)
    out = dict(tails=tails, adj=rows["adj"], dist=rows["geodDist"], src_edge=rows["nextToSrcEdge"],
               dst_edge=rows["nextToDstEdge"], src_x=rows["nextToSrcX"], dst_x=rows["nextToDstX"])
    pd, pi = [], []
    for (f0, f1), (p0, p1) in zip(faces, pts):
        r = O.sssd_point(m, rows, tails, K, int(f0), p0, int(f1), p1)
        pd.append([r["dist"], r["src_x"], r["dst_x"]])
        pi.append([r["src_edge"], r["dst_edge"], r["last_vert"], r["route"]])
    out["point_d"] = np.asarray(pd)
    out["point_i"] = np.asarray(pi, dtype=np.int32)
    return out


def fixtures():
    data = {}
    cases = {"sphere": (mg.noisy_sphere(seed=2, splits4=2, centroid_split=False), K_SPHERE, 3),
             "disk": (mg.planar_disk(n_interior=60, n_boundary=20, seed=5), K_DISK, 4)}
    for name, ((xyz, tri), K, seed) in cases.items():
        faces, pts = point_cases(xyz, tri, seed, 12)
        res = compute(xyz, tri, K, faces, pts)
        data[f"{name}_xyz"] = xyz
        data[f"{name}_tri"] = tri
        data[f"{name}_K"] = np.int32(K)
        data[f"{name}_faces"] = faces
        data[f"{name}_pts"] = pts
        for k, v in res.items():
            data[f"{name}_{k}"] = v
    return data


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "oracle_small.npz"), **fixtures())
    print("wrote", os.path.join(HERE, "oracle_small.npz"))
