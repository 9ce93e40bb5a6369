"""Host source orders of the batched build (svg_amd/csrc/source_order.h),
compiled with the host C++ compiler and run on the CPU: the costliest-first
order equals a comparison sort by (cost descending, id ascending), and the
order predicted from the mesh is a permutation that keeps id order without
saddle vertices and starts where the saddles are."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from svg_amd import meshgen as mg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "svg_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include "source_order.h"
int main(int argc, char** argv)
{
    FILE* f = std::fopen(argv[1], "rb");
    int hd[3];
    if (std::fread(hd, 4, 3, f) != 3) return 2;
    const int mode = hd[0], n = hd[1], m = hd[2];
    std::vector<int32_t> order;
    int ret = 1;
    if (mode == 0) {                       // costliest_first over n costs
        std::vector<int32_t> cost((size_t)n);
        if (std::fread(cost.data(), 4, (size_t)n, f) != (size_t)n) return 2;
        svg::costliest_first(cost, order);
    } else {                               // predicted_order: n vertices, m faces, K
        int K;
        if (std::fread(&K, 4, 1, f) != 1) return 2;
        svg::HostMesh h;
        h.nv = n; h.nf = m;
        h.tri.resize(3 * (size_t)m);
        h.flags.resize((size_t)n);
        if (std::fread(h.tri.data(), 4, h.tri.size(), f) != h.tri.size()) return 2;
        if (std::fread(h.flags.data(), 1, h.flags.size(), f) != h.flags.size()) return 2;
        ret = svg::predicted_order(h, K, order) ? 1 : 0;
    }
    std::fclose(f);
    FILE* g = std::fopen(argv[2], "wb");
    std::fwrite(&ret, 4, 1, g);
    std::fwrite(order.data(), 4, order.size(), g);
    std::fclose(g);
    return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    cxx = shutil.which("g++") or shutil.which("c++")
    if not cxx:
        pytest.skip("no host C++ compiler")
    d = tmp_path_factory.mktemp("order")
    src = d / "driver.cpp"
    src.write_text(DRIVER)
    exe = d / "driver"
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)], check=True)

    def run(header, *arrays):
        inp, out = d / "in.bin", d / "out.bin"
        with open(inp, "wb") as f:
            np.asarray(header, np.int32).tofile(f)
            for a in arrays:
                a.tofile(f)
        subprocess.run([str(exe), str(inp), str(out)], check=True)
        r = np.fromfile(out, np.int32)
        return int(r[0]), r[1:]
    return run


@pytest.mark.parametrize("n,hi", [(1, 5), (1000, 20), (72000, 12000), (5003, 2_000_000_000), (64, 1)])
def test_costliest_first_is_a_stable_sort(driver, n, hi):
    rng = np.random.default_rng(n)
    cost = rng.integers(0, hi, n, dtype=np.int64).astype(np.int32)
    _, order = driver([0, n, 0], cost)
    ref = sorted(range(n), key=lambda i: (-int(cost[i]), i))
    assert order.tolist() == ref


def _saddle_flags(xyz, tri):
    """VF_SADDLE (2) where the corner angles add up to more than 2*pi."""
    V = len(xyz)
    asum = np.zeros(V)
    for k in range(3):
        a, b, c = tri[:, k], tri[:, (k + 1) % 3], tri[:, (k + 2) % 3]
        u, v = xyz[b] - xyz[a], xyz[c] - xyz[a]
        cosang = (u * v).sum(1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1)
        np.add.at(asum, a, np.arccos(np.clip(cosang, -1.0, 1.0)))
    return np.where(asum > 2 * np.pi, 2, 1).astype(np.uint8)


def test_predicted_order_without_saddles_keeps_id_order(driver):
    xyz, tri = mg.noisy_sphere(splits4=3, centroid_split=False, sigma=0.0)
    flags = _saddle_flags(xyz, tri)
    assert not (flags & 2).any()
    ret, order = driver([1, len(xyz), len(tri), 80], np.ascontiguousarray(tri, np.int32), flags)
    assert ret == 0 and order.size == 0


def test_predicted_order_starts_at_an_isolated_saddle(driver):
    # one saddle flag in the middle of a flat grid: its share is the largest,
    # its ring of neighbours comes before every vertex two or more edges away
    xyz, tri = mg.planar_grid(n=21, jitter=0.0)
    V = len(xyz)
    flags = np.ones(V, np.uint8)
    c = int(np.argmin(((xyz[:, :2] - xyz[:, :2].mean(0)) ** 2).sum(1)))
    flags[c] = 2
    ret, order = driver([1, V, len(tri), 30], np.ascontiguousarray(tri, np.int32), flags)
    assert ret == 1
    assert sorted(order.tolist()) == list(range(V))
    ring = {int(v) for t in tri if c in t for v in t} - {c}
    assert order[0] == c
    assert set(order[1:1 + len(ring)].tolist()) == ring


def test_predicted_order_follows_source_cost_on_a_bumpy_torus(driver):
    # the prediction ranks sources like the oracle's time-weighted event
    # counts (window pops + 3 x pseudo pops) on a small bumpy torus
    import oracle as O
    xyz, tri = mg.bumpy_torus(nu=60, nv=36)
    K = 80
    flags = _saddle_flags(xyz, tri)
    ret, order = driver([1, len(xyz), len(tri), K], np.ascontiguousarray(tri, np.int32), flags)
    assert ret == 1 and sorted(order.tolist()) == list(range(len(xyz)))
    _, _, cnt, _ = O.svg_build(O.Mesh(xyz, tri), K)
    ci = {n: i for i, n in enumerate(O.COUNTERS)}
    cost = cnt[:, ci["win_pop"]] + 3 * cnt[:, ci["pseudo_pop"]]
    rank_pred = np.empty(len(xyz)); rank_pred[order] = np.arange(len(xyz))
    rank_cost = np.empty(len(xyz)); rank_cost[np.argsort(-cost, kind="stable")] = np.arange(len(xyz))
    assert np.corrcoef(rank_pred, rank_cost)[0, 1] > 0.6
