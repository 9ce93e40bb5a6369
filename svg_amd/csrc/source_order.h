// Source orders of the batched build (host code, shared with its unit test).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "svg_mesh.h"

namespace svg {

// Sources by recorded cost, costliest first, vertex id breaking ties: two
// stable counting passes over 16-bit digits of the inverted cost (O(V),
// where a comparison sort took milliseconds on the host).
inline void costliest_first(const std::vector<int32_t>& cost, std::vector<int32_t>& order)
{
    const size_t V = cost.size();
    std::vector<int32_t> tmp(V);
    order.resize(V);
    for (size_t i = 0; i < V; ++i) tmp[i] = (int32_t)i;
    std::vector<int32_t> start((1 << 16) + 1);
    for (int shift = 0; shift < 32; shift += 16) {
        std::fill(start.begin(), start.end(), 0);
        auto digit = [&](int32_t v) { return (~(uint32_t)std::max(cost[v], 0) >> shift) & 0xFFFFu; };
        for (size_t i = 0; i < V; ++i) start[digit(tmp[i]) + 1]++;
        for (int d = 0; d < (1 << 16); ++d) start[d + 1] += start[d];
        for (size_t i = 0; i < V; ++i) order[start[digit(tmp[i])]++] = tmp[i];
        tmp.swap(order);
    }
    order.swap(tmp);
}

// Source order for a build with no recorded event counts.  A source's
// event count grows with the saddle vertices among those it reaches: each
// reached saddle becomes a pseudo-source that emits windows over its whole
// fan.  The costliest-first order is therefore estimated from the share of
// saddles around each vertex: the saddle indicator averaged with the
// neighbours ceil(sqrt(K/3)) times (the number of rings that hold K
// vertices on a valence-6 mesh).  Sources run from the highest share down,
// ordered by a stable counting sort over 2^16 levels, so vertex id breaks ties.
// A mesh without saddles keeps the id order (returns false).
inline bool predicted_order(const HostMesh& h, int K, std::vector<int32_t>& order)
{
    const int V = h.nv;
    std::vector<float> x((size_t)V), acc((size_t)V);
    std::vector<float> inv((size_t)V);
    std::vector<int32_t> deg((size_t)V, 0);
    bool any = false;
    for (int e = 0; e < 3 * h.nf; ++e) deg[h.tri[e]]++;
    for (int v = 0; v < V; ++v) {
        x[v] = (h.flags[v] & VF_SADDLE) ? 1.0f : 0.0f;
        any |= x[v] != 0.0f;
        inv[v] = 1.0f / (float)(1 + deg[v]);
    }
    if (!any) return false;
    const int rings = std::min(32, std::max(1, (int)std::ceil(std::sqrt(K / 3.0))));
    for (int it = 0; it < rings; ++it) {
        acc = x;
        for (int f = 0; f < h.nf; ++f) {
            const int32_t* t = &h.tri[3 * (size_t)f];
            acc[t[0]] += x[t[1]];
            acc[t[1]] += x[t[2]];
            acc[t[2]] += x[t[0]];
        }
        for (int v = 0; v < V; ++v) x[v] = acc[v] * inv[v];
    }
    constexpr int kLevels = 1 << 16;
    std::vector<int32_t> start(kLevels + 1, 0);
    std::vector<uint16_t> lvl((size_t)V);
    for (int v = 0; v < V; ++v) {
        const float s = std::min(1.0f, std::max(0.0f, x[v]));
        lvl[v] = (uint16_t)(kLevels - 1 - (int)(s * (float)(kLevels - 1) + 0.5f));   // highest share first
        start[lvl[v] + 1]++;
    }
    for (int i = 0; i < kLevels; ++i) start[i + 1] += start[i];
    order.resize((size_t)V);
    for (int v = 0; v < V; ++v) order[start[lvl[v]]++] = v;
    return true;
}

}  // namespace svg
