// A host-side parallel loop shared by the mesh builder and the mesh upload.
#pragma once

#include <algorithm>
#include <system_error>
#include <thread>
#include <vector>

namespace svg {

// Runs body(begin, end) over [0, n) in chunks on up to 16 host threads (on
// the calling thread alone for small n or when no thread can be started);
// every element is computed the same way whatever the split.
template <class F>
inline void parallel_for(int n, F&& body)
{
    const unsigned hw = std::thread::hardware_concurrency();
    const int T = (int)std::max(1u, std::min(16u, hw));
    if (n < (1 << 16) || T == 1) { body(0, n); return; }
    const int chunk = (n + T - 1) / T;
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) {
        const int a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        try {
            th.emplace_back([&body, a, b] { body(a, b); });
        } catch (const std::system_error&) {
            body(a, b);
        }
    }
    for (std::thread& x : th) x.join();
}

}  // namespace svg
