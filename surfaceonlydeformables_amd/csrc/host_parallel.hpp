// Static split of an index range over the host's threads (at most 16), for the O(n^2) host loops of the
// facade's post-processing.  f(begin, end) is called once per chunk; chunks are disjoint.
#pragma once
#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

namespace cqbem {

template <class F>
void parallel_for(int64_t n, F f) {
    const unsigned t = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (n < 4096 || t == 1) {
        f((int64_t)0, n);
        return;
    }
    const int64_t chunk = (n + t - 1) / t;
    std::vector<std::thread> pool;
    for (unsigned i = 0; i < t; i++) {
        const int64_t a = (int64_t)i * chunk, b = std::min(n, a + chunk);
        if (a < b) pool.emplace_back(f, a, b);
    }
    for (auto &th : pool) th.join();
}

}  // namespace cqbem
