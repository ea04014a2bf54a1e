// Host-side helpers the batch calls share.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <functional>
#include <thread>
#include <vector>

namespace dinara {

// Runs body(lo, hi) over [0, count) cut into `threads` contiguous shares, each on a thread of its own; one share
// on the caller's thread when `threads` is one or less.
inline void parallel_for(int64_t count, int64_t threads, const std::function<void(int64_t, int64_t)> &body) {
    const int64_t workers = std::max<int64_t>(1, std::min<int64_t>(threads, count));
    if (workers <= 1) {
        body(0, count);
        return;
    }
    std::vector<std::thread> pool;
    for (int64_t w = 0; w < workers; ++w)
        pool.emplace_back([&, w]() { body(count * w / workers, count * (w + 1) / workers); });
    for (auto &t : pool) t.join();
}

// The decisions of one sweep launch may take at most this much device memory; a larger launch is cut in pieces.
constexpr size_t ARENA_BUDGET = size_t(16) << 30;

}  // namespace dinara
