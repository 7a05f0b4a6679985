// tos_parallel.h -- a small host thread pool for independent loop
// iterations.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <thread>
#include <vector>

namespace tos {

// f(i) for i in [0, n) on up to 16 threads, in chunks of `grain`; the
// iterations must be independent.
template <class F>
void parallel_for(int64_t n, int64_t grain, F f)
{
    const int64_t chunks = (n + grain - 1) / grain;
    const int nt = (int)std::max<int64_t>(
        1, std::min<int64_t>(std::min(16u, std::max(1u, std::thread::hardware_concurrency())), chunks));
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        for (;;) {
            const int64_t c = next.fetch_add(1);
            if (c >= chunks) break;
            const int64_t e = std::min(n, (c + 1) * grain);
            for (int64_t i = c * grain; i < e; ++i) f(i);
        }
    };
    std::vector<std::thread> th;
    for (int k = 1; k < nt; ++k) th.emplace_back(work);
    work();
    for (auto &x : th) x.join();
}

}  // namespace tos
