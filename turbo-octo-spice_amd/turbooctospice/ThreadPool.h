// ThreadPool.h -- runs a list of tasks on N threads and returns the AND of
// their results (same contract as the reference's ThreadPool.h:9-25, on
// std::thread instead of boost::asio).
#ifndef TOS_AMD_TURBOOCTOSPICE_THREAD_POOL
#define TOS_AMD_TURBOOCTOSPICE_THREAD_POOL

#include <atomic>
#include <functional>
#include <list>
#include <thread>
#include <vector>

namespace turbooctospice {

class ThreadPool {
public:
    typedef std::function<bool()> Task;
    explicit ThreadPool(int nthreads) : n_(nthreads < 1 ? 1 : nthreads) {}
    /// Runs every task, n threads at a time; returns the AND of the results.
    bool run(std::list<Task> tasks)
    {
        std::vector<Task> v(tasks.begin(), tasks.end());
        std::vector<char> ok(v.size(), 0);
        std::atomic<size_t> next{0};
        auto worker = [&]() {
            for (;;) {
                const size_t k = next.fetch_add(1);
                if (k >= v.size()) break;
                ok[k] = v[k]() ? 1 : 0;
            }
        };
        std::vector<std::thread> th;
        const int nt = (int)std::min<size_t>((size_t)n_, std::max<size_t>(1, v.size()));
        for (int i = 1; i < nt; ++i) th.emplace_back(worker);
        worker();
        for (auto &t : th) t.join();
        bool all = true;
        for (char c : ok) all = all && c;
        return all;
    }
private:
    int n_;
};

}  // namespace turbooctospice

#endif
