// Host-side data-parallel loop used by matrix preparation (scaling) and by the device layouts (block sorting).
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <exception>
#include <thread>
#include <vector>

namespace tpdhg {

// True on the worker threads of a running parallel_for (a nested call then runs serially on its caller).
inline bool& in_parallel_for() {
    thread_local bool on = false;
    return on;
}

// Runs f(i) for i in [0, n) on up to 16 host threads, `grain` consecutive indices at a time; f must write disjoint
// data for different i, so the results do not depend on the thread count. An exception thrown by f is rethrown
// here after every thread has stopped.
template <class F>
void parallel_for(int64_t n, int64_t grain, F&& f) {
    if (n <= 0) return;
    const int64_t chunks = (n + grain - 1) / grain;
    const int64_t hw = std::max<int64_t>(1, std::min<int64_t>(16, int64_t(std::thread::hardware_concurrency())));
    const int nt = in_parallel_for() ? 1 : int(std::min(chunks, hw));
    if (nt <= 1) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::atomic<bool> failed{false};
    std::exception_ptr err;
    auto work = [&] {
        const bool was = in_parallel_for();
        in_parallel_for() = true;
        try {
            for (int64_t c = next.fetch_add(1); c < chunks && !failed; c = next.fetch_add(1))
                for (int64_t i = c * grain, e = std::min(n, (c + 1) * grain); i < e; ++i) f(i);
        } catch (...) {
            if (!failed.exchange(true)) err = std::current_exception();
        }
        in_parallel_for() = was;
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (std::thread& t : th) t.join();
    if (err) std::rethrow_exception(err);
}

}  // namespace tpdhg
