// Short-lived host threads for the readers and generators, behind the C ABI:
// nothing they run may throw past it.
#pragma once

#include <atomic>
#include <thread>
#include <vector>

// Runs f(k) for k = 0 .. count-1: k = 0 on the calling thread, the others on
// threads of their own where one can be started and on the calling thread
// where not. An exception thrown by f stays inside its call. Returns false
// when any call threw, true otherwise.
template <typename F>
bool run_on_threads(unsigned count, F&& f)
{
    std::atomic<bool> ok{true};
    auto guarded = [&](unsigned k) {
        try {
            f(k);
        } catch (...) {
            ok.store(false);
        }
    };
    std::vector<std::thread> th;
    unsigned k = 1;
    try {
        th.reserve(count > 0 ? count - 1 : 0);
        for (; k < count; ++k) th.emplace_back(guarded, k);
    } catch (...) {
        // No more threads: the rest run here.
    }
    for (unsigned j = k; j < count; ++j) guarded(j);
    if (count > 0) guarded(0);
    for (auto& t : th) t.join();
    return ok.load();
}
