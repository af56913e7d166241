// Parallel host memcpy for the staging copies of the host entry points.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <mutex>
#include <thread>
#include <vector>

namespace fip {

// Host-to-host copies of large arrays (into pinned staging) split over a
// few worker threads and the caller: one core's memcpy reaches only part of
// the host's memory bandwidth, and these copies sit on the host drop-in's and
// the serving form's critical path.  copy() blocks until every byte is
// copied; calls must not overlap (the context's mutex serialises them).
class CopyPool {
public:
    explicit CopyPool(int workers);
    ~CopyPool();
    void copy(void* dst, const void* src, size_t bytes);

private:
    void worker();
    void run_chunks();
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable wake_, done_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0, chunk_ = 0, nchunks_ = 0;
    std::atomic<size_t> next_{0}, finished_{0};
    std::atomic<int> active_{0};
    unsigned long long gen_ = 0;
    bool stop_ = false;
};

}  // namespace fip
