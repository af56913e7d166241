// Parallel host memcpy (see fip_copy.hpp).
#include "fip_copy.hpp"

#include <algorithm>
#include <cstring>

namespace fip {

CopyPool::CopyPool(int workers)
{
    for (int i = 0; i < workers; ++i) threads_.emplace_back([this] { worker(); });
}

CopyPool::~CopyPool()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    wake_.notify_all();
    for (std::thread& t : threads_) t.join();
}

void CopyPool::run_chunks()
{
    for (;;) {
        const size_t k = next_.fetch_add(1);
        if (k >= nchunks_) break;
        const size_t o = k * chunk_;
        std::memcpy(dst_ + o, src_ + o, std::min(chunk_, bytes_ - o));
        if (finished_.fetch_add(1) + 1 == nchunks_) {
            std::lock_guard<std::mutex> lk(mu_);
            done_.notify_all();
        }
    }
}

void CopyPool::worker()
{
    unsigned long long seen = 0;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu_);
            wake_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            // Registered under the lock: the caller does not publish the
            // next job before every worker of this one has left it.
            active_.fetch_add(1);
        }
        run_chunks();
        {
            std::lock_guard<std::mutex> lk(mu_);
            active_.fetch_sub(1);
        }
        done_.notify_all();
    }
}

void CopyPool::copy(void* dst, const void* src, size_t bytes)
{
    constexpr size_t kMinChunk = size_t(256) << 10;
    if (threads_.empty() || bytes < 2 * kMinChunk) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t parts = threads_.size() + 1;
    {
        std::unique_lock<std::mutex> lk(mu_);
        // A worker that woke too late for the previous job may still be
        // leaving it: publish only when none is inside a job.
        done_.wait(lk, [&] { return active_.load() == 0; });
        dst_ = static_cast<char*>(dst);
        src_ = static_cast<const char*>(src);
        bytes_ = bytes;
        chunk_ = std::max(kMinChunk, ((bytes + parts - 1) / parts + 63) / 64 * 64);
        nchunks_ = (bytes + chunk_ - 1) / chunk_;
        next_.store(0);
        finished_.store(0);
        ++gen_;
    }
    wake_.notify_all();
    run_chunks();
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return finished_.load() == nchunks_ && active_.load() == 0; });
}

}  // namespace fip
