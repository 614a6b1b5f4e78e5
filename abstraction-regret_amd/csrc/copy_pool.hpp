// Persistent thread pool for large host memory copies (the serving scheduler's staging).
// Host-only C++; no HIP call.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace gfsm {

// Persistent helper threads for the scheduler's staging copies: a copy is cut into parts of
// kPart bytes that the caller and the helpers take in order, so no thread is created per
// copy (a launch stages its bytes in several chunks). One caller at a time.
class CopyPool {
  public:
    explicit CopyPool(size_t helpers) {
        for (size_t k = 0; k < helpers; ++k) threads_.emplace_back([this] { work(); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            quit_ = true;
        }
        go_.notify_all();
        for (auto &t : threads_) t.join();
    }
    void copy(uint8_t *dst, const uint8_t *src, size_t n) {
        if (n < 2 * kPart || threads_.empty()) {
            if (n) std::memcpy(dst, src, n);
            return;
        }
        {
            std::unique_lock<std::mutex> lk(mu_);
            // a helper that woke late for the previous copy may still hold its pointers and
            // would take this copy's part numbers: set the job up only once none is active
            done_.wait(lk, [&] { return active_ == 0; });
            dst_ = dst;
            src_ = src;
            n_ = n;
            parts_ = (n + kPart - 1) / kPart;
            next_.store(0, std::memory_order_relaxed);
            finished_ = 0;
            ++generation_;
        }
        go_.notify_all();
        const size_t mine = take(dst, src, n, parts_);
        std::unique_lock<std::mutex> lk(mu_);
        finished_ += mine;
        // every part copied and no helper still holding this job's pointers
        done_.wait(lk, [&] { return finished_ == parts_ && active_ == 0; });
    }

  private:
    static constexpr size_t kPart = size_t(1) << 20;
    // Copies parts until none is left; returns how many it copied.
    size_t take(uint8_t *dst, const uint8_t *src, size_t n, size_t parts) {
        size_t did = 0;
        for (size_t p; (p = next_.fetch_add(1, std::memory_order_relaxed)) < parts; ++did) {
            const size_t lo = p * kPart, len = std::min(kPart, n - lo);
            std::memcpy(dst + lo, src + lo, len);
        }
        return did;
    }
    void work() {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            go_.wait(lk, [&] { return quit_ || generation_ != seen; });
            if (quit_) return;
            seen = generation_;
            uint8_t *dst = dst_;
            const uint8_t *src = src_;
            const size_t n = n_, parts = parts_;
            ++active_;
            lk.unlock();
            const size_t did = take(dst, src, n, parts);
            lk.lock();
            finished_ += did;
            --active_;
            done_.notify_all();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable go_, done_;
    uint8_t *dst_ = nullptr;
    const uint8_t *src_ = nullptr;
    size_t n_ = 0, parts_ = 0, finished_ = 0, active_ = 0;
    std::atomic<size_t> next_{0};
    uint64_t generation_ = 0;
    bool quit_ = false;
};

}  // namespace gfsm
