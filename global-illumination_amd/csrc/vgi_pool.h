// A small persistent pool of host threads for the edit path's per-region
// loops. run(n, fn) calls fn(begin, end) over chunks of [0, n) on the pool and
// the calling thread and returns when all are done; callers write per-index
// results and merge them in index order, so the outcome does not depend on
// the thread count or the schedule. Calls from several host threads take
// turns (the pool serves one job at a time); fn must not call run.
#pragma once
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace vgi_host {

class WorkPool {
public:
    explicit WorkPool(unsigned threads) {
        for (unsigned i = 0; i < threads; ++i) workers_.emplace_back([this] { loop(); });
    }
    ~WorkPool() {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            stop_ = true;
        }
        wake_.notify_all();
        for (std::thread &w : workers_) w.join();
    }
    WorkPool(const WorkPool &) = delete;
    WorkPool &operator=(const WorkPool &) = delete;

    void run(size_t n, const std::function<void(size_t, size_t)> &fn, size_t chunk = 0) {
        if (n == 0) return;
        if (chunk == 0) chunk = std::max<size_t>(1, n / ((workers_.size() + 1) * 4));
        if (workers_.empty() || n <= chunk) {
            fn(0, n);
            return;
        }
        std::lock_guard<std::mutex> turn(turn_);
        {
            std::lock_guard<std::mutex> lock(mutex_);
            job_ = &fn;
            n_ = n;
            chunk_ = chunk;
            next_.store(0);
            active_ = workers_.size();
            ++generation_;
        }
        wake_.notify_all();
        work();
        std::unique_lock<std::mutex> lock(mutex_);
        done_.wait(lock, [this] { return active_ == 0; });
        job_ = nullptr;
    }

private:
    void work() {
        for (;;) {
            size_t begin = next_.fetch_add(chunk_);
            if (begin >= n_) return;
            (*job_)(begin, std::min(n_, begin + chunk_));
        }
    }
    void loop() {
        size_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lock(mutex_);
                wake_.wait(lock, [&] { return stop_ || generation_ != seen; });
                if (stop_) return;
                seen = generation_;
            }
            work();
            std::lock_guard<std::mutex> lock(mutex_);
            if (--active_ == 0) done_.notify_one();
        }
    }

    std::vector<std::thread> workers_;
    std::mutex turn_;  // one job at a time
    std::mutex mutex_;
    std::condition_variable wake_, done_;
    const std::function<void(size_t, size_t)> *job_ = nullptr;
    size_t n_ = 0, chunk_ = 1, active_ = 0, generation_ = 0;
    std::atomic<size_t> next_{0};
    bool stop_ = false;
};

}  // namespace vgi_host
