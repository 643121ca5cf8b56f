// host_pool.h -- fixed sets of host worker threads: HostPool runs one
// parallel loop at a time for its caller; JobQueue runs background jobs.
//
// run(n, f) calls f(0..n-1) spread over the workers and the calling thread
// and returns when every call is done; one run at a time (the owner calls
// it from a single thread).  Each worker runs `init` once when it starts
// (the server makes its device current there).  No device code and no HIP
// types: the same header is checked on the CPU under ThreadSanitizer
// (tests/native/pool_check.cpp).
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace pca {

class HostPool {
public:
    HostPool(unsigned workers, std::function<void()> init = {}) : init_(std::move(init))
    {
        for (unsigned i = 0; i < workers; i++) threads_.emplace_back([this] { loop(); });
    }
    HostPool(const HostPool &) = delete;
    HostPool &operator=(const HostPool &) = delete;
    ~HostPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        wake_.notify_all();
        for (std::thread &t : threads_) t.join();
    }
    size_t workers() const { return threads_.size(); }

    // f(0..n-1); only on the calling thread when `parallel` is false
    template <class F>
    void run(size_t n, F &&f, bool parallel = true)
    {
        if (!parallel || n <= 1 || threads_.empty()) {
            for (size_t i = 0; i < n; i++) f(i);
            return;
        }
        const std::function<void(size_t)> job(std::ref(f));
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &job;
            n_ = n;
            next_.store(0);
            pending_ = threads_.size();
            gen_++;
        }
        wake_.notify_all();
        for (size_t i; (i = next_.fetch_add(1)) < n;) job(i);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return pending_ == 0; });   // every worker has left the job
        job_ = nullptr;
    }

private:
    void loop()
    {
        if (init_) init_();
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(size_t)> *job;
            size_t n;
            {
                std::unique_lock<std::mutex> lk(mu_);
                wake_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                job = job_;
                n = n_;
            }
            for (size_t i; (i = next_.fetch_add(1)) < n;) (*job)(i);
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_.notify_one();
        }
    }
    std::function<void()> init_;
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable wake_, done_;
    const std::function<void(size_t)> *job_ = nullptr;
    size_t n_ = 0, pending_ = 0;
    uint64_t gen_ = 0;
    std::atomic<size_t> next_{0};
    bool stop_ = false;
};

// Background jobs: submit() queues a job and returns at once; a free
// worker runs it.  wait_all() returns when every job submitted so far has
// finished.  Jobs do not submit jobs.  Each worker runs `init` once when it
// starts.
class JobQueue {
public:
    JobQueue(unsigned workers, std::function<void()> init = {}) : init_(std::move(init))
    {
        for (unsigned i = 0; i < workers; i++) threads_.emplace_back([this] { loop(); });
    }
    JobQueue(const JobQueue &) = delete;
    JobQueue &operator=(const JobQueue &) = delete;
    ~JobQueue()
    {
        wait_all();
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        wake_.notify_all();
        for (std::thread &t : threads_) t.join();
    }
    size_t workers() const { return threads_.size(); }

    void submit(std::function<void()> job)
    {
        if (threads_.empty()) {   // no workers: run it here
            job();
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            jobs_.push_back(std::move(job));
            in_flight_++;
        }
        wake_.notify_one();
    }

    void wait_all()
    {
        std::unique_lock<std::mutex> lk(mu_);
        idle_.wait(lk, [this] { return in_flight_ == 0; });
    }

private:
    void loop()
    {
        if (init_) init_();
        for (;;) {
            std::function<void()> job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                wake_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
                if (jobs_.empty()) return;   // stop_ and nothing left
                job = std::move(jobs_.front());
                jobs_.erase(jobs_.begin());
            }
            job();
            std::lock_guard<std::mutex> lk(mu_);
            if (--in_flight_ == 0) idle_.notify_all();
        }
    }
    std::function<void()> init_;
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable wake_, idle_;
    std::vector<std::function<void()>> jobs_;
    size_t in_flight_ = 0;
    bool stop_ = false;
};

}  // namespace pca
