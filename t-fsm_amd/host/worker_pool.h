// Persistent worker threads for the host side of the miner: pattern growth
// of a level, either waited for at once (parallel_for) or run on the workers
// while the caller evaluates the level on the GPU (Background).
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace tfsm {

// Worker threads kept for the life of the process: a mining run grows
// several levels, and starting threads for each of them cost more than the
// growth of a small level.  Each job has its own state; the caller takes
// items too and waits only for the items, so a worker that wakes late finds
// its job drained and does nothing.  A new job wakes one worker, and each
// worker that takes the job wakes the next before working on it, so the
// caller pays for one wake-up rather than one per worker (it starts GPU work
// right after posting a background job).
class WorkerPool {
public:
    static WorkerPool &get() {
        static WorkerPool pool;
        return pool;
    }
    ~WorkerPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : workers_) t.join();
    }
    void run(size_t n, int threads, const std::function<void(size_t)> &fn) {
        std::lock_guard<std::mutex> serial(run_mu_);
        auto job = std::make_shared<Job>();
        job->fn = &fn;
        job->n = n;
        job->left.store(n);
        {
            std::lock_guard<std::mutex> lk(mu_);
            while ((int)workers_.size() < threads - 1) workers_.emplace_back([this]() { loop(); });
            cur_ = job;
            ++gen_;
        }
        cv_.notify_one();
        drain(*job);
        std::unique_lock<std::mutex> lk(job->m);
        job->cv.wait(lk, [&]() { return job->left.load() == 0; });
        std::lock_guard<std::mutex> lk2(mu_);
        cur_.reset();
    }

    struct Job {
        const std::function<void(size_t)> *fn = nullptr;
        std::function<void(size_t)> own;     // a background job's function
        size_t n = 0;
        std::atomic<size_t> next{0}, left{0};
        std::mutex m;
        std::condition_variable cv;
    };
    // A job on the workers alone: start() returns at once, so the caller can
    // do other work meanwhile, and finish() waits for its items.  The pool
    // runs no other job in between (the same thread must call both).
    std::shared_ptr<Job> start(size_t n, int threads, std::function<void(size_t)> fn) {
        run_mu_.lock();
        auto job = std::make_shared<Job>();
        job->own = std::move(fn);
        job->fn = &job->own;
        job->n = n;
        job->left.store(n);
        if (n == 0) return job;
        {
            std::lock_guard<std::mutex> lk(mu_);
            while ((int)workers_.size() < std::max(1, threads - 1)) workers_.emplace_back([this]() { loop(); });
            cur_ = job;
            ++gen_;
        }
        cv_.notify_one();
        return job;
    }
    void finish(const std::shared_ptr<Job> &job) {
        {
            std::unique_lock<std::mutex> lk(job->m);
            job->cv.wait(lk, [&]() { return job->left.load() == 0; });
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            cur_.reset();
        }
        run_mu_.unlock();
    }

private:
    static void drain(Job &job) {
        size_t i;
        while ((i = job.next.fetch_add(1)) < job.n) {
            (*job.fn)(i);
            if (job.left.fetch_sub(1) == 1) {
                std::lock_guard<std::mutex> lk(job.m);
                job.cv.notify_all();
            }
        }
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            std::shared_ptr<Job> job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&]() { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                job = cur_;
            }
            cv_.notify_one();           // the next waiting worker
            if (job) drain(*job);
        }
    }
    std::mutex run_mu_, mu_;
    std::condition_variable cv_;
    std::vector<std::thread> workers_;
    std::shared_ptr<Job> cur_;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

template <class F>
void parallel_for(size_t n, int threads, F &&f) {
    if (threads <= 1 || n < 2) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    const std::function<void(size_t)> fn = [&](size_t i) { f(i); };
    WorkerPool::get().run(n, std::min<int>(threads, (int)n), fn);
}

// fn over [0, n) on the pool's workers while the caller goes on; wait() (or
// the destructor) returns once every item ran.
class Background {
public:
    Background(size_t n, int threads, std::function<void(size_t)> fn)
        : job_(WorkerPool::get().start(n, std::min<int>(std::max(threads, 2), (int)n + 1), std::move(fn))) {}
    ~Background() { wait(); }
    void wait() {
        if (job_) WorkerPool::get().finish(job_);
        job_.reset();
    }

private:
    std::shared_ptr<WorkerPool::Job> job_;
};

}  // namespace tfsm
