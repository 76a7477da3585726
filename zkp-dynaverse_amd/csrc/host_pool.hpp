// A process-wide pool of host worker threads for the host epilogues (window
// combination of batched MSMs).  parallel_for(count, fn) runs fn(i) for every
// i < count on the pool and the calling thread and returns when all are done;
// calls from several threads at once are independent jobs.
#pragma once
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace dvgpu {

class HostPool {
  public:
    static HostPool& get() {
        static HostPool pool;
        return pool;
    }

    template <class Fn>
    void parallel_for(size_t count, Fn&& fn) {
        if (count == 0) return;
        if (count == 1 || workers_.empty()) {
            for (size_t i = 0; i < count; ++i) fn(i);
            return;
        }
        auto job = std::make_shared<Job>();
        job->count = count;
        job->fn = [&fn](size_t i) { fn(i); };
        {
            std::lock_guard<std::mutex> lk(mu_);
            jobs_.push_back(job);
        }
        cv_.notify_all();
        run(*job);  // the caller works too
        std::unique_lock<std::mutex> lk(job->mu);
        job->done_cv.wait(lk, [&] { return job->done.load() == job->count; });
    }

    // Asynchronous form: start(count, fn) queues fn(i) for every i < count
    // (claimed in increasing i) and returns; the caller may run claims
    // itself with help(), and must call finish() (which waits for every
    // claimed index) before fn or anything it references goes away.
    struct Job;
    template <class Fn>
    std::shared_ptr<Job> start(size_t count, Fn& fn) {
        auto job = std::make_shared<Job>();
        job->count = count;
        job->fn = [&fn](size_t i) { fn(i); };
        if (count && !workers_.empty()) {
            {
                std::lock_guard<std::mutex> lk(mu_);
                jobs_.push_back(job);
            }
            cv_.notify_all();
        }
        return job;
    }
    // Runs one unclaimed index of the job; false when none is left.
    static bool help(Job& job) {
        const size_t i = job.next.fetch_add(1);
        if (i >= job.count) return false;
        job.fn(i);
        if (job.done.fetch_add(1) + 1 == job.count) {
            std::lock_guard<std::mutex> lk(job.mu);
            job.done_cv.notify_all();
        }
        return true;
    }
    static void finish(Job& job) {
        while (help(job)) {
        }
        std::unique_lock<std::mutex> lk(job.mu);
        job.done_cv.wait(lk, [&] { return job.done.load() >= job.count; });
    }

    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }

    struct Job {
        size_t count = 0;
        std::function<void(size_t)> fn;
        std::atomic<size_t> next{0}, done{0};
        std::mutex mu;
        std::condition_variable done_cv;
    };

  private:

    HostPool() {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const unsigned n = std::min(hw, 16u) - 1;  // the caller is one more worker
        for (unsigned t = 0; t < n; ++t) workers_.emplace_back([this] { loop(); });
    }

    static void run(Job& job) {
        size_t i;
        while ((i = job.next.fetch_add(1)) < job.count) {
            job.fn(i);
            if (job.done.fetch_add(1) + 1 == job.count) {
                std::lock_guard<std::mutex> lk(job.mu);
                job.done_cv.notify_all();
            }
        }
    }

    void loop() {
        for (;;) {
            std::shared_ptr<Job> job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
                if (stop_) return;
                job = jobs_.front();
                if (job->next.load() >= job->count) {  // fully claimed: retire it
                    jobs_.pop_front();
                    continue;
                }
            }
            run(*job);
        }
    }

    std::vector<std::thread> workers_;
    std::deque<std::shared_ptr<Job>> jobs_;
    std::mutex mu_;
    std::condition_variable cv_;
    bool stop_ = false;
};

}  // namespace dvgpu
