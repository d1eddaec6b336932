// copy_pool.h -- a small persistent thread pool for host-side row copies
// of the host-memory drop-in (large images are split over it).  run() is
// blocking and may be called from many threads at once; the pool threads
// live until the process ends.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace goofy_impl {

class CopyPool {
public:
    static CopyPool& get()
    {
        static CopyPool* p = new CopyPool();  // never destroyed
        return *p;
    }
    int workers() const { return (int)threads_.size(); }
    // Runs fn(i) for i in [0, n) on the pool and the calling thread; returns
    // when every item is done and no worker still references the job.
    template <class F>
    void run(int n, F&& fn)
    {
        if (n <= 1 || threads_.empty()) {
            for (int i = 0; i < n; ++i) fn(i);
            return;
        }
        Job job;
        job.n = n;
        job.fn = [&fn](int i) { fn(i); };
        {
            std::lock_guard<std::mutex> lock(mu_);
            jobs_.push_back(&job);
        }
        cv_.notify_all();
        work_on(job);
        std::unique_lock<std::mutex> lock(mu_);
        auto it = std::find(jobs_.begin(), jobs_.end(), &job);
        if (it != jobs_.end()) jobs_.erase(it);  // no worker can pick it up any more
        cv_done_.wait(lock, [&] { return job.done.load() == job.n && job.active == 0; });
    }

private:
    struct Job {
        int n = 0;
        std::atomic<int> next{0};
        std::atomic<int> done{0};
        int active = 0;  // workers inside work_on (guarded by the pool mutex)
        std::function<void(int)> fn;
    };
    CopyPool()
    {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const unsigned n = std::min(7u, hw > 1 ? hw - 1 : 0u);
        for (unsigned i = 0; i < n; ++i) threads_.emplace_back([this] { loop(); });
        for (auto& t : threads_) t.detach();
    }
    static void work_on(Job& job)
    {
        int i, finished = 0;
        while ((i = job.next.fetch_add(1)) < job.n) {
            job.fn(i);
            ++finished;
        }
        if (finished) job.done.fetch_add(finished);
    }
    void loop()
    {
        for (;;) {
            Job* job = nullptr;
            {
                std::unique_lock<std::mutex> lock(mu_);
                cv_.wait(lock, [&] { return !jobs_.empty(); });
                job = jobs_.front();
                if (job->next.load() >= job->n) {
                    jobs_.pop_front();
                    continue;
                }
                ++job->active;
            }
            work_on(*job);
            {
                std::lock_guard<std::mutex> lock(mu_);
                --job->active;
            }
            cv_done_.notify_all();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_, cv_done_;
    std::deque<Job*> jobs_;
    std::vector<std::thread> threads_;
};

}  // namespace goofy_impl
