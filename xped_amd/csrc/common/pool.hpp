// Persistent host worker pool shared by the plan builders of the tensor layer and the device
// layer (one pool per process).
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace xp {

// A persistent pool of host worker threads for the plan builders: starting threads per plan
// costs more than the work of a typical plan.  run(n, f) calls f(i) for i in [0, n) on the
// workers and the calling thread, and returns when all calls are done; the first exception is
// rethrown.
class WorkerPool {
public:
    static WorkerPool& get()
    {
        static WorkerPool pool;
        return pool;
    }
    unsigned size() const { return static_cast<unsigned>(threads_.size()) + 1; }
    template <class F>
    void run(std::size_t n, F&& f)
    {
        if(n == 0) return;
        if(n == 1 || threads_.empty()) {
            for(std::size_t i = 0; i < n; ++i) f(i);
            return;
        }
        std::lock_guard<std::mutex> one(run_mutex_);  // one parallel region at a time
        std::function<void(std::size_t)> job = [&f](std::size_t i) { f(i); };
        {
            std::lock_guard<std::mutex> lk(m_);
            job_ = &job;
            n_ = n;
            next_ = 0;
            pending_ = n;
            err_ = nullptr;
            ++generation_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(m_);
        done_cv_.wait(lk, [this] { return pending_ == 0; });
        job_ = nullptr;
        if(err_) std::rethrow_exception(err_);
    }
    ~WorkerPool()
    {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for(auto& t : threads_) t.join();
    }

private:
    WorkerPool()
    {
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        for(unsigned i = 1; i < hw; ++i) threads_.emplace_back([this] { loop(); });
    }
    void work()
    {
        for(;;) {
            std::size_t i;
            const std::function<void(std::size_t)>* job;
            {
                std::lock_guard<std::mutex> lk(m_);
                if(!job_ || next_ >= n_) return;
                i = next_++;
                job = job_;
            }
            try {
                (*job)(i);
            } catch(...) {
                std::lock_guard<std::mutex> lk(m_);
                if(!err_) err_ = std::current_exception();
            }
            std::lock_guard<std::mutex> lk(m_);
            if(--pending_ == 0) done_cv_.notify_all();
        }
    }
    void loop()
    {
        std::uint64_t seen = 0;
        for(;;) {
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return stop_ || generation_ != seen; });
                if(stop_) return;
                seen = generation_;
            }
            work();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex m_, run_mutex_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(std::size_t)>* job_ = nullptr;
    std::size_t n_ = 0, next_ = 0, pending_ = 0;
    std::uint64_t generation_ = 0;
    std::exception_ptr err_;
    bool stop_ = false;
};

// f(begin, end) over [0, n) in up to pool-size chunks when n / min_per_chunk > 1
template <class F>
inline void parallel_ranges(std::size_t n, std::size_t min_per_chunk, F&& f)
{
    WorkerPool& pool = WorkerPool::get();
    const std::size_t nt = std::min<std::size_t>(pool.size(), std::max<std::size_t>(1, n / std::max<std::size_t>(min_per_chunk, 1)));
    if(nt <= 1) {
        f(std::size_t(0), n);
        return;
    }
    pool.run(nt, [&](std::size_t k) { f(n * k / nt, n * (k + 1) / nt); });
}


}  // namespace xp
