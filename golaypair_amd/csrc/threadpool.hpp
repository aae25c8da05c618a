// A process-wide pool of host worker threads for the library's data-parallel
// host loops (table preparation, text building, sorts).  parallel_for(n, f)
// runs f(0..n-1): the calling thread takes indices too and waits only for
// helpers that already started, so a call always completes even when every
// worker is busy (calls may nest or run concurrently from several host threads).
#pragma once

#include <stddef.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace gp {

class ThreadPool {
public:
    explicit ThreadPool(int n) {
        for (int i = 0; i < n; ++i) workers_.emplace_back([this] { run(); });
    }
    ~ThreadPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    int size() const { return (int)workers_.size(); }

    template <class F>
    void parallel_for(size_t n, F f, int max_helpers = 1 << 30) {
        if (n == 0) return;
        struct State {
            std::atomic<size_t> next{0};
            int active = 0;      // helpers that started and have not finished
            bool closed = false; // the caller is done: helpers not yet started do nothing
            std::exception_ptr err;  // first exception thrown by f, rethrown by the caller
            std::mutex mu;
            std::condition_variable cv;
        };
        auto st = std::make_shared<State>();
        std::function<void(size_t)> body = f;
        auto work = [st, body, n] {
            for (size_t i; (i = st->next.fetch_add(1)) < n;) {
                try {
                    body(i);
                } catch (...) {
                    std::lock_guard<std::mutex> l2(st->mu);
                    if (!st->err) st->err = std::current_exception();
                    st->next = n;  // stop handing out indices
                }
            }
        };
        const int helpers = (int)std::min<size_t>(std::min<size_t>(n - 1, (size_t)size()), (size_t)max_helpers);
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (int h = 0; h < helpers; ++h)
                q_.emplace_back([st, work] {
                    {
                        std::lock_guard<std::mutex> l2(st->mu);
                        if (st->closed) return;
                        ++st->active;
                    }
                    work();
                    std::lock_guard<std::mutex> l2(st->mu);
                    if (--st->active == 0) st->cv.notify_all();
                });
        }
        cv_.notify_all();
        work();
        // every index is taken; wait only for helpers still running one
        std::unique_lock<std::mutex> lk(st->mu);
        st->closed = true;
        st->cv.wait(lk, [&] { return st->active == 0; });
        if (st->err) std::rethrow_exception(st->err);
    }

private:
    void run() {
        for (;;) {
            std::function<void()> task;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
                if (stop_ && q_.empty()) return;
                task = std::move(q_.front());
                q_.pop_front();
            }
            task();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> q_;
    bool stop_ = false;
};

// The library's pool: 15 workers plus the calling thread.
inline ThreadPool& pool() {
    static ThreadPool p(15);
    return p;
}

}  // namespace gp
