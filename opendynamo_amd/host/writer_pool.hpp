// Background writers of the dumps (spawn_io_task, solver.rs:745-818).  Each
// writer holds its pinned snapshot slot from `wait` until its file is written
// and then hands the slot back with `release`, on every path, so the step
// loop's next snapshot never refills a buffer that is still being read.  The
// destructor joins every writer, so writers finish before whatever their
// callables use is destroyed, also when an exception leaves the step loop.
// Host code only: the CPU tests drive it with a stand-in for the device.
#pragma once

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <deque>
#include <exception>
#include <functional>
#include <memory>
#include <thread>
#include <utility>

namespace odh {

class WriterPool {
   public:
    // wait(ticket): the snapshot's host buffer once its copy has landed
    // (throws on failure); release(ticket): hands the slot back, false on
    // failure (reported, not thrown: the writer thread ends either way).
    using Wait = std::function<const void*(uint64_t)>;
    using Release = std::function<bool(uint64_t)>;

    WriterPool(Wait wait, Release release) : wait_(std::move(wait)), release_(std::move(release)) {}
    WriterPool(const WriterPool&) = delete;
    WriterPool& operator=(const WriterPool&) = delete;
    ~WriterPool() { join_all(); }

    template <class F>
    void spawn(uint64_t ticket, F write) {
        reap();
        auto done = std::make_shared<std::atomic<bool>>(false);
        writers_.push_back({std::thread([this, ticket, write, done]() {
                                try {
                                    write(wait_(ticket));
                                } catch (const std::exception& e) {
                                    std::fprintf(stderr, "IO task failed: %s\n", e.what());
                                }
                                release_(ticket);
                                done->store(true, std::memory_order_release);
                            }),
                            done});
    }

    void join_all() {
        for (auto& w : writers_)
            if (w.th.joinable()) w.th.join();
        writers_.clear();
    }

   private:
    struct Writer {
        std::thread th;
        std::shared_ptr<std::atomic<bool>> done;
    };
    // join the writers that have finished; at most as many as there are
    // snapshot slots are unfinished at a time, so the list stays short
    void reap() {
        for (auto it = writers_.begin(); it != writers_.end();) {
            if (it->done->load(std::memory_order_acquire)) {
                it->th.join();
                it = writers_.erase(it);
            } else {
                ++it;
            }
        }
    }
    Wait wait_;
    Release release_;
    std::deque<Writer> writers_;
};

}  // namespace odh
