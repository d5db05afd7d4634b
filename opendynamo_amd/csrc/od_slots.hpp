// Loan-and-return pool of the snapshot slots (the reference's pinned-buffer
// channel: a buffer is taken with pool_rx.recv() and handed back by the IO
// task once its file is written, solver.rs:163-175, 320, 816).  A slot is
// lent to one ticket at a time and comes back only through give_back; lend
// blocks while every slot is out, so a slot is never refilled while its
// previous snapshot is still being read.  Host code only (no HIP): the CPU
// tests build it on its own under the thread sanitizer.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <mutex>

namespace od {

template <int N>
class SlotPool {
   public:
    // Lends a free slot to a new ticket (never 0), waiting until one is free.
    int lend(uint64_t* ticket) {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [this] { return free_slot() >= 0; });
        const int s = free_slot();
        owner_[s] = next_++;
        *ticket = owner_[s];
        return s;
    }
    // The slot lent to `ticket`, or -1 (unknown, or already given back).
    int slot_of(uint64_t ticket) const {
        std::lock_guard<std::mutex> lk(mu_);
        return find(ticket);
    }
    // Returns the ticket's slot to the pool and wakes a waiting lend; false
    // when the ticket holds no slot.
    bool give_back(uint64_t ticket) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            const int s = find(ticket);
            if (s < 0) return false;
            owner_[s] = 0;
        }
        cv_.notify_all();
        return true;
    }
    // Slots currently lent.
    int lent() const {
        std::lock_guard<std::mutex> lk(mu_);
        int n = 0;
        for (int q = 0; q < N; ++q) n += owner_[q] != 0;
        return n;
    }

   private:
    int free_slot() const {
        for (int q = 0; q < N; ++q)
            if (!owner_[q]) return q;
        return -1;
    }
    int find(uint64_t ticket) const {
        for (int q = 0; q < N; ++q)
            if (ticket != 0 && owner_[q] == ticket) return q;
        return -1;
    }
    mutable std::mutex mu_;
    std::condition_variable cv_;
    uint64_t owner_[N] = {};  // ticket holding each slot; 0 = free
    uint64_t next_ = 1;
};

}  // namespace od
