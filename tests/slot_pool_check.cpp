// CPU check of the snapshot slot pool (opendynamo_amd/csrc/od_slots.hpp), the
// loan-and-return rule of od_get_state_async / od_wait / od_release
// (solver.rs:163-175, 320, 816).  A producer fills each lent slot with its
// snapshot number while slow consumers hold the slots and read them
// repeatedly: every consumer must see its own snapshot intact until it gives
// the slot back, so the producer must block whenever all slots are out.  A
// given-back or unknown ticket must be refused, and a full pool must block a
// lend until a slot comes back.  Built with the thread sanitizer by
// tests/test_slot_pool.py.  Exit status 0 = pass.
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

#include "../opendynamo_amd/csrc/od_slots.hpp"

namespace {
constexpr int kSlots = 3;
constexpr int kWords = 4096;
constexpr uint64_t kSnapshots = 200;
}  // namespace

int main() {
    od::SlotPool<kSlots> pool;
    std::vector<std::vector<uint64_t>> buf(kSlots, std::vector<uint64_t>(kWords, 0));
    std::atomic<int> torn{0}, misuse{0}, max_lent{0};
    std::vector<std::thread> consumers;
    for (uint64_t k = 1; k <= kSnapshots; ++k) {
        uint64_t t = 0;
        const int s = pool.lend(&t);  // blocks while all slots are held
        if (s < 0 || s >= kSlots || t == 0) {
            std::fprintf(stderr, "lend returned slot %d ticket %llu\n", s, (unsigned long long)t);
            return 1;
        }
        for (uint64_t& w : buf[s]) w = k;  // the snapshot
        const int l = pool.lent();
        for (int m = max_lent.load(); l > m && !max_lent.compare_exchange_weak(m, l);) {
        }
        consumers.emplace_back([&pool, &buf, &torn, &misuse, s, t, k] {
            std::mt19937 rng((unsigned)k);
            const int reads = 1 + (int)(rng() % 4);
            for (int r = 0; r < reads; ++r) {  // a slow writer reading its slot
                std::this_thread::sleep_for(std::chrono::microseconds(200 + rng() % 1500));
                for (uint64_t w : buf[s])
                    if (w != k) {
                        ++torn;
                        break;
                    }
            }
            if (pool.slot_of(t) != s) ++misuse;
            if (!pool.give_back(t)) ++misuse;
            if (pool.give_back(t)) ++misuse;  // a second give-back is refused
            if (pool.slot_of(t) != -1) ++misuse;
        });
    }
    for (std::thread& th : consumers) th.join();

    // a full pool blocks lend until a slot comes back
    uint64_t held[kSlots];
    for (uint64_t& h : held) pool.lend(&h);
    std::atomic<bool> got{false};
    uint64_t late = 0;
    std::thread waiter([&] {
        pool.lend(&late);
        got = true;
    });
    std::this_thread::sleep_for(std::chrono::milliseconds(50));
    const bool blocked = !got.load();
    pool.give_back(held[1]);
    waiter.join();
    const bool handed_over = got.load() && pool.slot_of(late) == 1;
    const bool refuses_unknown = !pool.give_back(0) && !pool.give_back(987654321ull);

    std::printf("snapshots %llu torn %d misuse %d max_lent %d blocked %d handed_over %d refuses_unknown %d\n",
                (unsigned long long)kSnapshots, torn.load(), misuse.load(), max_lent.load(), (int)blocked,
                (int)handed_over, (int)refuses_unknown);
    const bool ok = torn == 0 && misuse == 0 && max_lent == kSlots && blocked && handed_over && refuses_unknown;
    return ok ? 0 : 1;
}
