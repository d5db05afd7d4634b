// The CLI's dump pipeline on the CPU: the step loop of open_dynamo
// (opendynamo_amd/host/main.cpp) with save_freq = 1 against a stand-in for
// the device.  Each "step" rewrites a 16^3 f32 state; a snapshot borrows a
// slot of the three-slot pool (od_slots.hpp, the rule od_get_state_async /
// od_release follow), copies the state into the slot's buffer and hands the
// ticket to the CLI's writer pool (writer_pool.hpp), whose threads write the
// file with the real HDF5 writer at gzip level 9 and give the slot back.
// Writes take far longer than a step, so all three slots are out at once and
// the loop waits for a writer, as in the CLI case with slow writes.  Every
// file must hold exactly the state of its step.
//
//   io_pipeline_driver <scratch dir>        exit status 0 = pass
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../opendynamo_amd/csrc/od_slots.hpp"
#include "../../opendynamo_amd/host/state_io.hpp"
#include "../../opendynamo_amd/host/writer_pool.hpp"

namespace {
constexpr uint64_t kN = 16, kSteps = 12;
constexpr uint64_t kValues = 5 * kN * kN * kN * 64;

// The stand-in state after step k: exact in f32, different for every step.
float state_value(uint64_t k, uint64_t i) { return (float)k + 0.25f * (float)(i % 4093); }
}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: io_pipeline_driver <scratch dir>\n");
        return 2;
    }
    const std::string out = argv[1];
    const odh::Compression gz9 = odh::parse_compression("gzip9");
    od::SlotPool<3> pool;
    std::vector<std::vector<float>> slot(3, std::vector<float>(kValues));
    std::vector<float> device(kValues);
    int max_lent = 0;
    const auto t0 = std::chrono::steady_clock::now();
    {
        odh::WriterPool writers(
            [&](uint64_t ticket) -> const void* {
                const int s = pool.slot_of(ticket);
                if (s < 0) throw std::runtime_error("wait: unknown ticket");
                return slot[s].data();
            },
            [&](uint64_t ticket) { return pool.give_back(ticket); });
        for (uint64_t k = 1; k <= kSteps; ++k) {
            for (uint64_t i = 0; i < kValues; ++i) device[i] = state_value(k, i);  // the step
            uint64_t ticket = 0;
            const int s = pool.lend(&ticket);  // blocks while all three slots are out
            std::memcpy(slot[s].data(), device.data(), kValues * sizeof(float));
            const int l = pool.lent();
            if (l > max_lent) max_lent = l;
            writers.spawn(ticket, [k, &out, &gz9](const void* host) {
                odh::save_state_h5(k, 0.002 * (double)k, "f32", out, "flow", host, kN, kN, kN, 6.0, 2.0, 3.0, gz9);
            });
        }
        writers.join_all();
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    int failures = 0;
    for (uint64_t k = 1; k <= kSteps; ++k) {
        char name[64];
        std::snprintf(name, sizeof(name), "/flow_f32_%06llu.h5", (unsigned long long)k);
        try {
            const odh::LoadedState st = odh::load_state_h5(out + name, kN, kN, kN, "f32");
            const float* v = reinterpret_cast<const float*>(st.data.data());
            bool same = st.step == k && st.data.size() == kValues * sizeof(float);
            for (uint64_t i = 0; same && i < kValues; ++i) same = v[i] == state_value(k, i);
            if (!same) {
                std::printf("dump %llu differs from the state of its step\n", (unsigned long long)k);
                ++failures;
            }
        } catch (const std::exception& e) {
            std::printf("dump %llu: %s\n", (unsigned long long)k, e.what());
            ++failures;
        }
    }
    if (max_lent != 3) {
        std::printf("the three slots were never all out at once (max %d): writes too fast to test waiting\n",
                    max_lent);
        ++failures;
    }
    std::printf("%llu dumps in %.2f s, max slots out %d; done, %d failures\n", (unsigned long long)kSteps, secs,
                max_lent, failures);
    return failures ? 1 : 0;
}
