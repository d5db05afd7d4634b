// Host layer (libodhost sources) under AddressSanitizer, UBSan and LeakSanitizer:
// config.toml parsing (valid and malformed), state-file round trips in every
// compression mode and both dtypes, and the loader's error paths.  Built and
// run by tests/test_host_sanitizers.py; argv[1] is a scratch directory.
#include "odhost.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

static int fails = 0;
#define CHECK(c)                                                                              \
    do {                                                                                      \
        if (!(c)) {                                                                           \
            std::printf("FAIL %s:%d %s (%s)\n", __FILE__, __LINE__, #c, odh_last_error());   \
            ++fails;                                                                          \
        }                                                                                     \
    } while (0)

static std::string dir;

static void write_text(const std::string& path, const char* text) {
    FILE* f = std::fopen(path.c_str(), "w");
    std::fputs(text, f);
    std::fclose(f);
}

template <class T>
static void round_trip(const char* kind, const char* comp, uint64_t nx, uint64_t ny, uint64_t nz) {
    const size_t n = 5 * nx * ny * nz * 64;
    std::vector<T> a(n), b(n);
    for (size_t i = 0; i < n; ++i) a[i] = (T)std::sin(0.001 * (double)i) * (T)(1 + i % 7);
    char path[512];
    CHECK(odh_save_state(3, 1.25, kind, dir.c_str(), "st", a.data(), nx, ny, nz, 1.0, 2.0, 3.0, comp, path,
                         sizeof path) == 0);
    uint64_t step = 0;
    double t = 0;
    CHECK(odh_load_state(path, nx, ny, nz, kind, b.data(), n * sizeof(T), &step, &t) == 0);
    CHECK(std::memcmp(a.data(), b.data(), n * sizeof(T)) == 0);
    CHECK(step == 3 && t == 1.25);
    uint64_t u[5];
    double f[4];
    char k[8] = {0};
    CHECK(odh_read_header(path, u, f, k) == 0);
    CHECK(u[0] == nx && u[1] == ny && u[2] == nz && u[3] == 3 && u[4] == 3);
    CHECK(std::strcmp(k, kind) == 0);
    // wrong mesh, wrong dtype, short buffer: errors, not crashes
    CHECK(odh_load_state(path, nx + 1, ny, nz, kind, b.data(), n * sizeof(T), &step, &t) != 0);
    CHECK(odh_load_state(path, nx, ny, nz, sizeof(T) == 4 ? "f64" : "f32", b.data(), n * sizeof(T), &step, &t) != 0);
    CHECK(odh_load_state(path, nx, ny, nz, kind, b.data(), n * sizeof(T) - 1, &step, &t) != 0);
    // a path buffer too small for the file name
    char small[4];
    (void)odh_save_state(4, 1.0, kind, dir.c_str(), "st", a.data(), nx, ny, nz, 1.0, 2.0, 3.0, comp, small,
                         sizeof small);
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    dir = argv[1];
    const char* comps[] = {"none", "gzip1", "gzip9", "blosc-lz4", "blosc-zstd"};
    for (const char* c : comps) {
        round_trip<float>("f32", c, 3, 2, 5);
        round_trip<double>("f64", c, 2, 3, 1);
        round_trip<float>("f32", c, 17, 9, 11);  // several 262,144-value chunks
    }
    CHECK(odh_save_state(1, 0.0, "f32", dir.c_str(), "st", nullptr, 1, 1, 1, 1.0, 1.0, 1.0, "gzip", nullptr, 0) != 0);
    CHECK(odh_load_state((dir + "/missing.h5").c_str(), 1, 1, 1, "f32", nullptr, 0, nullptr, nullptr) != 0);
    write_text(dir + "/garbage.h5", "not hdf5");
    uint64_t u[5];
    double f[4];
    char k[8];
    CHECK(odh_read_header((dir + "/garbage.h5").c_str(), u, f, k) != 0);

    const char* cfgs[] = {
        "[mesh]\nnx=4\nny=4\nnz=4\n[physics]\nmach_bulk=0.1\ninitial_state=\"a.h5\"\n[solver]\nn_steps=1\n"
        "dt=0.001\nflux_scheme=\"slau2\"\n[io]\nsave_freq=1\noutput_dir=\"o\"\nprefix=\"p\"\n[precision]\n"
        "u_real=\"f32\"\nq_real=\"f32\"\nrhs_real=\"f32\"\ncompute_real=\"f32\"\n",
        "[mesh]\nnx=\n",
        "[mesh\nnx=4",
        "garbage = = =\n[[x]]\n\"unterminated\n",
        "",
    };
    for (int i = 0; i < 5; ++i) {
        const std::string p = dir + "/c" + std::to_string(i) + ".toml";
        write_text(p, cfgs[i]);
        char js[8192];
        char tiny[8];
        const int r = odh_config_json(p.c_str(), js, sizeof js);
        if (i == 0)
            CHECK(r == 0);
        else
            CHECK(r != 0);
        (void)odh_config_json(p.c_str(), tiny, sizeof tiny);
    }
    std::printf("done, %d failures\n", fails);
    return fails != 0;
}
