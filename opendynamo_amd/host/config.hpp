// config.toml reader: the keys, types and defaults of drunkseal/opendynamo's
// src/config.rs:1-209 (serde structs), plus optional keys with no reference
// counterpart ([physics] body_force_x, [solver] cfl, [precision] rk_accum).
#pragma once
#include <cstdint>
#include <map>
#include <optional>
#include <string>

namespace odh {

enum class Real { F32 = 4, F64 = 8 };

struct MeshConfig {  // config.rs:12-38
    uint64_t nx = 0, ny = 0, nz = 0;
    float lx = 4.0f * 3.14159265358979323846f;
    float ly = 2.0f;
    float lz = (4.0f / 3.0f) * 3.14159265358979323846f;
};

struct PhysicsConfig {  // config.rs:101-142
    float mach_bulk = 0.f;
    std::string initial_condition = "channel";  // channel | uniform | fourier (new)
    std::string y_boundary = "wall";            // wall | periodic
    std::string viscosity_model = "fixed";      // fixed | sutherland
    double mu_ref = 0.00016;                    // read in f64 (see DESIGN)
    float t_ref = 273.15f;
    float s_suth = 110.4f;
    float perturbation_amplitude = 0.f;
    uint32_t perturbation_modes = 16;
    uint32_t perturbation_seed = 7;
    std::optional<std::string> initial_state;
    double body_force_x = 0.001;  // no reference counterpart (rhs.cu:508 hard-codes it)
};

struct SolverConfig {  // config.rs:153-160
    uint64_t n_steps = 0;
    float dt = 0.f;
    bool has_dt = false;
    std::string flux_scheme;  // slau2 | ausm+ | rusanov
    bool fast_math = true;
    std::optional<double> cfl;  // no reference counterpart: dt from the initial state
};

struct IOConfig {  // config.rs:177-203
    uint64_t save_freq = 0;
    std::string output_dir;
    std::string prefix;
    uint64_t checkpoint_freq = 0;
    std::string compression = "blosc-lz4";
};

struct PrecisionConfig {  // config.rs:169-175
    Real u_real = Real::F32, q_real = Real::F32, rhs_real = Real::F32, compute_real = Real::F32;
    std::string rk_accum = "compute";  // compute | state (no reference counterpart)
};

struct Config {
    MeshConfig mesh;
    PhysicsConfig physics;
    SolverConfig solver;
    IOConfig io;
    PrecisionConfig precision;
};

// A parsed TOML document restricted to what config.toml uses: [section]
// tables of key = value with strings, integers, floats and booleans.
struct TomlValue {
    enum Kind { String, Integer, Float, Bool } kind;
    std::string s;
    int64_t i = 0;
    double f = 0.0;
    bool b = false;
};
using TomlDoc = std::map<std::string, std::map<std::string, TomlValue>>;

TomlDoc parse_toml(const std::string& text);          // throws std::runtime_error
Config config_from_toml(const TomlDoc& doc);          // throws on missing / mistyped keys
Config load_config(const std::string& path);          // config.rs:205-209

const char* real_name(Real r);  // "f32" / "f64"

}  // namespace odh
