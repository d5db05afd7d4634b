#include "config.hpp"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace odh {
namespace {

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

[[noreturn]] void fail(int line, const std::string& msg) {
    throw std::runtime_error("config.toml line " + std::to_string(line) + ": " + msg);
}

// Value text up to an unquoted '#'.
TomlValue parse_value(const std::string& raw, int line) {
    std::string v = trim(raw);
    if (v.empty()) fail(line, "missing value");
    TomlValue out{};
    if (v[0] == '"' || v[0] == '\'') {
        const char q = v[0];
        std::string s;
        size_t i = 1;
        for (; i < v.size() && v[i] != q; ++i) {
            if (q == '"' && v[i] == '\\' && i + 1 < v.size()) {
                char e = v[++i];
                s += e == 'n' ? '\n' : e == 't' ? '\t' : e;
            } else {
                s += v[i];
            }
        }
        if (i >= v.size()) fail(line, "unterminated string");
        std::string rest = trim(v.substr(i + 1));
        if (!rest.empty() && rest[0] != '#') fail(line, "trailing characters after string");
        out.kind = TomlValue::String;
        out.s = s;
        return out;
    }
    size_t hash = v.find('#');
    if (hash != std::string::npos) v = trim(v.substr(0, hash));
    if (v == "true" || v == "false") {
        out.kind = TomlValue::Bool;
        out.b = v == "true";
        return out;
    }
    std::string num;
    for (char c : v)
        if (c != '_') num += c;
    bool is_float = num.find_first_of(".eE") != std::string::npos || num == "inf" || num == "+inf" ||
                    num == "-inf" || num == "nan";
    char* end = nullptr;
    if (is_float) {
        out.kind = TomlValue::Float;
        out.f = std::strtod(num.c_str(), &end);
    } else {
        out.kind = TomlValue::Integer;
        out.i = std::strtoll(num.c_str(), &end, 10);
        out.f = (double)out.i;
    }
    if (num.empty() || end == nullptr || *end != '\0') fail(line, "cannot parse value '" + v + "'");
    return out;
}

const TomlValue* find(const TomlDoc& d, const std::string& sec, const std::string& key) {
    auto s = d.find(sec);
    if (s == d.end()) return nullptr;
    auto k = s->second.find(key);
    return k == s->second.end() ? nullptr : &k->second;
}

[[noreturn]] void missing(const std::string& sec, const std::string& key) {
    throw std::runtime_error("config.toml: missing field `" + key + "` in [" + sec + "]");
}
[[noreturn]] void mistyped(const std::string& sec, const std::string& key, const char* want) {
    throw std::runtime_error("config.toml: [" + sec + "] " + key + ": invalid type, expected " + want);
}

uint64_t get_u64(const TomlDoc& d, const std::string& sec, const std::string& key, const uint64_t* def) {
    const TomlValue* v = find(d, sec, key);
    if (!v) {
        if (!def) missing(sec, key);
        return *def;
    }
    if (v->kind != TomlValue::Integer || v->i < 0) mistyped(sec, key, "a non-negative integer");
    return (uint64_t)v->i;
}

// f32 fields: the TOML number rounded to f32, as serde does.
float get_f32(const TomlDoc& d, const std::string& sec, const std::string& key, const float* def) {
    const TomlValue* v = find(d, sec, key);
    if (!v) {
        if (!def) missing(sec, key);
        return *def;
    }
    if (v->kind != TomlValue::Integer && v->kind != TomlValue::Float) mistyped(sec, key, "f32");
    return (float)v->f;
}

double get_f64(const TomlDoc& d, const std::string& sec, const std::string& key, double def) {
    const TomlValue* v = find(d, sec, key);
    if (!v) return def;
    if (v->kind != TomlValue::Integer && v->kind != TomlValue::Float) mistyped(sec, key, "a number");
    return v->f;
}

std::string get_str(const TomlDoc& d, const std::string& sec, const std::string& key, const char* def) {
    const TomlValue* v = find(d, sec, key);
    if (!v) {
        if (!def) missing(sec, key);
        return def;
    }
    if (v->kind != TomlValue::String) mistyped(sec, key, "a string");
    return v->s;
}

bool get_bool(const TomlDoc& d, const std::string& sec, const std::string& key, bool def) {
    const TomlValue* v = find(d, sec, key);
    if (!v) return def;
    if (v->kind != TomlValue::Bool) mistyped(sec, key, "a boolean");
    return v->b;
}

void check_enum(const std::string& sec, const std::string& key, const std::string& v,
                std::initializer_list<const char*> allowed) {
    for (const char* a : allowed)
        if (v == a) return;
    std::string list;
    for (const char* a : allowed) list += std::string(list.empty() ? "" : ", ") + "`" + a + "`";
    throw std::runtime_error("config.toml: [" + sec + "] " + key + ": unknown variant `" + v + "`, expected one of " +
                             list);
}

Real get_real(const TomlDoc& d, const std::string& key) {
    std::string v = get_str(d, "precision", key, nullptr);
    check_enum("precision", key, v, {"f32", "f64"});
    return v == "f32" ? Real::F32 : Real::F64;
}

}  // namespace

TomlDoc parse_toml(const std::string& text) {
    TomlDoc doc;
    std::istringstream in(text);
    std::string line, section;
    int n = 0;
    while (std::getline(in, line)) {
        ++n;
        std::string t = trim(line);
        if (t.empty() || t[0] == '#') continue;
        if (t[0] == '[') {
            size_t close = t.find(']');
            if (close == std::string::npos) fail(n, "unterminated table header");
            section = trim(t.substr(1, close - 1));
            if (section.empty()) fail(n, "empty table name");
            doc[section];
            continue;
        }
        size_t eq = t.find('=');
        if (eq == std::string::npos) fail(n, "expected key = value");
        std::string key = trim(t.substr(0, eq));
        if (key.size() >= 2 && (key[0] == '"' || key[0] == '\'')) key = key.substr(1, key.size() - 2);
        if (key.empty()) fail(n, "empty key");
        if (doc[section].count(key)) fail(n, "duplicate key `" + key + "`");
        doc[section][key] = parse_value(t.substr(eq + 1), n);
    }
    return doc;
}

Config config_from_toml(const TomlDoc& d) {
    for (const char* sec : {"mesh", "physics", "solver", "io", "precision"})
        if (!d.count(sec)) throw std::runtime_error(std::string("config.toml: missing table [") + sec + "]");
    Config c;
    MeshConfig dm;
    c.mesh.nx = get_u64(d, "mesh", "nx", nullptr);
    c.mesh.ny = get_u64(d, "mesh", "ny", nullptr);
    c.mesh.nz = get_u64(d, "mesh", "nz", nullptr);
    c.mesh.lx = get_f32(d, "mesh", "lx", &dm.lx);
    c.mesh.ly = get_f32(d, "mesh", "ly", &dm.ly);
    c.mesh.lz = get_f32(d, "mesh", "lz", &dm.lz);

    PhysicsConfig dp;
    c.physics.mach_bulk = get_f32(d, "physics", "mach_bulk", nullptr);
    c.physics.initial_condition = get_str(d, "physics", "initial_condition", "channel");
    check_enum("physics", "initial_condition", c.physics.initial_condition, {"channel", "uniform", "fourier"});
    c.physics.y_boundary = get_str(d, "physics", "y_boundary", "wall");
    check_enum("physics", "y_boundary", c.physics.y_boundary, {"wall", "periodic"});
    c.physics.viscosity_model = get_str(d, "physics", "viscosity_model", "fixed");
    check_enum("physics", "viscosity_model", c.physics.viscosity_model, {"fixed", "sutherland"});
    c.physics.mu_ref = get_f64(d, "physics", "mu_ref", dp.mu_ref);
    c.physics.t_ref = get_f32(d, "physics", "t_ref", &dp.t_ref);
    c.physics.s_suth = get_f32(d, "physics", "s_suth", &dp.s_suth);
    c.physics.perturbation_amplitude = get_f32(d, "physics", "perturbation_amplitude", &dp.perturbation_amplitude);
    uint64_t pm = dp.perturbation_modes, ps = dp.perturbation_seed;
    c.physics.perturbation_modes = (uint32_t)get_u64(d, "physics", "perturbation_modes", &pm);
    c.physics.perturbation_seed = (uint32_t)get_u64(d, "physics", "perturbation_seed", &ps);
    if (find(d, "physics", "initial_state")) c.physics.initial_state = get_str(d, "physics", "initial_state", nullptr);
    c.physics.body_force_x = get_f64(d, "physics", "body_force_x", dp.body_force_x);

    c.solver.n_steps = get_u64(d, "solver", "n_steps", nullptr);
    if (find(d, "solver", "cfl")) c.solver.cfl = get_f64(d, "solver", "cfl", 0.0);
    if (find(d, "solver", "dt") || !c.solver.cfl) {
        c.solver.dt = get_f32(d, "solver", "dt", nullptr);
        c.solver.has_dt = true;
    }
    c.solver.flux_scheme = get_str(d, "solver", "flux_scheme", nullptr);
    check_enum("solver", "flux_scheme", c.solver.flux_scheme, {"slau2", "ausm+", "rusanov"});
    c.solver.fast_math = get_bool(d, "solver", "fast_math", true);

    c.io.save_freq = get_u64(d, "io", "save_freq", nullptr);
    c.io.output_dir = get_str(d, "io", "output_dir", nullptr);
    c.io.prefix = get_str(d, "io", "prefix", nullptr);
    uint64_t zero = 0;
    c.io.checkpoint_freq = get_u64(d, "io", "checkpoint_freq", &zero);
    c.io.compression = get_str(d, "io", "compression", "blosc-lz4");

    c.precision.u_real = get_real(d, "u_real");
    c.precision.q_real = get_real(d, "q_real");
    c.precision.rhs_real = get_real(d, "rhs_real");
    c.precision.compute_real = get_real(d, "compute_real");
    c.precision.rk_accum = get_str(d, "precision", "rk_accum", "compute");
    check_enum("precision", "rk_accum", c.precision.rk_accum, {"compute", "state"});
    if (c.io.save_freq == 0) throw std::runtime_error("config.toml: [io] save_freq must be >= 1");
    return c;
}

Config load_config(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot read " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return config_from_toml(parse_toml(ss.str()));
}

const char* real_name(Real r) { return r == Real::F32 ? "f32" : "f64"; }

}  // namespace odh
