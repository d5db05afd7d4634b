// C ABI of the host layer (include/odhost.h).
#include <cstdio>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>

#include "../../include/odhost.h"
#include "config.hpp"
#include "state_io.hpp"

namespace {
thread_local std::string g_err;

int guard(const std::exception& e) {
    g_err = e.what();
    return 1;
}

void copy_out(const std::string& s, char* dst, size_t cap) {
    if (!dst || cap == 0) return;
    size_t n = s.size() < cap - 1 ? s.size() : cap - 1;
    std::memcpy(dst, s.data(), n);
    dst[n] = '\0';
}

std::string jstr(const std::string& s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o + "\"";
}
std::string jnum(double v) {
    char b[64];
    std::snprintf(b, sizeof(b), "%.17g", v);
    return b;
}
}  // namespace

extern "C" {

const char* odh_last_error(void) { return g_err.c_str(); }

int odh_save_state(uint64_t label, double t, const char* kind, const char* out_dir, const char* prefix,
                   const void* data, uint64_t nx, uint64_t ny, uint64_t nz, double lx, double ly, double lz,
                   const char* compression, char* path_out, size_t path_cap) {
    try {
        if (!kind || !out_dir || !prefix || !data || !compression) throw std::runtime_error("null argument");
        std::string p = odh::save_state_h5(label, t, kind, out_dir, prefix, data, nx, ny, nz, lx, ly, lz,
                                           odh::parse_compression(compression));
        copy_out(p, path_out, path_cap);
        return 0;
    } catch (const std::exception& e) {
        return guard(e);
    }
}

int odh_load_state(const char* path, uint64_t nx, uint64_t ny, uint64_t nz, const char* kind, void* data_out,
                   size_t bytes, uint64_t* step, double* t) {
    try {
        if (!path || !kind || !data_out) throw std::runtime_error("null argument");
        odh::LoadedState s = odh::load_state_h5(path, nx, ny, nz, kind);
        if (s.data.size() != bytes) throw std::runtime_error("output buffer size does not match the state size");
        std::memcpy(data_out, s.data.data(), bytes);
        if (step) *step = s.step;
        if (t) *t = s.t;
        return 0;
    } catch (const std::exception& e) {
        return guard(e);
    }
}

int odh_read_header(const char* path, uint64_t u64_out[5], double f64_out[4], char* kind_out) {
    try {
        if (!path) throw std::runtime_error("null argument");
        odh::LoadedState s = odh::read_state_h5(path);
        if (u64_out) {
            u64_out[0] = s.nx;
            u64_out[1] = s.ny;
            u64_out[2] = s.nz;
            u64_out[3] = s.fr_order;
            u64_out[4] = s.step;
        }
        if (f64_out) {
            f64_out[0] = s.t;
            f64_out[1] = s.lx;
            f64_out[2] = s.ly;
            f64_out[3] = s.lz;
        }
        copy_out(s.u_real, kind_out, 4);
        return 0;
    } catch (const std::exception& e) {
        return guard(e);
    }
}

int odh_config_json(const char* path, char* json_out, size_t cap) {
    try {
        if (!path) throw std::runtime_error("null argument");
        odh::Config c = odh::load_config(path);
        std::ostringstream o;
        o << "{\"mesh\":{\"nx\":" << c.mesh.nx << ",\"ny\":" << c.mesh.ny << ",\"nz\":" << c.mesh.nz
          << ",\"lx\":" << jnum(c.mesh.lx) << ",\"ly\":" << jnum(c.mesh.ly) << ",\"lz\":" << jnum(c.mesh.lz) << "},";
        o << "\"physics\":{\"mach_bulk\":" << jnum(c.physics.mach_bulk)
          << ",\"initial_condition\":" << jstr(c.physics.initial_condition)
          << ",\"y_boundary\":" << jstr(c.physics.y_boundary) << ",\"viscosity_model\":" << jstr(c.physics.viscosity_model)
          << ",\"mu_ref\":" << jnum(c.physics.mu_ref) << ",\"t_ref\":" << jnum(c.physics.t_ref)
          << ",\"s_suth\":" << jnum(c.physics.s_suth)
          << ",\"perturbation_amplitude\":" << jnum(c.physics.perturbation_amplitude)
          << ",\"perturbation_modes\":" << c.physics.perturbation_modes
          << ",\"perturbation_seed\":" << c.physics.perturbation_seed << ",\"initial_state\":"
          << (c.physics.initial_state ? jstr(*c.physics.initial_state) : "null")
          << ",\"body_force_x\":" << jnum(c.physics.body_force_x) << "},";
        o << "\"solver\":{\"n_steps\":" << c.solver.n_steps << ",\"dt\":" << (c.solver.has_dt ? jnum(c.solver.dt) : "null")
          << ",\"flux_scheme\":" << jstr(c.solver.flux_scheme) << ",\"fast_math\":" << (c.solver.fast_math ? "true" : "false")
          << ",\"cfl\":" << (c.solver.cfl ? jnum(*c.solver.cfl) : "null") << "},";
        o << "\"io\":{\"save_freq\":" << c.io.save_freq << ",\"output_dir\":" << jstr(c.io.output_dir)
          << ",\"prefix\":" << jstr(c.io.prefix) << ",\"checkpoint_freq\":" << c.io.checkpoint_freq
          << ",\"compression\":" << jstr(c.io.compression) << "},";
        o << "\"precision\":{\"u_real\":" << jstr(odh::real_name(c.precision.u_real))
          << ",\"q_real\":" << jstr(odh::real_name(c.precision.q_real))
          << ",\"rhs_real\":" << jstr(odh::real_name(c.precision.rhs_real))
          << ",\"compute_real\":" << jstr(odh::real_name(c.precision.compute_real))
          << ",\"rk_accum\":" << jstr(c.precision.rk_accum) << "}}";
        std::string s = o.str();
        if (s.size() + 1 > cap) throw std::runtime_error("json buffer too small");
        copy_out(s, json_out, cap);
        return 0;
    } catch (const std::exception& e) {
        return guard(e);
    }
}

}  // extern "C"
