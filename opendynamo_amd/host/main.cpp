// open_dynamo — command-line driver (drunkseal/opendynamo src/main.rs and the
// host loop of src/solver.rs:139-374).  Reads ./config.toml (or the path in
// argv[1]), loads [physics].initial_state, saves the startup state, advances
// n_steps RK3 steps on the GPU through libopendynamo_hip, and writes dumps and
// checkpoints from background IO threads with the reference's labels.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/opendynamo.h"
#include "config.hpp"
#include "state_io.hpp"
#include "writer_pool.hpp"

namespace {

void od_check(od_status s, const char* what) {
    if (s != OD_OK) throw std::runtime_error(std::string(what) + ": " + od_last_error());
}

const char* flux_debug_name(const std::string& s) {  // main.rs:17 prints the enum
    return s == "slau2" ? "Slau2" : s == "ausm+" ? "AusmPlus" : "Rusanov";
}



// solver.rs:820-839: min/max of rho and u over the final state, in f32.
template <class T>
void physical_state_report(const T* u, uint64_t np) {
    float rmin = 3.4e38f, rmax = -3.4e38f, umin = 3.4e38f, umax = -3.4e38f;
    for (uint64_t q = 0; q < np; ++q) {
        float r = (float)u[q];
        float vel = (float)u[np + q] / r;
        rmin = std::fmin(rmin, r);
        rmax = std::fmax(rmax, r);
        umin = std::fmin(umin, vel);
        umax = std::fmax(umax, vel);
    }
    std::printf("Physical State Check (Final):\n  Rho: [%.5f, %.5f]\n  U:   [%.5f, %.5f]\n", rmin, rmax, umin, umax);
}

int run(const std::string& cfg_path) {
    odh::Config cfg = odh::load_config(cfg_path);
    std::printf("Configuring pipeline with precisions:\n  U:       %s\n  Q:       %s\n  RHS:     %s\n  COMPUTE: %s\n"
                "  SCHEME:  %s\n",
                cfg.precision.u_real == odh::Real::F32 ? "F32" : "F64",
                cfg.precision.q_real == odh::Real::F32 ? "F32" : "F64",
                cfg.precision.rhs_real == odh::Real::F32 ? "F32" : "F64",
                cfg.precision.compute_real == odh::Real::F32 ? "F32" : "F64", flux_debug_name(cfg.solver.flux_scheme));

    od_params p;
    od_default_params(&p);
    p.nx = (int32_t)cfg.mesh.nx;
    p.ny = (int32_t)cfg.mesh.ny;
    p.nz = (int32_t)cfg.mesh.nz;
    p.lx = (double)cfg.mesh.lx;
    p.ly = (double)cfg.mesh.ly;
    p.lz = (double)cfg.mesh.lz;
    p.u_real = (int32_t)cfg.precision.u_real;
    p.q_real = (int32_t)cfg.precision.q_real;
    p.rhs_real = (int32_t)cfg.precision.rhs_real;
    p.compute_real = (int32_t)cfg.precision.compute_real;
    p.flux_scheme = cfg.solver.flux_scheme == "slau2" ? OD_FLUX_SLAU2
                    : cfg.solver.flux_scheme == "ausm+" ? OD_FLUX_AUSMPLUS
                                                        : OD_FLUX_RUSANOV;
    p.y_wall = cfg.physics.y_boundary == "wall" ? 1 : 0;
    p.viscosity_model = cfg.physics.viscosity_model == "sutherland" ? 1 : 0;
    p.viscous = cfg.physics.mu_ref != 0.0 ? 1 : 0;  // mu_ref = 0: Euler
    p.mu_ref = cfg.physics.mu_ref;
    p.t_ref = (double)cfg.physics.t_ref;
    p.s_suth = (double)cfg.physics.s_suth;
    p.body_force_x = cfg.physics.body_force_x;
    p.rk_accum = cfg.precision.rk_accum == "state" ? 1 : 0;
    p.fast_math = cfg.solver.fast_math ? 1 : 0;

    const uint64_t E = cfg.mesh.nx * cfg.mesh.ny * cfg.mesh.nz, NP = E * 64, NV = 5 * NP;
    std::printf("Allocating device memory for %llu elements, total DOFs: %llu\n", (unsigned long long)E,
                (unsigned long long)NP);
    od_ctx* ctx = nullptr;
    od_check(od_create(&p, 0, &ctx), "od_create");
    struct Destroy {
        od_ctx* c;
        ~Destroy() { od_destroy(c); }
    } destroy{ctx};

    if (!cfg.physics.initial_state)
        throw std::runtime_error(
            "[physics].initial_state is required: point it at a solver state file (a fresh IC from "
            "python -m opendynamo_amd.make_initial_state, or a checkpoint <output_dir>/checkpoint_..._.h5 to resume from).");
    const std::string kind = odh::real_name(cfg.precision.u_real);
    std::printf("Loading initial state from %s ...\n", cfg.physics.initial_state->c_str());
    odh::LoadedState st = odh::load_state_h5(*cfg.physics.initial_state, cfg.mesh.nx, cfg.mesh.ny, cfg.mesh.nz, kind);
    od_check(od_set_state(ctx, st.data.data(), NV, kind == "f32" ? OD_F32 : OD_F64), "od_set_state");
    std::printf("Initial state loaded: label %llu (t = %.6f).\n", (unsigned long long)st.step, st.t);
    const uint64_t base_step = st.step;
    const double base_t = st.t;
    const odh::Compression comp = odh::parse_compression(cfg.io.compression);
    const double lx = (double)cfg.mesh.lx, ly = (double)cfg.mesh.ly, lz = (double)cfg.mesh.lz;

    // startup state saved at the base label (solver.rs:228-254)
    odh::save_state_h5(base_step, base_t, kind, cfg.io.output_dir, cfg.io.prefix, st.data.data(), cfg.mesh.nx,
                       cfg.mesh.ny, cfg.mesh.nz, lx, ly, lz, comp);

    float dt = cfg.solver.dt;
    if (cfg.solver.cfl) {
        double d = 0.0;
        od_check(od_cfl_dt(ctx, *cfg.solver.cfl, &d), "od_cfl_dt");
        dt = (float)d;
        std::printf("dt from CFL %.3g: %.9g\n", *cfg.solver.cfl, (double)dt);
    }
    const uint64_t n_steps = cfg.solver.n_steps;
    // declared after `destroy`: the writers are joined before the context goes
    odh::WriterPool writers(
        [ctx](uint64_t ticket) {
            const void* host = nullptr;
            od_check(od_wait(ctx, ticket, &host), "od_wait");
            return host;
        },
        [ctx](uint64_t ticket) {
            if (od_release(ctx, ticket) == OD_OK) return true;
            std::fprintf(stderr, "IO task failed: od_release: %s\n", od_last_error());
            return false;
        });
    std::printf("Starting RK3 explicit time integration...\n");
    const auto t0 = std::chrono::steady_clock::now();
    for (uint64_t step = 0; step < n_steps; ++step) {
        od_check(od_step(ctx, (double)dt, 1), "od_step");
        if (step % cfg.io.save_freq == 0 || step == n_steps - 1) {
            uint64_t ticket = 0;
            // blocks while all 3 pinned slots are still held by writers
            od_check(od_get_state_async(ctx, &ticket), "od_get_state_async");
            const bool want_ckpt = step == n_steps - 1 || (cfg.io.checkpoint_freq > 0 && step % cfg.io.checkpoint_freq == 0);
            // labels count completed steps; t = base_t + (k+1) dt (solver.rs:769-810)
            const uint64_t label = base_step + step + 1;
            const double t = base_t + ((double)step + 1.0) * (double)dt;
            writers.spawn(ticket, [=, &cfg](const void* host) {
                odh::save_state_h5(label, t, kind, cfg.io.output_dir, cfg.io.prefix, host, cfg.mesh.nx, cfg.mesh.ny,
                                   cfg.mesh.nz, lx, ly, lz, comp);
                if (want_ckpt)
                    odh::save_state_h5(label, t, kind, cfg.io.output_dir, "checkpoint", host, cfg.mesh.nx,
                                       cfg.mesh.ny, cfg.mesh.nz, lx, ly, lz, comp);
            });
        }
    }
    writers.join_all();
    od_check(od_synchronize(ctx), "od_synchronize");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Simulation finished in %.3f seconds.\n", secs);
    std::printf("Time per step: %.3f ms\n", n_steps ? secs * 1000.0 / (double)n_steps : 0.0);

    std::vector<unsigned char> fin(NV * (kind == "f32" ? 4 : 8));
    od_check(od_get_state(ctx, fin.data(), NV, kind == "f32" ? OD_F32 : OD_F64), "od_get_state");
    if (kind == "f32")
        physical_state_report((const float*)fin.data(), NP);
    else
        physical_state_report((const double*)fin.data(), NP);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(argc > 1 ? argv[1] : "config.toml");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
}
