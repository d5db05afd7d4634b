// Host-side pieces shared by the single-box (od_api.hip) and the ensemble
// (od_batch.hip) front ends: error reporting, the FR operators, stage
// arguments, CFL step and type conversion.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/opendynamo.h"
#include "od_internal.hpp"

namespace od::host {

// Sets the thread-local message read by od_last_error and returns s.
od_status fail(od_status s, const std::string& msg);
const char* last_error();

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            return ::od::host::fail(OD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// FR P = 3 operators in f64 (fr_math.rs:3-142) and GL weights.
struct HostOps {
    double x[4], D[16], eL[4], eR[4], gL[4], gR[4], w[4];
};
HostOps fr_operators();

template <class C>
od::Ops<C> cast_ops(const HostOps& h) {
    od::Ops<C> o;
    for (int q = 0; q < 16; ++q) o.D[q] = (C)h.D[q];
    for (int q = 0; q < 4; ++q) {
        o.eL[q] = (C)h.eL[q];
        o.eR[q] = (C)h.eR[q];
        o.gL[q] = (C)h.gL[q];
        o.gR[q] = (C)h.gR[q];
    }
    return o;
}

inline bool real_ok(int r) { return r == OD_F32 || r == OD_F64; }

template <class Dst, class Src>
void convert(Dst* d, const Src* s, size_t n) {
    for (size_t q = 0; q < n; ++q) d[q] = (Dst)s[q];
}

// SSP-RK3 Shu-Osher coefficients u = a u^n + b u^(k) + c dt L (solver.rs:275-316).
struct Coeffs {
    double a, b, cm;
    bool stage1;
};
constexpr Coeffs kStage[3] = {{1.0, 0.0, 1.0, true}, {0.75, 0.25, 0.25, false}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, false}};

// Per-box scalars, formed exactly as the reference forms them.
template <class CT>
CT jacobian(double l, int n) {  // jac = 2 / (L / n) in the compute type (rhs.cu:366-369; rk.cu:93-95)
    return ((CT)2.0) / (((CT)l) / ((CT)n));
}
template <class T>
T stage_dt_coef(int stage, double dt) {  // c * dt in type T from the f32 dt (solver.rs:256)
    return (T)kStage[stage].cm * (T)dt;
}

// Kernel arguments of one stage for a box whose state buffers are u[0], u[1]
// and whose traces / fv / debug arrays start at the given pointers.
// Buffer roles: stage 1 u0 -> u1, stage 2 u1 -> u1, stage 3 u1 -> u0; traces
// tr[s] -> tr[s+1 mod 3].
template <class UT, class CT>
od::StageArgs<UT, CT> make_stage_args(const od_params& p, int nx, int ny, int nz, double lx, double ly, double lz,
                                      void* const u[2], void* const tr[3], void* fv, void* dbg, int stage,
                                      double dt) {
    od::StageArgs<UT, CT> a{};
    UT* u0 = (UT*)u[0];
    UT* u1 = (UT*)u[1];
    a.uin = stage == 0 ? u0 : u1;
    a.uprev = u0;
    a.uout = stage == 2 ? u0 : u1;
    a.tr_in = (const CT*)tr[stage];
    a.tr_out = (CT*)tr[(stage + 1) % 3];
    a.fv = (CT*)fv;
    a.dbg = (CT*)dbg;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.wall = p.y_wall ? 1 : 0;
    a.flux = p.flux_scheme;
    a.sutherland = p.viscosity_model == 1 ? 1 : 0;
    a.round_q = (sizeof(CT) == 8 && p.q_real == OD_F32) ? 1 : 0;
    a.round_rhs = (sizeof(CT) == 8 && p.rhs_real == OD_F32) ? 1 : 0;
    a.accum_state = p.rk_accum ? 1 : 0;
    a.stage1 = kStage[stage].stage1 ? 1 : 0;
    a.mode = 0;
    a.jx = jacobian<CT>(lx, nx);
    a.jy = jacobian<CT>(ly, ny);
    a.jz = jacobian<CT>(lz, nz);
    a.mu = (CT)p.mu_ref;
    a.mu_ref = (CT)p.mu_ref;
    a.t_ref = (CT)p.t_ref;
    a.s_suth = (CT)p.s_suth;
    a.bf = (CT)p.body_force_x;
    const Coeffs& k = kStage[stage];
    a.ca = (CT)k.a;
    a.cb = (CT)k.b;
    a.cc = stage_dt_coef<CT>(stage, dt);
    a.ua = (UT)k.a;
    a.ub = (UT)k.b;
    a.uc = stage_dt_coef<UT>(stage, dt);
    return a;
}

// dt = cfl / ((2P+1) max_points sum_d (|u_d| + c) / dx_d), rounded to f32,
// from an f64 state [5][NP].
double cfl_dt_from_state(const double* h, int64_t NP, double dx, double dy, double dz, double cfl);

}  // namespace od::host
