// C-ABI front end of the HIP library (include/opendynamo.h).
//
// Owns the device buffers, the compute stream, an IO stream with three
// pinned staging slots, and one captured hipGraph per RK3 step.  Replaces the
// orchestration of drunkseal/opendynamo's src/solver.rs (context, NVRTC
// module, operators, allocations, launch configs, the RK3 loop with its
// buffer swap and the async D2H of the IO task).
//
// Buffer roles over one step (no host-side swap; the state ends in u0):
//   stage 1: u0 -> u1          traces tr[0] -> tr[1]
//   stage 2: u1 -> u1 (u^n=u0) traces tr[1] -> tr[2]
//   stage 3: u1 -> u0 (u^n=u0) traces tr[2] -> tr[0]
// Each element's work-group reads and writes only its own state entries, so
// the in-place stage updates are race free; neighbours communicate through
// the trace and viscous-flux buffers only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/opendynamo.h"
#include "od_ctx.hpp"

namespace od::host {

thread_local std::string g_err;

od_status fail(od_status s, const std::string& msg) {
    g_err = msg;
    return s;
}

const char* last_error() { return g_err.c_str(); }

// ---- FR operators for P = 3 (fr_math.rs:3-142), f64 host arithmetic ------
double legendre(int p, double x) {
    if (p == 0) return 1.0;
    if (p == 1) return x;
    double a = 1.0, b = x, c = 0.0;
    for (int k = 2; k <= p; ++k) {
        double kf = (double)k;
        c = ((2.0 * kf - 1.0) * x * b - (kf - 1.0) * a) / kf;
        a = b;
        b = c;
    }
    return c;
}
double legendre_d(int p, double x) {
    if (p == 0) return 0.0;
    double pf = (double)p;
    return pf / (x * x - 1.0) * (x * legendre(p, x) - legendre(p - 1, x));
}

HostOps fr_operators() {
    HostOps o{};
    const int P = 3, n = 4;
    for (int i = 0; i < n; ++i) {  // Newton on P_4 from cos(pi (i + 3/4) / (n + 1/2))
        double x = std::cos(M_PI * ((double)i + 0.75) / ((double)n + 0.5));
        while (true) {
            double s = legendre(n, x) / legendre_d(n, x);
            x -= s;
            if (std::fabs(s) < 1e-14) break;
        }
        o.x[i] = x;
    }
    std::sort(o.x, o.x + n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            if (i == j) {
                double s = 0.0;
                for (int k = 0; k < n; ++k)
                    if (k != i) s += 1.0 / (o.x[i] - o.x[k]);
                o.D[i * n + i] = s;
            } else {
                double num = 1.0, den = 1.0;
                for (int k = 0; k < n; ++k)
                    if (k != i && k != j) num *= o.x[i] - o.x[k];
                for (int k = 0; k < n; ++k)
                    if (k != j) den *= o.x[j] - o.x[k];
                o.D[i * n + j] = num / den;
            }
        }
    for (int i = 0; i < n; ++i) {
        double l = 1.0, r = 1.0;
        for (int j = 0; j < n; ++j)
            if (j != i) {
                l *= (-1.0 - o.x[j]) / (o.x[i] - o.x[j]);
                r *= (1.0 - o.x[j]) / (o.x[i] - o.x[j]);
            }
        o.eL[i] = l;
        o.eR[i] = r;
    }
    const double sign = (P % 2 == 0) ? 1.0 : -1.0;
    for (int i = 0; i < n; ++i) {
        double a = legendre_d(P, o.x[i]), b = legendre_d(P + 1, o.x[i]);
        o.gR[i] = 0.5 * (a + b);
        o.gL[i] = 0.5 * sign * (a - b);
        double dl = legendre_d(n, o.x[i]);
        o.w[i] = 2.0 / ((1.0 - o.x[i] * o.x[i]) * dl * dl);
    }
    return o;
}

double cfl_dt_from_state(const double* h, int64_t NP, double dx, double dy, double dz, double cfl) {
    double lam = 0.0;
    for (int64_t q = 0; q < NP; ++q) {
        double r = h[q], ru = h[NP + q], rv = h[2 * NP + q], rw = h[3 * NP + q], e = h[4 * NP + q];
        double ke = 0.5 * (ru * ru + rv * rv + rw * rw) / r;
        double pp = std::fmax(1e-7, 0.4 * (e - ke));
        double cs = std::sqrt(1.4 * pp / r);
        double l = (std::fabs(ru / r) + cs) / dx + (std::fabs(rv / r) + cs) / dy + (std::fabs(rw / r) + cs) / dz;
        lam = std::fmax(lam, l);
    }
    return (double)(float)(cfl / (7.0 * lam));
}

}  // namespace od::host

using namespace od::host;

namespace {

template <class UT, class CT>
od::StageArgs<UT, CT> stage_args(od_ctx* c, int stage, double dt) {
    od::StageArgs<UT, CT> a = make_stage_args<UT, CT>(c->p, c->p.nx, c->p.ny, c->p.nz, c->p.lx, c->p.ly, c->p.lz,
                                                      c->u, c->tr, c->fv, c->dbg, stage, dt);
    a.mu_t = c->mu_t_on ? (const CT*)c->mu_t : nullptr;
    return a;
}

// kind: 0 = visc, 1 = update, 2 = trace (of u0 into tr[0])
template <class UT, class CT>
int launch_kind(od_ctx* c, int kind, int stage, double dt, int mode) {
    od::StageArgs<UT, CT> a = stage_args<UT, CT>(c, stage, dt);
    a.mode = mode;
    if (kind == 2) {
        a.uin = (const UT*)c->u[0];
        a.tr_out = (CT*)c->tr[0];
        // the fast build's traces (one x order, fused arithmetic) come from its own kernel
        if constexpr (sizeof(CT) == 4)
            if (c->p.fast_math) return od::fastmath::launch_trace<UT, CT>(a, c->E, c->stream);
        return od::launch_trace<UT, CT>(a, c->E, c->stream);
    }
    if constexpr (sizeof(CT) == 4) {
        if (c->p.fast_math) {
            if constexpr (sizeof(UT) == 4) {
                if (kind == 0 && c->blocks) return od::fastmath::launch_visc_block(a, c->stream);
                if (kind == 1 && c->blocks)
                    return c->p.viscous ? od::fastmath::launch_update_block<true>(a, true, c->stream)
                                        : od::fastmath::launch_update_block<false>(a, true, c->stream);
            }
            if (kind == 0) return od::fastmath::launch_visc<UT, CT>(a, c->E, c->stream);
            if (c->p.viscous) return od::fastmath::launch_update<UT, CT, true>(a, c->E, c->stream);
            return od::fastmath::launch_update<UT, CT, false>(a, c->E, c->stream);
        }
    }
    if (kind == 0) return od::launch_visc<UT, CT>(a, c->E, c->stream);
    if (c->p.viscous) return od::launch_update<UT, CT, true>(a, c->E, c->stream);
    return od::launch_update<UT, CT, false>(a, c->E, c->stream);
}

int launch(od_ctx* c, int kind, int stage, double dt, int mode = 0) {
    const bool uf = c->p.u_real == OD_F32, cf = c->p.compute_real == OD_F32;
    if (uf && cf) return launch_kind<float, float>(c, kind, stage, dt, mode);
    if (!uf && !cf) return launch_kind<double, double>(c, kind, stage, dt, mode);
    if (!uf && cf) return launch_kind<double, float>(c, kind, stage, dt, mode);
    return launch_kind<float, double>(c, kind, stage, dt, mode);
}

od_status enqueue_step(od_ctx* c, double dt) {
    for (int s = 0; s < 3; ++s) {
        if (c->p.viscous && launch(c, 0, s, dt)) return fail(OD_ERR_HIP, "od_visc_kernel launch failed");
        if (launch(c, 1, s, dt)) return fail(OD_ERR_HIP, "od_update_kernel launch failed");
    }
    return OD_OK;
}

// Steps per replay of the multi-step graph: consecutive graph launches leave
// ~10 us of idle GPU between them, the kernels inside one graph none, and
// every step has the same buffer roles, so n steps run as n / 8 replays of an
// 8-step graph plus n % 8 of the one-step graph (same kernels, same order).
constexpr uint32_t kStepsPerGraph = 8;

od_status capture_steps(od_ctx* c, double dt, uint32_t steps, hipGraphExec_t* out) {
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    od_status s = OD_OK;
    for (uint32_t q = 0; q < steps && s == OD_OK; ++q) s = enqueue_step(c, dt);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (s != OD_OK) {
        if (g) (void)hipGraphDestroy(g);
        return s;
    }
    if (e != hipSuccess) return fail(OD_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(OD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    // the one-time device-side set-up of the graph now, not in its first
    // replay (which otherwise runs ~1 ms long)
    e = hipGraphUpload(*out, c->stream);
    if (e != hipSuccess) return fail(OD_ERR_HIP, std::string("hipGraphUpload: ") + hipGetErrorString(e));
    return OD_OK;
}

od_status ensure_graph(od_ctx* c, double dt) {
    if (c->graph && c->graph_k && c->graph_dt == dt && c->graph_mu_t == c->mu_t_on) return OD_OK;
    for (hipGraphExec_t* gx : {&c->graph, &c->graph_k}) {
        if (*gx) {
            HIP_TRY(hipGraphExecDestroy(*gx));
            *gx = nullptr;
        }
    }
    od_status s = capture_steps(c, dt, 1, &c->graph);
    if (s == OD_OK) s = capture_steps(c, dt, kStepsPerGraph, &c->graph_k);
    if (s != OD_OK) return s;
    c->graph_dt = dt;
    c->graph_mu_t = c->mu_t_on;
    return OD_OK;
}

// Enqueue n steps of the captured graphs on the library stream.
od_status launch_steps(od_ctx* c, uint32_t n) {
    uint32_t q = 0;
    for (; q + kStepsPerGraph <= n; q += kStepsPerGraph) HIP_TRY(hipGraphLaunch(c->graph_k, c->stream));
    for (; q < n; ++q) HIP_TRY(hipGraphLaunch(c->graph, c->stream));
    return OD_OK;
}

}  // namespace

extern "C" {

const char* od_last_error(void) { return last_error(); }

void od_default_params(od_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->nx = p->ny = p->nz = 1;
    p->lx = (double)(4.0f * 3.14159265358979323846f);          // config.rs:30-32
    p->ly = 2.0;                                               // config.rs:33-35
    p->lz = (double)((4.0f / 3.0f) * 3.14159265358979323846f); // config.rs:36-38
    p->u_real = p->q_real = p->rhs_real = p->compute_real = OD_F32;
    p->flux_scheme = OD_FLUX_SLAU2;
    p->y_wall = 1;
    p->viscous = 1;
    p->viscosity_model = 0;
    p->mu_ref = 0.00016;  // constants.cuh:42, config.rs:77-79
    p->t_ref = 273.15;
    p->s_suth = 110.4;
    p->body_force_x = 0.001;  // rhs.cu:508
    p->rk_accum = 0;
    p->fast_math = 1;
}

od_status od_create(const od_params* p, int device, od_ctx** out) {
    if (!p || !out) return fail(OD_ERR_INVALID, "od_create: null argument");
    *out = nullptr;
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail(OD_ERR_INVALID, "mesh dimensions must be >= 1");
    if (!real_ok(p->u_real) || !real_ok(p->q_real) || !real_ok(p->rhs_real) || !real_ok(p->compute_real))
        return fail(OD_ERR_INVALID, "precision types must be OD_F32 or OD_F64");
    if (p->flux_scheme < 0 || p->flux_scheme > 2) return fail(OD_ERR_INVALID, "flux_scheme must be 0..2");
    if (p->viscosity_model < 0 || p->viscosity_model > 1) return fail(OD_ERR_INVALID, "viscosity_model must be 0 or 1");
    if (!(p->lx > 0 && p->ly > 0 && p->lz > 0)) return fail(OD_ERR_INVALID, "domain extents must be positive");
    const int64_t E = (int64_t)p->nx * p->ny * p->nz;
    if (E > (int64_t)0x7fffffff) return fail(OD_ERR_INVALID, "too many elements for one launch grid");
    if (p->ny > 65535 || p->nz > 65535) return fail(OD_ERR_INVALID, "ny and nz must be <= 65535 (launch grid)");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(OD_ERR_INVALID, "no such HIP device");
    HIP_TRY(hipSetDevice(device));

    od_ctx* c = new od_ctx();
    c->p = *p;
    c->device = device;
    c->E = E;
    c->NP = E * 64;
    c->ubytes = (size_t)p->u_real;
    c->cbytes = (size_t)p->compute_real;
    // fast f32 with even dimensions: the step kernels on 2x2x2-element blocks
    c->blocks = p->fast_math && p->u_real == OD_F32 && p->compute_real == OD_F32 && p->nx % 2 == 0 &&
                p->ny % 2 == 0 && p->nz % 2 == 0;
    c->ops = fr_operators();
    {
        const od::Ops<float> of = cast_ops<float>(c->ops);
        const od::Ops<double> od64 = cast_ops<double>(c->ops);
        if (od::upload_ops(of, od64) || od::fastmath::upload_ops(of, od64)) {
            delete c;
            return fail(OD_ERR_HIP, "copying the FR operators to the device failed");
        }
    }
    auto cleanup = [&](od_status s) {
        od_destroy(c);
        return s;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->io_stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(OD_ERR_HIP, "stream creation failed"));
    const size_t state_bytes = (size_t)5 * c->NP * c->ubytes;
    const size_t tr_bytes = (size_t)E * od::kTraceStride * c->cbytes;
    const size_t fv_bytes = (size_t)E * od::kFvStride * c->cbytes;
    auto dmalloc = [&](void** ptr, size_t bytes) {
        if (hipMalloc(ptr, bytes) != hipSuccess) return false;
        c->device_bytes += bytes;
        return true;
    };
    bool ok = dmalloc(&c->u[0], state_bytes) && dmalloc(&c->u[1], state_bytes);
    for (int q = 0; q < 3 && ok; ++q) ok = dmalloc(&c->tr[q], tr_bytes);
    ok = ok && (!p->viscous || dmalloc(&c->fv, fv_bytes));
    ok = ok && dmalloc((void**)&c->sums_partial, (size_t)5 * E * sizeof(double)) &&
         dmalloc((void**)&c->sums_out, 8 * sizeof(double));
    if (!ok) return cleanup(fail(OD_ERR_NOMEM, "device allocation failed"));
    for (int q = 0; q < 3; ++q)
        if (hipHostMalloc(&c->pinned[q], state_bytes, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&c->slot_done[q], hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(OD_ERR_NOMEM, "pinned staging allocation failed"));
    *out = c;
    return OD_OK;
}

void od_destroy(od_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->io_stream) (void)hipStreamSynchronize(c->io_stream);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    if (c->graph_k) (void)hipGraphExecDestroy(c->graph_k);
    for (void* q : c->u)
        if (q) (void)hipFree(q);
    for (void* q : c->tr)
        if (q) (void)hipFree(q);
    if (c->fv) (void)hipFree(c->fv);
    if (c->dbg) (void)hipFree(c->dbg);
    if (c->sums_partial) (void)hipFree(c->sums_partial);
    if (c->sums_out) (void)hipFree(c->sums_out);
    od_closure_release_internal(c);
    for (int q = 0; q < 3; ++q) {
        if (c->pinned[q]) (void)hipHostFree(c->pinned[q]);
        if (c->slot_done[q]) (void)hipEventDestroy(c->slot_done[q]);
        if (c->staged[q]) (void)hipFree(c->staged[q]);
        if (c->staged_done[q]) (void)hipEventDestroy(c->staged_done[q]);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->io_stream) (void)hipStreamDestroy(c->io_stream);
    delete c;
}

od_status od_set_state(od_ctx* c, const void* host, size_t n, od_real dtype) {
    if (!c || !host) return fail(OD_ERR_INVALID, "od_set_state: null argument");
    if (n != (size_t)(5 * c->NP)) return fail(OD_ERR_INVALID, "od_set_state: wrong number of values");
    if (!real_ok(dtype)) return fail(OD_ERR_INVALID, "od_set_state: bad dtype");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<unsigned char> tmp;
    const void* src = host;
    if ((int)dtype != c->p.u_real) {
        tmp.resize(n * c->ubytes);
        if (c->p.u_real == OD_F64)
            convert((double*)tmp.data(), (const float*)host, n);
        else
            convert((float*)tmp.data(), (const double*)host, n);
        src = tmp.data();
    }
    HIP_TRY(hipMemcpyAsync(c->u[0], src, n * c->ubytes, hipMemcpyHostToDevice, c->stream));
    if (launch(c, 2, 0, 0.0)) return fail(OD_ERR_HIP, "od_trace_kernel launch failed");
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->state_set = true;
    return OD_OK;
}

od_status od_get_state(od_ctx* c, void* host, size_t n, od_real dtype) {
    if (!c || !host) return fail(OD_ERR_INVALID, "od_get_state: null argument");
    if (n != (size_t)(5 * c->NP)) return fail(OD_ERR_INVALID, "od_get_state: wrong number of values");
    if (!real_ok(dtype)) return fail(OD_ERR_INVALID, "od_get_state: bad dtype");
    HIP_TRY(hipSetDevice(c->device));
    if ((int)dtype == c->p.u_real) {
        HIP_TRY(hipMemcpyAsync(host, c->u[0], n * c->ubytes, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        return OD_OK;
    }
    std::vector<unsigned char> tmp(n * c->ubytes);
    HIP_TRY(hipMemcpyAsync(tmp.data(), c->u[0], n * c->ubytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->p.u_real == OD_F64)
        convert((float*)host, (const double*)tmp.data(), n);
    else
        convert((double*)host, (const float*)tmp.data(), n);
    return OD_OK;
}

od_status od_get_state_async(od_ctx* c, uint64_t* ticket) {
    if (!c || !ticket) return fail(OD_ERR_INVALID, "od_get_state_async: null argument");
    HIP_TRY(hipSetDevice(c->device));
    // Borrow a free pinned slot, blocking until a consumer returns one
    // (pool_rx.recv(), solver.rs:320): a slot is never refilled while its
    // previous snapshot is still being read.
    uint64_t t = 0;
    const int slot = c->slots.lend(&t);
    // On an error below the slot goes back to the pool before returning.
    auto give_back = [c, t](od_status st) {
        c->slots.give_back(t);
        return st;
    };
    const size_t bytes = (size_t)5 * c->NP * c->ubytes;
    if (!c->staged[slot]) {
        if (hipMalloc(&c->staged[slot], bytes) != hipSuccess) {
            c->staged[slot] = nullptr;
            return give_back(fail(OD_ERR_NOMEM, "od_get_state_async: snapshot buffer allocation failed"));
        }
        c->device_bytes += bytes;
        if (hipEventCreateWithFlags(&c->staged_done[slot], hipEventDisableTiming) != hipSuccess)
            return give_back(fail(OD_ERR_HIP, "od_get_state_async: event creation failed"));
    }
    // a consumer may return the slot without waiting for its copy: the slot's
    // previous device-to-host copy must be complete before either buffer is
    // refilled
    hipError_t e = hipEventSynchronize(c->slot_done[slot]);
    // the snapshot in stream order on the compute stream (before the next
    // step rewrites u[0]), then to the pinned slot on the IO stream
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->staged[slot], c->u[0], bytes, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->staged_done[slot], c->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(c->io_stream, c->staged_done[slot], 0);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->pinned[slot], c->staged[slot], bytes, hipMemcpyDeviceToHost, c->io_stream);
    if (e == hipSuccess) e = hipEventRecord(c->slot_done[slot], c->io_stream);
    if (e != hipSuccess)
        return give_back(fail(OD_ERR_HIP, std::string("od_get_state_async: ") + hipGetErrorString(e)));
    *ticket = t;
    return OD_OK;
}

od_status od_wait(od_ctx* c, uint64_t ticket, const void** host_state) {
    if (!c) return fail(OD_ERR_INVALID, "od_wait: null context");
    const int slot = c->slots.slot_of(ticket);
    if (slot < 0) return fail(OD_ERR_STATE, "od_wait: unknown or already released ticket");
    HIP_TRY(hipEventSynchronize(c->slot_done[slot]));
    if (host_state) *host_state = c->pinned[slot];
    return OD_OK;
}

od_status od_release(od_ctx* c, uint64_t ticket) {
    if (!c) return fail(OD_ERR_INVALID, "od_release: null context");
    if (!c->slots.give_back(ticket)) return fail(OD_ERR_STATE, "od_release: unknown or already released ticket");
    return OD_OK;
}

od_status od_step(od_ctx* c, double dt, uint32_t n_steps) {
    if (!c) return fail(OD_ERR_INVALID, "od_step: null context");
    if (!c->state_set) return fail(OD_ERR_STATE, "od_step: state not set");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(OD_ERR_INVALID, "od_step: dt must be positive and finite");
    HIP_TRY(hipSetDevice(c->device));
    od_status s = ensure_graph(c, dt);
    if (s != OD_OK) return s;
    if (!c->mlp) {
        s = launch_steps(c, n_steps);
        if (s == OD_OK) c->steps_done += n_steps;
        return s;
    }
    for (uint32_t q = 0; q < n_steps; ++q) {  // the closure refreshes mu_t between steps
        if (c->steps_done % (uint64_t)c->closure.update_interval == 0) {
            s = od_closure_update_internal(c, dt * c->closure.update_interval);
            if (s != OD_OK) return s;
        }
        HIP_TRY(hipGraphLaunch(c->graph, c->stream));
        c->steps_done += 1;
    }
    return OD_OK;
}

od_status od_synchronize(od_ctx* c) {
    if (!c) return fail(OD_ERR_INVALID, "od_synchronize: null context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipStreamSynchronize(c->io_stream));
    return OD_OK;
}

od_status od_rhs(od_ctx* c, void* host_out, size_t n) {
    if (!c || !host_out) return fail(OD_ERR_INVALID, "od_rhs: null argument");
    if (!c->state_set) return fail(OD_ERR_STATE, "od_rhs: state not set");
    if (n != (size_t)(5 * c->NP)) return fail(OD_ERR_INVALID, "od_rhs: wrong number of values");
    HIP_TRY(hipSetDevice(c->device));
    if (!c->dbg) {
        if (hipMalloc(&c->dbg, (size_t)5 * c->NP * c->cbytes) != hipSuccess)
            return fail(OD_ERR_NOMEM, "od_rhs: allocation failed");
        c->device_bytes += (size_t)5 * c->NP * c->cbytes;
    }
    // stage-1 operands: uin = u0, tr_in = tr[0]; mode 1 writes the corrected RHS
    if (c->p.viscous && launch(c, 0, 0, 1.0)) return fail(OD_ERR_HIP, "od_visc_kernel launch failed");
    if (launch(c, 1, 0, 1.0, 1)) return fail(OD_ERR_HIP, "od_update_kernel launch failed");
    HIP_TRY(hipMemcpyAsync(host_out, c->dbg, n * c->cbytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return OD_OK;
}

od_status od_conservation_sums(od_ctx* c, double out[5]) {
    if (!c || !out) return fail(OD_ERR_INVALID, "od_conservation_sums: null argument");
    HIP_TRY(hipSetDevice(c->device));
    double w3[64];
    for (int pt = 0; pt < 64; ++pt) w3[pt] = c->ops.w[pt & 3] * c->ops.w[(pt >> 2) & 3] * c->ops.w[pt >> 4];
    const double jac = (c->p.lx / c->p.nx) * (c->p.ly / c->p.ny) * (c->p.lz / c->p.nz) / 8.0;
    int rc = c->p.u_real == OD_F32
                 ? od::launch_sums<float>((const float*)c->u[0], c->sums_partial, c->sums_out, c->E, w3, jac, c->stream)
                 : od::launch_sums<double>((const double*)c->u[0], c->sums_partial, c->sums_out, c->E, w3, jac,
                                           c->stream);
    if (rc) return fail(OD_ERR_HIP, "conservation sum kernels failed to launch");
    HIP_TRY(hipMemcpyAsync(out, c->sums_out, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return OD_OK;
}

od_status od_cfl_dt(od_ctx* c, double cfl, double* dt_out) {
    if (!c || !dt_out) return fail(OD_ERR_INVALID, "od_cfl_dt: null argument");
    if (!(cfl > 0.0)) return fail(OD_ERR_INVALID, "od_cfl_dt: cfl must be positive");
    const size_t n = (size_t)5 * c->NP;
    std::vector<double> h(n);
    od_status s = od_get_state(c, h.data(), n, OD_F64);
    if (s != OD_OK) return s;
    *dt_out = cfl_dt_from_state(h.data(), c->NP, c->p.lx / c->p.nx, c->p.ly / c->p.ny, c->p.lz / c->p.nz, cfl);
    return OD_OK;
}

od_status od_time_steps(od_ctx* c, double dt, uint32_t n_steps, double* ms_total) {
    if (!c || !ms_total) return fail(OD_ERR_INVALID, "od_time_steps: null argument");
    if (!c->state_set) return fail(OD_ERR_STATE, "od_time_steps: state not set");
    HIP_TRY(hipSetDevice(c->device));
    od_status s = ensure_graph(c, dt);
    if (s != OD_OK) return s;
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(a, c->stream));
    s = launch_steps(c, n_steps);
    if (s != OD_OK) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return s;
    }
    HIP_TRY(hipEventRecord(b, c->stream));
    HIP_TRY(hipEventSynchronize(b));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, a, b));
    HIP_TRY(hipEventDestroy(a));
    HIP_TRY(hipEventDestroy(b));
    *ms_total = ms;
    return OD_OK;
}

od_status od_time_kernels(od_ctx* c, double dt, uint32_t n_steps, double out_ms[4], const char* names[4],
                          int* n_kinds) {
    if (!c || !out_ms) return fail(OD_ERR_INVALID, "od_time_kernels: null argument");
    if (!c->state_set) return fail(OD_ERR_STATE, "od_time_kernels: state not set");
    HIP_TRY(hipSetDevice(c->device));
    const int kinds = 2;
    std::vector<hipEvent_t> ev;
    std::vector<int> kind_of;
    auto rec = [&](int kind) -> od_status {
        hipEvent_t a, b;
        HIP_TRY(hipEventCreate(&a));
        HIP_TRY(hipEventCreate(&b));
        ev.push_back(a);
        ev.push_back(b);
        kind_of.push_back(kind);
        return OD_OK;
    };
    for (uint32_t q = 0; q < n_steps; ++q)
        for (int s = 0; s < 3; ++s)
            for (int kind = 0; kind < kinds; ++kind) {
                if (kind == 0 && !c->p.viscous) continue;
                od_status st = rec(kind);
                if (st != OD_OK) return st;
                HIP_TRY(hipEventRecord(ev[ev.size() - 2], c->stream));
                if (launch(c, kind, s, dt)) return fail(OD_ERR_HIP, "kernel launch failed");
                HIP_TRY(hipEventRecord(ev.back(), c->stream));
            }
    HIP_TRY(hipStreamSynchronize(c->stream));
    double sum[4] = {0, 0, 0, 0};
    int cnt[4] = {0, 0, 0, 0};
    for (size_t q = 0; q < kind_of.size(); ++q) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ev[2 * q], ev[2 * q + 1]));
        sum[kind_of[q]] += ms;
        cnt[kind_of[q]] += 1;
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    for (int k = 0; k < 4; ++k) {
        out_ms[k] = (k < kinds && cnt[k]) ? sum[k] / cnt[k] : 0.0;
        const char* nm = k < kinds ? od::kernel_name(k) : "";
        if (k == 0 && c->blocks) nm = "od_visc_block_kernel";
        if (k == 1 && c->blocks) nm = "od_update_block_kernel";
        if (names) names[k] = nm;
    }
    if (n_kinds) *n_kinds = kinds;
    return OD_OK;
}

od_status od_info(od_ctx* c, double out[4]) {
    if (!c || !out) return fail(OD_ERR_INVALID, "od_info: null argument");
    out[0] = (double)c->NP;
    out[1] = (double)c->device_bytes;
    out[2] = c->p.viscous ? 6.0 : 3.0;
    out[3] = 0.0;
    return OD_OK;
}

}  // extern "C"

// ---- Conservation sums without allocation -----------------------------------
// od_conservation_sums_noalloc keeps its per-element partial sums ([5][E]
// doubles, the buffer od_create allocates for od_conservation_sums) in this
// fixed array.  Host code never uses the name `scratch` as an address: each
// GPU holds its own copy, whose device address hipGetSymbolAddress returns
// once per GPU (scratch_on below).
static __device__ __align__(256) double scratch[5 * OD_NOALLOC_MAX_ELEMENTS];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current GPU, looked up on the first call
// on that GPU (with that GPU current) and kept.
od_status scratch_on(int device, double** out) {
    if (device < 0 || device >= kMaxDevices) return fail(OD_ERR_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(OD_ERR_HIP, "hipGetSymbolAddress returned a null address for the scratch array");
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[device];
    return OD_OK;
}

// GL weights of the 64 points (tensor products of the 4-point rule), as
// od_conservation_sums forms them.
const double* gl_weights64() {
    static const std::vector<double> w3 = [] {
        const HostOps o = fr_operators();
        std::vector<double> w(64);
        for (int pt = 0; pt < 64; ++pt) w[pt] = o.w[pt & 3] * o.w[(pt >> 2) & 3] * o.w[pt >> 4];
        return w;
    }();
    return w3.data();
}

}  // namespace

extern "C" od_status od_conservation_sums_noalloc(int32_t nx, int32_t ny, int32_t nz, double lx, double ly,
                                                  double lz, const void* d_state, od_real u_real, double* d_out,
                                                  void* stream) {
    if (!d_state || !d_out) return fail(OD_ERR_INVALID, "od_conservation_sums_noalloc: null argument");
    if (nx < 1 || ny < 1 || nz < 1) return fail(OD_ERR_INVALID, "mesh dimensions must be >= 1");
    if (!real_ok(u_real)) return fail(OD_ERR_INVALID, "u_real must be OD_F32 or OD_F64");
    if (!(lx > 0 && ly > 0 && lz > 0)) return fail(OD_ERR_INVALID, "domain extents must be positive");
    const int64_t E = (int64_t)nx * ny * nz;
    if (E > (int64_t)OD_NOALLOC_MAX_ELEMENTS)
        return fail(OD_ERR_INVALID, "od_conservation_sums_noalloc: " + std::to_string(E) +
                                        " elements exceed the maximum of " +
                                        std::to_string((long long)OD_NOALLOC_MAX_ELEMENTS));
    int device = 0;
    HIP_TRY(hipGetDevice(&device));
    double* partial = nullptr;
    od_status s = scratch_on(device, &partial);
    if (s != OD_OK) return s;
    const double jac = (lx / nx) * (ly / ny) * (lz / nz) / 8.0;
    const int rc = u_real == OD_F32 ? od::launch_sums<float>((const float*)d_state, partial, d_out, E, gl_weights64(),
                                                             jac, stream)
                                    : od::launch_sums<double>((const double*)d_state, partial, d_out, E,
                                                              gl_weights64(), jac, stream);
    if (rc) return fail(OD_ERR_HIP, "conservation sum kernels failed to launch");
    return OD_OK;
}
