// The single-box context (include/opendynamo.h), shared by the front-end
// translation units.
#pragma once
#include <hip/hip_runtime.h>


#include "od_host.hpp"
#include "od_slots.hpp"

struct od_ctx {
    od_params p;
    int device = 0;
    hipStream_t stream = nullptr, io_stream = nullptr;
    int64_t E = 0, NP = 0;
    size_t ubytes = 0, cbytes = 0;
    void* u[2] = {nullptr, nullptr};
    void* tr[3] = {nullptr, nullptr, nullptr};
    void* fv = nullptr;
    void* dbg = nullptr;
    double* sums_partial = nullptr;
    double* sums_out = nullptr;
    od::host::HostOps ops;
    bool state_set = false;
    bool blocks = false;               // step kernels on 2x2x2-element blocks (fast f32, even dims),
                                       // which store the traces of the blocks' external faces only
    hipGraphExec_t graph = nullptr;    // one RK3 step
    hipGraphExec_t graph_k = nullptr;  // kStepsPerGraph RK3 steps
    double graph_dt = 0.0;
    void* pinned[3] = {nullptr, nullptr, nullptr};
    hipEvent_t slot_done[3] = {nullptr, nullptr, nullptr};
    // Device copies of the snapshots (allocated at the first snapshot): the
    // next step's last stage rewrites u[0] while the IO stream may still be
    // copying, so the compute stream first copies u[0] here (device to
    // device, in stream order) and the IO stream copies from here.
    void* staged[3] = {nullptr, nullptr, nullptr};
    hipEvent_t staged_done[3] = {nullptr, nullptr, nullptr};
    // Pinned staging pool (solver.rs:163-175): a slot is lent to a ticket by
    // od_get_state_async and given back by od_release.
    od::SlotPool<3> slots;
    size_t device_bytes = 0;
    // learned closure: per-element eddy viscosity [E] in the compute type
    void* mu_t = nullptr;
    bool mu_t_on = false;
    bool graph_mu_t = false;      // whether the captured graph reads mu_t
    od_mlp* mlp = nullptr;        // attached network (not owned)
    od_closure_opts closure{};    // attach options
    uint64_t steps_done = 0;      // steps since attach (update cadence)
    double ke_prev = -1.0;        // kinetic energy at the previous update
    double eps_prev = 0.0;        // model dissipation rate applied since then
    bool closure_active = false;  // false: budget fallback, plain FR until the next update
    double closure_stats[6] = {0, 0, 0, 0, 0, 0};
    void* feat = nullptr;         // [E][16] features, then eps [E] and out-of-range flags [E] (f32)
    double* red = nullptr;        // budget-check reductions (device)
    double* red_host = nullptr;   // pinned copy
};

// Closure refresh and release (od_closure.hip), used by od_step / od_destroy.
od_status od_closure_update_internal(od_ctx* c, double dt_interval);
void od_closure_release_internal(od_ctx* c);
