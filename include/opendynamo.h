/*
 * opendynamo.h — C ABI of the MI355X (gfx950) FR P=3 compressible
 * Navier-Stokes RK3 step.
 *
 * Plain C: opaque context, plain pointers and sizes, status codes.  The
 * library owns every device buffer, the stream and the captured step graph;
 * the caller owns host arrays.  No call aborts or throws across the ABI: a
 * non-zero status comes with a message from od_last_error().
 *
 * What each entry point replaces in drunkseal/opendynamo (file:line):
 *   od_create            solver.rs:147-178  (CudaContext::new, compile_module,
 *                        load_kernels, setup_fr_operators, mesh_dims,
 *                        allocate_buffers, launch_configs); fr_math.rs:98-142
 *   od_set_state         solver.rs:185-201  (planes -> device state u)
 *   od_get_state         solver.rs:229-230, 355-365 (device -> host copy)
 *   od_get_state_async   solver.rs:757-759  (event + io_stream D2H)
 *   od_wait              solver.rs:785      (io_stream.synchronize in the IO task)
 *   od_release           solver.rs:816      (done.send(buf): the buffer back to the pool)
 *   od_step              solver.rs:262-317  (RK3 loop: launch_rhs_stage +
 *                        launch_rk_stage x3, buffer swap) -> one hipGraph per step
 *   od_synchronize       solver.rs:334, 343 (stream.synchronize)
 *   od_rhs               rhs.cu:344-739 + rk.cu:12-25 (the corrected RHS the
 *                        stage update consumes, exposed for parity checks)
 *   od_destroy           drop of the CudaSlice buffers / context
 * Entry points without a reference counterpart: od_conservation_sums,
 * od_cfl_dt, od_time_steps, od_time_kernels, od_info, and the od_batch_*
 * ensemble API (the reference runs one box per process).
 */
#ifndef OPENDYNAMO_H
#define OPENDYNAMO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum od_status {
    OD_OK = 0,
    OD_ERR_INVALID = 1,   /* bad argument or parameter combination */
    OD_ERR_HIP = 2,       /* a HIP runtime call failed */
    OD_ERR_NOMEM = 3,     /* device or pinned host allocation failed */
    OD_ERR_STATE = 4,     /* call not valid in the context's current state */
    OD_ERR_UNSUPPORTED = 5
} od_status;

/* Real types, as the [precision] keys of config.toml (config.rs:162-175). */
typedef enum od_real { OD_F32 = 4, OD_F64 = 8 } od_real;

/* [solver] flux_scheme (config.rs:144-151). */
typedef enum od_flux { OD_FLUX_SLAU2 = 0, OD_FLUX_AUSMPLUS = 1, OD_FLUX_RUSANOV = 2 } od_flux;

typedef struct od_params {
    int32_t nx, ny, nz;        /* [mesh] element counts */
    double lx, ly, lz;         /* [mesh] extents: the f32 config values widened */
    int32_t u_real;            /* [precision] u_real: state type */
    int32_t q_real;            /* [precision] q_real: gradient round trip */
    int32_t rhs_real;          /* [precision] rhs_real: RHS round trip */
    int32_t compute_real;      /* [precision] compute_real: arithmetic type */
    int32_t flux_scheme;       /* od_flux */
    int32_t y_wall;            /* [physics] y_boundary: 1 = "wall", 0 = "periodic" */
    int32_t viscous;           /* 1 = Navier-Stokes, 0 = Euler (mu_ref = 0 selects it) */
    int32_t viscosity_model;   /* [physics] viscosity_model: 0 = fixed, 1 = sutherland */
    double mu_ref, t_ref, s_suth;
    double body_force_x;       /* [physics] body_force_x, reference value 0.001 */
    int32_t rk_accum;          /* [precision] rk_accum: 0 = compute type, 1 = state type */
    int32_t fast_math;         /* [solver] fast_math: 1 = fast-math kernels for f32 compute */
} od_params;

typedef struct od_ctx od_ctx;

/* Fill `p` with the reference defaults (config.rs:17-99). */
void od_default_params(od_params* p);

od_status od_create(const od_params* p, int device, od_ctx** out);
void od_destroy(od_ctx* ctx);

/* State planes, var-major [5][nx*ny*nz*64] (rho, rhou, rhov, rhow, E),
 * element-major 64-point blocks, pt = i + 4j + 16k (io.rs:54-58).
 * `dtype` may differ from u_real: values are converted on the host. */
od_status od_set_state(od_ctx* ctx, const void* host, size_t n_values, od_real dtype);
od_status od_get_state(od_ctx* ctx, void* host, size_t n_values, od_real dtype);

/* Loan-and-return pool of 3 library-owned pinned staging slots
 * (solver.rs:163-175, 320, 816).
 * od_get_state_async borrows a free slot -- blocking until od_release
 * returns one if all 3 are out -- and snapshots the current state into it:
 * a device-to-device copy in the compute stream's order (so steps enqueued
 * later cannot change the snapshot), then the copy to the pinned slot on a
 * second stream, overlapping later steps.  od_wait(ticket)
 * blocks until the copy is complete and returns the pinned pointer, which
 * stays valid until od_release(ticket) hands the slot back.  Every ticket
 * must be released exactly once (also after a failed write), or a later
 * od_get_state_async blocks for ever.  od_wait / od_release may be called
 * from any thread; od_get_state_async from the thread that calls od_step. */
od_status od_get_state_async(od_ctx* ctx, uint64_t* ticket);
od_status od_wait(od_ctx* ctx, uint64_t ticket, const void** host_state);
od_status od_release(od_ctx* ctx, uint64_t ticket);

/* Enqueue n_steps RK3 steps of size dt (the f32 config value widened).
 * Asynchronous; replays of a captured 8-step hipGraph plus a 1-step one
 * (one 1-step replay per step while a closure network is attached). */
od_status od_step(od_ctx* ctx, double dt, uint32_t n_steps);
od_status od_synchronize(od_ctx* ctx);

/* Corrected RHS (RHS + FR face corrections) of the current state in the
 * compute type, [5][N]; the state is not changed. */
od_status od_rhs(od_ctx* ctx, void* host_out, size_t n_values);

/* GL-quadrature integrals of the 5 conserved variables (deterministic
 * fixed-order FP64 reduction on the device). */
od_status od_conservation_sums(od_ctx* ctx, double out[5]);

/* Allocation-free variant of od_conservation_sums.  No reference
 * counterpart: it is for callers that cannot allocate device memory (code
 * under hipGraph stream capture, serving loops and harnesses that forbid
 * hipMalloc) and hold the state in their own device memory.
 *   d_state  device pointer, var-major [5][nx*ny*nz*64] of type u_real
 *   d_out    device pointer to 5 doubles: the GL integrals (same kernels,
 *            same fixed reduction order: bitwise equal to
 *            od_conservation_sums on the same state and extents)
 *   stream   a hipStream_t (NULL: the null stream); the call only enqueues
 *            two kernels on it and returns, the caller synchronizes.
 * The per-element partial sums live in a fixed device array of the library
 * (5 * OD_NOALLOC_MAX_ELEMENTS doubles = 40 MiB on each GPU the library is
 * loaded on, held until the process ends); meshes above
 * OD_NOALLOC_MAX_ELEMENTS elements are refused with OD_ERR_INVALID.  The call
 * makes no hipMalloc / hipFree.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (different streams, or threads) race. */
#define OD_NOALLOC_MAX_ELEMENTS 1048576
od_status od_conservation_sums_noalloc(int32_t nx, int32_t ny, int32_t nz, double lx, double ly, double lz,
                                       const void* d_state, od_real u_real, double* d_out, void* stream);

/* CFL time step from the current state:
 *   dt = cfl / ((2P+1) * max_points sum_d (|u_d| + c) / dx_d), rounded to f32. */
od_status od_cfl_dt(od_ctx* ctx, double cfl, double* dt_out);

/* Time n_steps graph launches with HIP events on the library stream. */
od_status od_time_steps(od_ctx* ctx, double dt, uint32_t n_steps, double* ms_total);

/* Time every kernel launch of n_steps steps individually (events around each
 * launch, no graph).  out_ms[q] = mean duration of kernel kind q; names[q]
 * its name.  Kinds: 0 = od_visc_kernel, 1 = od_update_kernel. */
od_status od_time_kernels(od_ctx* ctx, double dt, uint32_t n_steps, double out_ms[4],
                          const char* names[4], int* n_kinds);

/* Sizes and bytes for reporting: out[0] = DOF, out[1] = device bytes held,
 * out[2] = kernel launches per step. */
od_status od_info(od_ctx* ctx, double out[4]);

const char* od_last_error(void);

/* ---- Ensembles --------------------------------------------------------
 * No reference counterpart.  A batch holds independent periodic or
 * wall-bounded boxes of different sizes sharing one od_params (physics,
 * scheme, precision; its nx..lz are ignored).  Every RK3 stage advances all
 * boxes with one launch of the gradient kernel and one of the update kernel
 * per kernel form (a grid walking the tiles of every box; in fast f32
 * batches, a second pair walking the 2x2x2-element blocks
 * of every box's core -- the box less its last layer along each odd
 * dimension -- with the tiles left for those layers), or, with schedule
 * order 2, with per-box launches; either way
 * one hipGraph per step.  Each box's result is byte-identical to the same
 * box run alone through od_step with the same dt, for every schedule. */
typedef struct od_batch od_batch;

typedef struct od_box {
    int32_t nx, ny, nz;        /* element counts */
    double lx, ly, lz;         /* extents */
} od_box;

/* How the batch kernels map tiles (two neighbouring elements, 128 threads)
 * to work-groups; block boxes always take one 512-thread work-group per
 * block.  The autotuner searches these; results never depend on it. */
typedef struct od_schedule {
    int32_t orient;            /* tile = element pair along 0 x, 1 y, 2 z */
    int32_t pairs;             /* tiles per work-group: 1 (128 threads) or 2 (256) */
    int32_t order;             /* 0 round robin; 1 contiguous tile range per XCD;
                                  2 box by box: the single-box kernels launched per box
                                  and stage inside the step graph (other fields ignored) */
    int32_t groups_per_cu;     /* 0: one work-group per work unit; k > 0: a
                                  persistent grid of k (at most the resident
                                  maximum) work-groups per CU */
} od_schedule;

od_status od_batch_create(const od_params* common, const od_box* boxes, int32_t n_boxes, int device,
                          od_batch** out);
void od_batch_destroy(od_batch* b);

/* State of one box, layout as od_set_state with that box's dimensions. */
od_status od_batch_set_state(od_batch* b, int32_t box, const void* host, size_t n_values, od_real dtype);
od_status od_batch_get_state(od_batch* b, int32_t box, void* host, size_t n_values, od_real dtype);

/* n_steps RK3 steps of every box; dt[box] is that box's step. Asynchronous. */
od_status od_batch_step(od_batch* b, const double* dt, uint32_t n_steps);
od_status od_batch_synchronize(od_batch* b);

/* Per-box CFL step (as od_cfl_dt) -> dt_out[n_boxes];
 * per-box GL integrals -> out[n_boxes][5]. */
od_status od_batch_cfl_dt(od_batch* b, double cfl, double* dt_out);
od_status od_batch_conservation_sums(od_batch* b, double* out);

od_status od_batch_set_schedule(od_batch* b, const od_schedule* s);
od_status od_batch_get_schedule(od_batch* b, od_schedule* s);

/* Time every candidate schedule over steps_per_trial steps (after one
 * untimed step) from the current state, restore the state after each trial,
 * and keep the fastest.  Up to `cap` trials are reported in trials[] /
 * trial_ms[] (ms per step); *n_trials = number of candidates timed. */
od_status od_batch_autotune(od_batch* b, const double* dt, uint32_t steps_per_trial, od_schedule* best,
                            double* best_ms, od_schedule* trials, double* trial_ms, int32_t cap,
                            int32_t* n_trials);

/* Time n_steps graph launches (ms total); per_box != 0 times the same steps
 * launched box by box with the single-box kernels instead (6 launches per
 * box per step, 3 without viscosity, no graph), for comparison. */
od_status od_batch_time_steps(od_batch* b, const double* dt, uint32_t n_steps, int32_t per_box, double* ms_total);

/* out[0] = total DOF, out[1] = device bytes held, out[2] = kernel launches
 * per step of the batched schedules (2 or 1 per stage and kernel form, the
 * forms being tiles and blocks), out[3] = number of boxes. */
od_status od_batch_info(od_batch* b, double out[4]);

/* ---- Learned closure ---------------------------------------------------
 * No reference counterpart.  An element-local MLP maps 12 invariants of the
 * element's resolved field (features below) through 3 hidden layers of 64
 * GELU units to a coefficient C = softplus(.) >= 0; the element's eddy
 * viscosity mu_t = rho_e * Delta^2 * |S|_e * C is added to the molecular
 * viscosity of every viscous flux of the element (mu = mu_mol + mu_t[e]);
 * C = s * softplus(.) with the output scale s of od_mlp_set_output_scale.
 * Weights and activations are BF16 on MFMA with FP32 accumulation.
 * Features (element means with GL weights, Delta = element size / (P+1),
 * |S| = sqrt(2 S^d:S^d)):
 *   0 log(rho Delta^2 |S| / mu_mol)   1 |Omega| / |S|     2 Q / |S|^2
 *   3 R / |S|^3                        4 div u / |S|       5..7 fraction of the
 *   velocity fluctuation energy in Legendre modes of highest degree 1, 2, 3
 *   8 rms(rho) / rho   9 rms(p) / p   10 Delta |S| / c   11 rms(u) / (Delta |S|)
 * The feature kernel also returns rho_e, Delta^2 |S|_e, the element's
 * kinetic energy and its integral of 2 S^d:S^d (16 floats per element). */
#define OD_MLP_FEATURES 12
#define OD_MLP_HIDDEN 64
#define OD_FEATURE_STRIDE 16
/* Parameter vector (f32): W1[64][12] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4 */
#define OD_MLP_PARAMS 9217

typedef struct od_mlp od_mlp;

typedef struct od_closure_opts {
    int32_t update_interval;   /* steps between eddy-viscosity updates (>= 1) */
    double budget_kappa;       /* kinetic-energy budget check: when, over an interval in
                                  which the kinetic energy decayed, the energy the model
                                  dissipated exceeds kappa times the observed decay, the next
                                  interval runs the plain FR step; <= 0 disables it */
} od_closure_opts;

od_status od_mlp_create(int device, od_mlp** out);
void od_mlp_destroy(od_mlp* m);
od_status od_mlp_set_params(od_mlp* m, const float* params, size_t n);
od_status od_mlp_get_params(od_mlp* m, float* params, size_t n);
/* Input standardisation x' = (x - shift) * scale and the training range
 * [lo, hi] of each raw feature: an element with a feature outside it gets
 * mu_t = 0 (the plain FR step). */
od_status od_mlp_set_scaling(od_mlp* m, const float shift[12], const float scale[12], const float lo[12],
                             const float hi[12]);
/* Output scale s: C = s * softplus(z) (default 1), so the network trains on
 * targets of order one. */
od_status od_mlp_set_output_scale(od_mlp* m, float s);
/* Forward pass on raw features x[n][12] -> C[n] (no range check). */
od_status od_mlp_forward(od_mlp* m, const float* x, size_t n, float* y);
/* Mean squared error of C against t[n] and its gradient (OD_MLP_PARAMS). */
od_status od_mlp_loss_grad(od_mlp* m, const float* x, const float* t, size_t n, double* loss, float* grad);
/* Adam (beta 0.9 / 0.999) on the device over `epochs` passes of shuffled
 * mini-batches of (x, t); loss_per_epoch[epochs] may be NULL.  Deterministic
 * for a given seed. */
od_status od_mlp_train(od_mlp* m, const float* x, const float* t, size_t n, uint32_t epochs, uint32_t batch,
                       float lr, uint64_t seed, float* loss_per_epoch);

/* Features of every element of the current state: out[E][16]. */
od_status od_closure_features(od_ctx* ctx, float* out, size_t n);
/* Attach (m != NULL) or detach the network: od_step then refreshes mu_t
 * from the state every update_interval steps (features -> MLP -> range
 * check -> budget check).  The network must outlive the attachment. */
od_status od_closure_attach(od_ctx* ctx, od_mlp* m, const od_closure_opts* opts);
/* Set a per-element eddy viscosity host[E] directly, or switch it off
 * (host == NULL: the plain FR step). */
od_status od_set_eddy_viscosity(od_ctx* ctx, const void* host, size_t n, od_real dtype);
od_status od_get_eddy_viscosity(od_ctx* ctx, void* host, size_t n, od_real dtype);
/* out[0] updates, out[1] updates that fell back on the budget check,
 * out[2] element-updates outside the training range, out[3] kinetic energy
 * at the last update, out[4] model dissipation rate at the last update,
 * out[5] 1 if the model is active for the current interval. */
od_status od_closure_stats(od_ctx* ctx, double out[6]);

#ifdef __cplusplus
}
#endif
#endif
