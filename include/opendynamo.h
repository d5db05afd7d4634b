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
 *   od_step              solver.rs:262-317  (RK3 loop: launch_rhs_stage +
 *                        launch_rk_stage x3, buffer swap) -> one hipGraph per step
 *   od_synchronize       solver.rs:334, 343 (stream.synchronize)
 *   od_rhs               rhs.cu:344-739 + rk.cu:12-25 (the corrected RHS the
 *                        stage update consumes, exposed for parity checks)
 *   od_destroy           drop of the CudaSlice buffers / context
 * Entry points without a reference counterpart: od_conservation_sums,
 * od_cfl_dt, od_time_steps, od_time_kernels, od_info.
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

/* Snapshot the current state into library-owned pinned memory on a second
 * stream ordered after the work enqueued so far; od_wait(ticket) blocks until
 * the copy is complete and returns the pinned pointer (valid until the next
 * od_get_state_async that reuses the slot; 3 slots rotate). */
od_status od_get_state_async(od_ctx* ctx, uint64_t* ticket);
od_status od_wait(od_ctx* ctx, uint64_t ticket, const void** host_state);

/* Enqueue n_steps RK3 steps of size dt (the f32 config value widened).
 * Asynchronous; one hipGraph launch per step. */
od_status od_step(od_ctx* ctx, double dt, uint32_t n_steps);
od_status od_synchronize(od_ctx* ctx);

/* Corrected RHS (RHS + FR face corrections) of the current state in the
 * compute type, [5][N]; the state is not changed. */
od_status od_rhs(od_ctx* ctx, void* host_out, size_t n_values);

/* GL-quadrature integrals of the 5 conserved variables (deterministic
 * fixed-order FP64 reduction on the device). */
od_status od_conservation_sums(od_ctx* ctx, double out[5]);

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

#ifdef __cplusplus
}
#endif
#endif
