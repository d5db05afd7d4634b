/*
 * CPU oracle for the FR P=3 compressible Navier-Stokes RK3 step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * library is compared against; nothing in the product (opendynamo_amd/,
 * include/) may include, link or call it.  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py use it.
 *
 * It restates, op for op and in the same association order, the kernels of
 * drunkseal/opendynamo:
 *   src/fr_math.rs:1-142                 FR operators
 *   src/cuda/device_functions.cuh:1-265  pressure, Euler/viscous fluxes,
 *                                        Rusanov / AUSM+ / SLAU2
 *   src/cuda/rhs.cu:1-739                volume kernel, face x/y/wall/z
 *   src/cuda/rk.cu:1-115                 FR corrections + SSP-RK3 blend
 *   src/solver.rs:256-317                stage order and coefficients
 * with floating-point contraction disabled (built -ffp-contract=off), so
 * every product is rounded before it is added.
 *
 * Additions with no counterpart in the reference (each defined here and in
 * the HIP library identically): inviscid mode (viscous terms skipped),
 * runtime mu_ref, Sutherland viscosity, configurable body force, optional
 * RK accumulation in the state type, CFL-derived dt, conservation sums.
 */
#ifndef OD_ORACLE_H
#define OD_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int32_t nx, ny, nz;
    double lx, ly, lz;          /* f32 config values widened to f64 */
    int32_t u_real;             /* 4 = f32, 8 = f64 (state)              */
    int32_t q_real;             /* gradient round trip type              */
    int32_t rhs_real;           /* RHS round trip type                   */
    int32_t compute_real;       /* arithmetic type                       */
    int32_t flux_scheme;        /* 0 = slau2, 1 = ausm+, 2 = rusanov     */
    int32_t y_wall;             /* 1 = slip walls in y, 0 = periodic     */
    int32_t viscous;            /* 1 = Navier-Stokes, 0 = Euler          */
    int32_t viscosity_model;    /* 0 = fixed mu_ref, 1 = Sutherland      */
    double mu_ref, t_ref, s_suth;
    double body_force_x;        /* added to the rho*u RHS (reference: 0.001) */
    int32_t rk_accum;           /* 0 = blend in compute type (reference), 1 = in state type */
    int32_t n_threads;          /* 0 = all available cores */
} orc_params;

/* FR operators for P=3 in f64 (fr_math.rs:98-142) plus GL weights. */
void orc_fr_operators(double points[4], double d_matrix[16], double extrap_left[4],
                      double extrap_right[4], double g_left_prime[4],
                      double g_right_prime[4], double weights[4]);

/* Advance the state (var-major [5][E*64], type u_real) n_steps RK3 steps of
 * size dt in place.  dt is the f32 config value widened to f64. */
int orc_step(const orc_params* p, void* state, double dt, int n_steps);

/* orc_step with a per-element eddy viscosity mu_t[E] (f64, converted to the
 * compute type) added to the molecular viscosity at every viscous flux of the
 * element: mu = mu_mol + mu_t[e].  No reference counterpart. */
int orc_step_mut(const orc_params* p, void* state, const double* mu_t, double dt, int n_steps);

/* Corrected RHS of one stage evaluated on `state`: RHS + sum of the FR face
 * corrections, exactly the value the RK blend consumes (compute_real type). */
int orc_rhs_total(const orc_params* p, const void* state, void* out);

/* Intermediate arrays of one stage, for debugging:
 *   which 0: raw volume RHS   [5][N]        (rhs_real type)
 *   which 1..6: gradient slices xL,xR,yL,yR,zL,zR [E][16][12] (compute type)
 *   which 7..12: j arrays xL,xR,yL,yR,zL,zR       [5][E][16]  (compute type) */
int orc_debug_array(const orc_params* p, const void* state, int which, void* out);

/* Point physics in the compute type given by `real` (4/8), doubles in/out.
 *   kind 0: pressure(in[5]) -> out[0]
 *   kind 1..3: Euler flux x/y/z (in[5]) -> out[5]
 *   kind 4..6: viscous flux x/y/z (in[5] state, in[5..19] gradient, in[20] mu) -> out[5]
 *   kind 7..9: interface flux scheme 0/1/2 (in[0..4] left, in[5..9] right,
 *              in[10..12] normal) -> out[5]
 *   kind 10: Sutherland mu (in[5] state, in[5..8] mu_ref,t_ref,s_suth) -> out[0] */
int orc_point(int real, int kind, const double* in, double* out);

/* GL-quadrature integrals of the 5 conserved variables over the box,
 * compensated summation in f64. */
int orc_sums(const orc_params* p, const void* state, double out[5]);

/* CFL time step: cfl / ((2P+1) * max over points of sum_d (|u_d|+c)/dx_d),
 * evaluated in f64 and rounded to f32 like the config value. */
double orc_cfl_dt(const orc_params* p, const void* state, double cfl);

#ifdef __cplusplus
}
#endif
#endif
