// gfx950 kernels of the FR P=3 compressible Navier-Stokes RK3 step.
//
// Per RK stage (drunkseal/opendynamo: solver.rs:610-738, rhs.cu, rk.cu):
//   od_visc_kernel    one wave64 per element: BR1-corrected gradient from the
//                     element's LDS tile and its neighbours' stored face
//                     traces, gradient face extrapolation, viscous normal flux
//                     at the element's 6 faces -> fv  (replaces the gradient
//                     half of compute_rhs_volume_kernel and the viscous half of
//                     the face kernels)
//                     In the fast f32 build it also forms the divergence of
//                     the volume viscous fluxes (-> fv, after the face part)
//   od_update_kernel  one wave64 per element: gradient again (register only;
//                     not in the fast f32 build, which takes the volume
//                     viscous divergence from the visc kernel), volume flux
//                     divergence, interface fluxes at all 6 own
//                     faces (SLAU2 / AUSM+ / Rusanov, slip wall), FR
//                     correction, SSP-RK3 blend, face traces of the new state
//                     (replaces the rest of the volume kernel, the x/y/wall/z
//                     face kernels, the j arrays and rk_stage*_kernel)
// Each face flux is evaluated by both elements that share the face, from the
// same operands in the same order, so both sides see identical bits and no
// face array, atomic or second pass is needed.
//
// Lane = solution point pt = i + 4j + 16k of the element.  Line contractions
// stage through LDS.  The arithmetic of every quantity (association order,
// no contraction: built with -ffp-contract=off) is that of the reference
// kernel that computes it; comments cite the reference line.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "od_internal.hpp"

// OD_FAST = 1 builds the fast-math variant of the same kernels (selected by
// the [solver] fast_math key for f32 compute): divisions become reciprocal
// multiplies, square roots the hardware approximation, normal-vector zero
// terms are dropped, and the translation unit is compiled with contraction.
#ifndef OD_FAST
#define OD_FAST 0
#endif

namespace od {
#if OD_FAST
namespace fastmath {
#endif

// FR operators in constant memory, set once per device by upload_ops (the
// same P = 3 values for every solver): kept out of the kernel arguments so
// they do not occupy scalar registers for the whole kernel.
__constant__ Ops<float> g_ops_f32;
__constant__ Ops<double> g_ops_f64;

namespace {

template <class C>
__device__ __forceinline__ const Ops<C>& dev_ops() {
    if constexpr (sizeof(C) == 4)
        return g_ops_f32;
    else
        return g_ops_f64;
}

// a / b and sqrt(x): IEEE (correctly rounded) in the strict build; in the
// fast build for f32, v_rcp_f32 / v_sqrt_f32 (1 ulp).
template <class C>
__device__ __forceinline__ C dv(C a, C b) {
#if OD_FAST
    if constexpr (sizeof(C) == 4) return a * __builtin_amdgcn_rcpf(b);
#endif
    return a / b;
}
template <class C>
__device__ __forceinline__ C sq(C x) {
#if OD_FAST
    if constexpr (sizeof(C) == 4) return __builtin_amdgcn_sqrtf(x);
#endif
    return sqrt(x);
}

// Fast build: the arithmetic of the point and interface fluxes may fuse a
// multiply into an add only within one source expression (as the C standard
// allows), not across statements.  Fusion across statements is decided per
// kernel from the shape of the whole computation (for a*b + c*d, which
// product to fuse depends on how many other expressions use each), so the
// same flux could round differently in the two-element-tile and the block
// kernels; within an expression the choice is fixed by the source.
#if OD_FAST
#define OD_CONTRACT_IN_EXPRESSIONS _Pragma("clang fp contract(on)")
#else
#define OD_CONTRACT_IN_EXPRESSIONS
#endif

template <class T>
__device__ __forceinline__ T sel3(T a0, T a1, T a2, int d) {
    return d == 0 ? a0 : (d == 1 ? a1 : a2);
}

// The RK stage blend kA u^n + kB u + kC r (stage 1: kA u^n + kC r; rk.cu:12-25).
// The fast build writes the fusion out, fma(kC, r, fma(kB, u, kA u^n)),
// so that every kernel form rounds it alike; the strict build rounds each
// product and sum, left to right.
template <class T>
__device__ __forceinline__ T fma_t(T a, T b, T c) {
    if constexpr (sizeof(T) == 4)
        return __builtin_fmaf(a, b, c);
    else
        return __builtin_fma(a, b, c);
}
template <class T>
__device__ __forceinline__ T rk_blend1(T kA, T kC, T un, T r) {
#if OD_FAST
    return fma_t(kC, r, kA * un);
#else
    return kA * un + kC * r;
#endif
}
template <class T>
__device__ __forceinline__ T rk_blend(T kA, T kB, T kC, T un, T u, T r) {
#if OD_FAST
    return fma_t(kC, r, fma_t(kB, u, kA * un));
#else
    return kA * un + kB * u + kC * r;
#endif
}

// ---------------------------------------------------------------------------
// Point physics in the compute type C (device_functions.cuh:1-257).
// Direction d = 0/1/2 is an argument; all three reference variants
// (compute_euler_flux_{x,y,z}, compute_viscous_flux_{x,y,z}) are the same
// expression tree with the direction's operands, which is what the selects
// pick, so the result is bitwise that of the matching variant.
// ---------------------------------------------------------------------------
template <class C>
__device__ __forceinline__ C gamma_m1() { return ((C)1.4) - ((C)1.0); }

template <class C>
__device__ __forceinline__ C pressure(C rho, C ru, C rv, C rw, C e) {  // :1-5
    OD_CONTRACT_IN_EXPRESSIONS
#if OD_FAST
    if constexpr (sizeof(C) == 4) {  // fast build: the fusions written out
        const C h = ((C)0.5) * __builtin_fmaf(rw, rw, __builtin_fmaf(rv, rv, ru * ru));
        const C p = gamma_m1<C>() * __builtin_fmaf(-h, __builtin_amdgcn_rcpf(rho), e);
        return fmax(((C)1e-7), p);
    }
#endif
    C ke = dv(((C)0.5) * (ru * ru + rv * rv + rw * rw), rho);
    C p = gamma_m1<C>() * (e - ke);
    return fmax(((C)1e-7), p);
}

template <class C>
__device__ __forceinline__ void euler_flux(int d, const C* s, C* f) {  // :7-21
    OD_CONTRACT_IN_EXPRESSIONS
    C rho = s[0], e = s[4];
    C p = pressure(rho, s[1], s[2], s[3], e);
    C md = sel3(s[1], s[2], s[3], d);
    C ud = dv(md, rho);
    f[0] = md;
#pragma unroll
    for (int c = 0; c < 3; ++c) f[1 + c] = (c == d) ? s[1 + c] * ud + p : s[1 + c] * ud;
    f[4] = ud * (e + p);
}

// :23-102.  g[3*var + dir]; G[a][b] = d(u_a)/d(x_b).
// The direction-independent part of the three directional viscous fluxes of
// a point: velocity, stress tensor and the heat flux along each direction.
// Every entry is the expression the directional flux evaluates, so the
// fluxes assembled from it are bitwise those of compute_viscous_flux_{x,y,z};
// the volume pass forms it once for its three directions.
template <class C>
struct ViscPoint {
    C u[3];
    C t00, t11, t22, t01, t02, t12;
    C q[3];
};
template <class C>
__device__ __forceinline__ C heat_flux(int d, const C* s, const C* g, C mu, C rho, C ke, C p) {
    C gd = sel3(g[0], g[1], g[2], d);
    C gu = sel3(g[3], g[4], g[5], d);
    C gv = sel3(g[6], g[7], g[8], d);
    C gw = sel3(g[9], g[10], g[11], d);
    C ge = sel3(g[12], g[13], g[14], d);
    C dke = dv(s[1] * gu + s[2] * gv + s[3] * gw, rho) - dv(ke * gd, rho);
    C dp = gamma_m1<C>() * (ge - dke);
    C coeff = dv(mu * ((C)1.4), ((C)0.72) * gamma_m1<C>());
    return -coeff * (dv(dp, rho) - dv(p * gd, rho * rho));
}
// dirs: bit d set = the heat flux along d is needed.
template <class C>
__device__ __forceinline__ ViscPoint<C> viscous_point(const C* s, const C* g, C mu, int dirs) {
    ViscPoint<C> vp;
    C rho = s[0];
    vp.u[0] = dv(s[1], rho);
    vp.u[1] = dv(s[2], rho);
    vp.u[2] = dv(s[3], rho);
    C G[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) G[a][b] = dv(g[3 * (a + 1) + b] - vp.u[a] * g[b], rho);
    C div = G[0][0] + G[1][1] + G[2][2];
    const C two = (C)2.0, tt = ((C)2.0) / ((C)3.0);
    vp.t00 = mu * (two * G[0][0] - tt * div);
    vp.t11 = mu * (two * G[1][1] - tt * div);
    vp.t22 = mu * (two * G[2][2] - tt * div);
    vp.t01 = mu * (G[0][1] + G[1][0]);
    vp.t02 = mu * (G[0][2] + G[2][0]);
    vp.t12 = mu * (G[1][2] + G[2][1]);
    C ke = dv(((C)0.5) * (s[1] * s[1] + s[2] * s[2] + s[3] * s[3]), rho);
    C p = gamma_m1<C>() * (s[4] - ke);
#pragma unroll
    for (int d = 0; d < 3; ++d) vp.q[d] = (dirs >> d) & 1 ? heat_flux(d, s, g, mu, rho, ke, p) : (C)0;
    return vp;
}
template <class C>
__device__ __forceinline__ void viscous_flux_dir(int d, const ViscPoint<C>& vp, C* f) {
    C r0 = sel3(vp.t00, vp.t01, vp.t02, d), r1 = sel3(vp.t01, vp.t11, vp.t12, d), r2 = sel3(vp.t02, vp.t12, vp.t22, d);
    f[0] = (C)0.0;
    f[1] = r0;
    f[2] = r1;
    f[3] = r2;
    f[4] = vp.u[0] * r0 + vp.u[1] * r1 + vp.u[2] * r2 - vp.q[d];
}
template <class C>
__device__ __forceinline__ void viscous_flux(int d, const C* s, const C* g, C mu, C* f) {
    viscous_flux_dir(d, viscous_point(s, g, mu, 1 << d), f);
}

template <class C>
__device__ __forceinline__ C sound_speed(C p, C rho) {
    OD_CONTRACT_IN_EXPRESSIONS
#if OD_FAST
    if constexpr (sizeof(C) == 4) {
        // 1 / max(1e-7, rho) as a select: the reciprocal of rho is the one
        // the fluxes form anyway (one transcendental fewer per state)
        const C inv = rho >= (C)1e-7 ? __builtin_amdgcn_rcpf(rho) : __builtin_amdgcn_rcpf((C)1e-7);
        return __builtin_amdgcn_sqrtf(fmax(((C)0.0), (((C)1.4) * p) * inv));
    }
#endif
    return sq(fmax(((C)0.0), dv(((C)1.4) * p, fmax(((C)1e-7), rho))));
}

// Normal velocity u . n and momentum flux component m v_n + p n_c for the
// face normal n = e_D.  The strict build keeps the reference's arithmetic
// with the zero components (u*1 + v*0 + w*0, p*0), which is what fixes the
// sign of zero results; the fast build drops the zero terms.
template <int D, class C>
__device__ __forceinline__ C nvel(C u, C v, C w) {
    OD_CONTRACT_IN_EXPRESSIONS
#if OD_FAST
    return D == 0 ? u : (D == 1 ? v : w);
#else
    const C n0 = D == 0 ? (C)1.0 : (C)0.0, n1 = D == 1 ? (C)1.0 : (C)0.0, n2 = D == 2 ? (C)1.0 : (C)0.0;
    return u * n0 + v * n1 + w * n2;
#endif
}
template <int D, class C>
__device__ __forceinline__ C mflux(int c, C a, C b, C p) {
    OD_CONTRACT_IN_EXPRESSIONS
#if OD_FAST
    return c == D ? a * b + p : a * b;
#else
    return a * b + p * (c == D ? (C)1.0 : (C)0.0);
#endif
}

template <int D, class C>
__device__ __forceinline__ void rusanov(const C* l, const C* r, C* f) {  // :104-128
    OD_CONTRACT_IN_EXPRESSIONS
    C ul = dv(l[1], l[0]), vl = dv(l[2], l[0]), wl = dv(l[3], l[0]);
    C ur = dv(r[1], r[0]), vr = dv(r[2], r[0]), wr = dv(r[3], r[0]);
    C pl = pressure(l[0], l[1], l[2], l[3], l[4]);
    C pr = pressure(r[0], r[1], r[2], r[3], r[4]);
    C vnl = nvel<D>(ul, vl, wl);
    C vnr = nvel<D>(ur, vr, wr);
    C cl = sound_speed(pl, l[0]), cr = sound_speed(pr, r[0]);
    C lam = fmax(fabs(vnl) + cl, fabs(vnr) + cr);
    C fl[5], fr[5];
    fl[0] = l[0] * vnl;
    fr[0] = r[0] * vnr;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        fl[1 + c] = mflux<D>(c, l[1 + c], vnl, pl);
        fr[1 + c] = mflux<D>(c, r[1 + c], vnr, pr);
    }
    fl[4] = vnl * (l[4] + pl);
    fr[4] = vnr * (r[4] + pr);
#pragma unroll
    for (int v = 0; v < 5; ++v) f[v] = ((C)0.5) * (fl[v] + fr[v] - lam * (r[v] - l[v]));
}

template <int D, class C>
__device__ __forceinline__ void ausm_plus(const C* l, const C* r, C* f) {  // :130-188
    OD_CONTRACT_IN_EXPRESSIONS
    C ul = dv(l[1], l[0]), vl = dv(l[2], l[0]), wl = dv(l[3], l[0]);
    C ur = dv(r[1], r[0]), vr = dv(r[2], r[0]), wr = dv(r[3], r[0]);
    C pl = pressure(l[0], l[1], l[2], l[3], l[4]);
    C pr = pressure(r[0], r[1], r[2], r[3], r[4]);
    C vnl = nvel<D>(ul, vl, wl);
    C vnr = nvel<D>(ur, vr, wr);
    C hl = dv(l[4] + pl, l[0]), hr = dv(r[4] + pr, r[0]);
    C cl = sound_speed(pl, l[0]), cr = sound_speed(pr, r[0]);
    C cm = fmax(cl, cr);
    C ml = dv(vnl, cm), mr = dv(vnr, cm);
    const C one = (C)1.0, q = (C)0.25, alpha = (C)0.1875, beta = (C)0.125, zero = (C)0.0;
    C mp, mm, pp, pm;
    if (fabs(ml) >= one) {
        mp = ((C)0.5) * (ml + fabs(ml));
        pp = (ml > zero) ? one : zero;
    } else {
        mp = q * (ml + one) * (ml + one) + beta * (ml * ml - one) * (ml * ml - one);
        pp = q * (ml + one) * (ml + one) * (((C)2.0) - ml) + alpha * ml * (ml * ml - one) * (ml * ml - one);
    }
    if (fabs(mr) >= one) {
        mm = ((C)0.5) * (mr - fabs(mr));
        pm = (mr < zero) ? one : zero;
    } else {
        mm = -q * (mr - one) * (mr - one) - beta * (mr * mr - one) * (mr * mr - one);
        pm = q * (mr - one) * (mr - one) * (((C)2.0) + mr) - alpha * mr * (mr * mr - one) * (mr * mr - one);
    }
    C mach = mp + mm;
    C p = pp * pl + pm * pr;
    bool pos = mach > zero;
    C mf = mach * cm * (pos ? l[0] : r[0]);
    C vs[3] = {pos ? ul : ur, pos ? vl : vr, pos ? wl : wr};
    C hs = pos ? hl : hr;
    f[0] = mf;
#pragma unroll
    for (int c = 0; c < 3; ++c) f[1 + c] = mflux<D>(c, mf, vs[c], p);
    f[4] = mf * hs;
}

template <int D, class C>
__device__ __forceinline__ void slau2(const C* l, const C* r, C* f) {  // :190-257
    OD_CONTRACT_IN_EXPRESSIONS
    C ul = dv(l[1], l[0]), vl = dv(l[2], l[0]), wl = dv(l[3], l[0]);
    C ur = dv(r[1], r[0]), vr = dv(r[2], r[0]), wr = dv(r[3], r[0]);
    C pl = pressure(l[0], l[1], l[2], l[3], l[4]);
    C pr = pressure(r[0], r[1], r[2], r[3], r[4]);
    C vnl = nvel<D>(ul, vl, wl);
    C vnr = nvel<D>(ur, vr, wr);
    C hl = dv(l[4] + pl, l[0]), hr = dv(r[4] + pr, r[0]);
    C sql = ul * ul + vl * vl + wl * wl;
    C sqr = ur * ur + vr * vr + wr * wr;
    C cl = sound_speed(pl, l[0]), cr = sound_speed(pr, r[0]);
    const C half = (C)0.5, one = (C)1.0, zero = (C)0.0;
    C cf = half * (cl + cr);
    C ml = dv(vnl, cf), mr = dv(vnr, cf);
    C mt = fmin(one, dv(one, cf) * sq(half * (sql + sqr)));
    C chi = one - mt;
    chi = chi * chi;
    C g = -fmax(fmin(ml, zero), -one) * fmin(fmax(mr, zero), one);
    C vavg = dv(l[0] * fabs(vnl) + r[0] * fabs(vnr), l[0] + r[0]);
    C val = (one - g) * vavg + g * fabs(vnl);
    C var = (one - g) * vavg + g * fabs(vnr);
    C mf = half * (l[0] * (vnl + val) + r[0] * (vnr - var) - dv(chi, cf) * (pr - pl));
    C bl = (fabs(ml) < one) ? ((C)0.25) * (((C)2.0) - ml) * (ml + one) * (ml + one)
                            : ((ml >= zero) ? one : zero);
    C br = (fabs(mr) < one) ? ((C)0.25) * (((C)2.0) + mr) * (mr - one) * (mr - one)
                            : ((mr >= zero) ? zero : one);
    C pf = half * (pl + pr) + half * (bl - br) * (pl - pr);
    pf += sq(half * (sql + sqr)) * (bl + br - one) * cf * half * (l[0] + r[0]);
    bool pos = mf > zero;
    C vs[3] = {pos ? ul : ur, pos ? vl : vr, pos ? wl : wr};
    C hs = pos ? hl : hr;
    f[0] = mf;
#pragma unroll
    for (int c = 0; c < 3; ++c) f[1 + c] = mflux<D>(c, mf, vs[c], pf);
    f[4] = mf * hs;
}

template <int D, class C>
__device__ __forceinline__ void interface_flux(int scheme, const C* l, const C* r, C* f) {
    if (scheme == 0)
        slau2<D>(l, r, f);
    else if (scheme == 1)
        ausm_plus<D>(l, r, f);
    else
        rusanov<D>(l, r, f);
}

// Sutherland law, theta = gamma p / rho (no reference counterpart; the
// reference parses the keys but uses a constant mu, constants.cuh:42).
template <class C>
__device__ __forceinline__ C sutherland(const C* s, C mu_ref, C t_ref, C s_suth) {
    C theta = dv(((C)1.4) * pressure(s[0], s[1], s[2], s[3], s[4]), s[0]);
    return dv((mu_ref * (theta * sq(theta))) * (t_ref + s_suth), t_ref * theta + s_suth);
}

// ---------------------------------------------------------------------------
// Work-group layout: two elements per 128-thread work-group, wave w owns
// element 2*blockIdx.x + w for the point-wise phases (lane = solution point
// pt = i + 4j + 16k).  Face-point phases work on both elements at once,
// direction by direction: a direction has 2 elements x 2 faces x 16 points
// = 64 tasks, one full wave pass with direction-specialised code.  Wave 0
// takes x and z, wave 1 takes y (3 full passes per 2 elements; a wave that
// finishes early waits at the barrier and frees its SIMD for other waves).
// ---------------------------------------------------------------------------
constexpr int kEpb = 2;

struct Elem {
    int32_t xl, xr, yl, yr, zl, zr;  // neighbours across each face; -1 = slip wall (box-local, < 2^31)
};

// Element (ex, ey, ez) and its face neighbours; coordinates come from the
// launch grid, so no integer division is needed.
__device__ __forceinline__ Elem make_elem(int ex, int ey, int ez, int nx, int ny, int nz, int wall) {
    Elem r;
    const int sy = nx, sz = nx * ny;
    const int e = ex + sy * ey + sz * ez;
    r.xl = e + (ex == 0 ? nx - 1 : -1);
    r.xr = e + (ex == nx - 1 ? 1 - nx : 1);
    const int yl = e + (ey == 0 ? (ny - 1) * sy : -sy);
    const int yr = e + (ey == ny - 1 ? -(ny - 1) * sy : sy);
    r.yl = (wall && ey == 0) ? -1 : yl;
    r.yr = (wall && ey == ny - 1) ? -1 : yr;
    r.zl = e + (ez == 0 ? (nz - 1) * sz : -sz);
    r.zr = e + (ez == nz - 1 ? -(nz - 1) * sz : sz);
    return r;
}

// Points of the line through face point fp of a face normal to d:
// x: fp = j+4k, y: fp = i+4k, z: fp = i+4j.
__device__ __forceinline__ void face_line(int d, int fp, int& base, int& stride) {
    base = d == 0 ? 4 * fp : (d == 1 ? (fp & 3) + 16 * (fp >> 2) : fp);
    stride = d == 0 ? 1 : (d == 1 ? 4 : 16);
}

// Trace orders of the reference: 0 = pairwise (t0+t1)+(t2+t3) (shuffle
// butterfly, rhs.cu:68-85, 595-599, 665-674), 1 = sequential from 0
// (rhs.cu:89-101, 171-176, 716-727), 2 = sequential from the first product
// (rhs.cu:131-138, 540-543).
template <class C>
__device__ __forceinline__ C trace4(const C* q, int base, int stride, C c0, C c1, C c2, C c3, int order) {
#if OD_FAST
    // Fast build: explicit fused forms, so that a trace comes out bitwise the
    // same in every kernel that forms it (the trace kernel and the update
    // kernel write the traces the next stage reads) whatever the compiler
    // would contract in each context.
    if constexpr (sizeof(C) == 4) {
        const C q0 = q[base], q1 = q[base + stride], q2 = q[base + 2 * stride], q3 = q[base + 3 * stride];
        if (order == 0) return __builtin_fmaf(q1, c1, q0 * c0) + __builtin_fmaf(q3, c3, q2 * c2);
        return __builtin_fmaf(q3, c3, __builtin_fmaf(q2, c2, __builtin_fmaf(q1, c1, q0 * c0)));
    }
#endif
    C t0 = q[base] * c0, t1 = q[base + stride] * c1, t2 = q[base + 2 * stride] * c2, t3 = q[base + 3 * stride] * c3;
    if (order == 0) return (t0 + t1) + (t2 + t3);
    if (order == 1) {
        C s0 = (C)0;
        s0 += t0;
        s0 += t1;
        s0 += t2;
        s0 += t3;
        return s0;
    }
    return t0 + t1 + t2 + t3;
}

// Bank-conflict-free position of point p = i + 4j + 16k in a 64-value LDS
// row (the visc kernel's gradient tile is stored in this order): a face pass
// reads, per half-wave, the 16 face lines of one element at one line
// position (fixed i, j or k), which in the plain layout hit 8 of the 32
// banks twice for x and y faces.  Bits of the position: i0^j0, i0^k1,
// i1, j1, k0, i0 -- a bijection under which each of those 16-point sets, and
// the 32 points a half-wave writes, fall in 16 or 32 distinct banks.
__device__ __forceinline__ int swz64(int p) {
    const int i0 = p & 1, j0 = (p >> 2) & 1, k1 = (p >> 5) & 1;
    return (i0 ^ j0) | ((i0 ^ k1) << 1) | (p & 2) << 1 | (p & 8) | (p & 16) | (i0 << 5);
}
// trace4 over a swizzled row: line points base + m * stride.
template <class C>
__device__ __forceinline__ C trace4_swz(const C* q, int base, int stride, C c0, C c1, C c2, C c3, int order) {
    const C r[4] = {q[swz64(base)], q[swz64(base + stride)], q[swz64(base + 2 * stride)], q[swz64(base + 3 * stride)]};
    return trace4(r, 0, 1, c0, c1, c2, c3, order);
}

// The fast build forms every trace in the pairwise (fused) form and keeps a
// single x trace per face (faces 0/1) for BR1 and the Riemann flux alike;
// the strict build keeps the reference's orders below, with two x traces
// (0/1 pairwise for BR1, 6/7 sequential for the interface flux).  Both
// sides of a face see the same two traces either way, so the interface
// flux stays single-valued (conservative).
constexpr bool kOneX = OD_FAST;
// Fast f32 build: the visc kernel, which holds the corrected gradient anyway,
// also forms the divergence of the volume viscous fluxes and stores it after
// the element's face viscous fluxes (fv + kFvFaces); the update kernel then
// takes the divergence of the Euler fluxes only and needs no gradient of its
// own.  The strict build keeps the reference's single divergence of
// F_e - F_v (rhs.cu:229-266) and its rounding.
template <class CT>
constexpr bool kSplitVisc = OD_FAST && std::is_same<CT, float>::value;
// The x line through point (i, j, k) is the four consecutive lanes of the
// point's quad, so the fast f32 build contracts it from registers with DPP
// quad broadcasts instead of an LDS round trip: the same products summed in
// the same order, so the results are those of the LDS form bit for bit.
template <class CT>
constexpr bool kDppX = OD_FAST && std::is_same<CT, float>::value;
template <class CT>
__device__ __forceinline__ CT quad_line(const CT* Dd, CT f) {
    static_assert(std::is_same<CT, float>::value, "DPP lanes are 32-bit");
    const int x = __float_as_int(f);
    const CT b0 = __int_as_float(__builtin_amdgcn_mov_dpp(x, 0x00, 0xF, 0xF, false));  // quad_perm [0,0,0,0]
    const CT b1 = __int_as_float(__builtin_amdgcn_mov_dpp(x, 0x55, 0xF, 0xF, false));  // [1,1,1,1]
    const CT b2 = __int_as_float(__builtin_amdgcn_mov_dpp(x, 0xAA, 0xF, 0xF, false));  // [2,2,2,2]
    const CT b3 = __int_as_float(__builtin_amdgcn_mov_dpp(x, 0xFF, 0xF, 0xF, false));  // [3,3,3,3]
    CT dd = 0;
    dd += Dd[0] * b0;
    dd += Dd[1] * b1;
    dd += Dd[2] * b2;
    dd += Dd[3] * b3;
    return dd;
}
// Fast f32 build: the interface correction takes the viscous fluxes of the
// two sides as their half-difference (F_vL - F_vR) / 2, which the block visc
// kernel stores once per interior face instead of the two fluxes; the strict
// build keeps the reference's F* = F_riemann - (F_vL + F_vR) / 2.
template <class CT>
constexpr bool kHalfDiff = OD_FAST && std::is_same<CT, float>::value;
// Order of the BR1 traces (rhs.cu:385-455): x, y pairwise; z sequential from 0.
__device__ __forceinline__ constexpr int br1_order(int d) { return (d == 2 && !kOneX) ? 1 : 0; }
// Order of the Riemann inputs and gradient slices (rhs.cu:126-180, 540-543,
// 663-674, 716-727): x sequential from the first product, y pairwise, z from 0.
__device__ __forceinline__ constexpr int face_order(int d) { return kOneX ? 0 : (d == 0 ? 2 : (d == 1 ? 0 : 1)); }
// Order of the state and gradient traces the viscous face fluxes take: the
// reference's face order; the fast build, whose viscous traces are formed and
// used inside one kernel, takes the fused sequential form (one instruction
// fewer per trace than the pairwise one).
__device__ __forceinline__ constexpr int fv_order(int d) { return kOneX ? 1 : face_order(d); }
// Faces of the trace layout the build writes, and the direction and order of face f.
constexpr int kOwnFaces = kOneX ? 6 : 8;
__device__ __forceinline__ int trace_face_dir(int f) { return f < 6 ? f >> 1 : 0; }
__device__ __forceinline__ int trace_face_order(int f) { return kOneX ? 0 : (f < 4 ? 0 : (f < 6 ? 1 : 2)); }

template <class C>
struct Coef4 {
    C c0, c1, c2, c3;
};

// The FR operators in LDS, laid out so that a lane reads the row or the side
// it needs directly: selecting among 4 wave-uniform scalars costs 3
// v_cndmask and 2 v_mov per coefficient and lane (one scalar operand per
// VALU instruction).
template <class C>
struct alignas(16) CoefTable {
    C D[4][4];  // D[i][m] = l_m'(x_i)
    C g[4][2];  // {g_left'(x_i), g_right'(x_i)}
    C e[2][4];  // {l_m(-1)}, {l_m(+1)}
};
template <class C>
__device__ __forceinline__ CoefTable<C>& coef_lds() {
    __shared__ CoefTable<C> t;
    return t;
}
// Threads 0..31 of the work-group copy the table from constant memory: the
// load is issued first thing (coef_load), the LDS store comes just before the
// body's first barrier (coef_store), which orders it before every read, so
// the load's latency hides behind the body's own loads.
template <class C>
__device__ __forceinline__ C coef_load(int tid) {
    static_assert(sizeof(Ops<C>) == 32 * sizeof(C) && sizeof(CoefTable<C>) == 32 * sizeof(C), "table layout");
    // Ops order: D[16], eL[4], eR[4], gL[4], gR[4]; the index stays inside
    // Ops for every thread, whether or not the load is speculated
    const int t = tid & 31;
    const int src = t < 16 ? t : (t < 24 ? 24 + ((t - 16) & 1) * 4 + ((t - 16) >> 1) : t - 8);
    return tid < 32 ? reinterpret_cast<const C*>(&dev_ops<C>())[src] : (C)0;
}
template <class C>
__device__ __forceinline__ void coef_store(int tid, C v) {
    if (tid < 32) reinterpret_cast<C*>(&coef_lds<C>())[tid] = v;
}
template <class C>
__device__ __forceinline__ Coef4<C> side_coef(bool right) {
    const C* r = coef_lds<C>().e[right];
    return {r[0], r[1], r[2], r[3]};
}

// Task of a direction-D face pass: lane -> (local element, side, face point).
struct FaceTask {
    int le, side, fp, f;
};
template <int D>
__device__ __forceinline__ FaceTask face_task(int lane) {
    FaceTask t;
    t.le = lane >> 5;
    t.side = (lane >> 4) & 1;
    t.fp = lane & 15;
    t.f = 2 * D + t.side;
    return t;
}

// The work-group's two elements, held in named scalars (no private arrays).
struct Tile {
    int32_t e0, e1;
    bool v0, v1;
    int32_t walls;  // bit 2 le + side: face yL/yR of element le is a slip wall
    Elem el0, el1;
};

// Whether the side `side` face of direction D of element le is a wall (only
// y faces can be); needs no neighbour index.
template <int D>
__device__ __forceinline__ bool tile_wall(const Tile& t, int le, int side) {
    return D == 1 && ((t.walls >> (2 * le + side)) & 1);
}

// Tile (t0, t1, t2) of a box: two elements adjacent along `orient`
// (0 x, 1 y, 2 z); the tile coordinate along `orient` counts pairs.
template <class UT, class CT>
__device__ __forceinline__ Tile make_tile_at(const StageArgs<UT, CT>& a, int orient, int t0, int t1, int t2) {
    Tile t;
    int ex = t0, ey = t1, ez = t2;
    int step;
    bool v1;
    if (orient == 0) {
        ex = kEpb * t0;
        v1 = ex + 1 < a.nx;
        step = 1;
    } else if (orient == 1) {
        ey = kEpb * t1;
        v1 = ey + 1 < a.ny;
        step = a.nx;
    } else {
        ez = kEpb * t2;
        v1 = ez + 1 < a.nz;
        step = a.nx * a.ny;
    }
    t.e0 = ex + a.nx * (ey + a.ny * ez);
    t.e1 = t.e0 + step;
    t.v0 = true;
    t.v1 = v1;
    t.el0 = make_elem(ex, ey, ez, a.nx, a.ny, a.nz, a.wall);
    const int d = v1 ? 1 : 0;
    t.el1 = make_elem(ex + (orient == 0 ? d : 0), ey + (orient == 1 ? d : 0), ez + (orient == 2 ? d : 0), a.nx, a.ny,
                      a.nz, a.wall);
    t.walls = (t.el0.yl < 0 ? 1 : 0) | (t.el0.yr < 0 ? 2 : 0) | (t.el1.yl < 0 ? 4 : 0) | (t.el1.yr < 0 ? 8 : 0);
    return t;
}

// Single-box launch grid: blockIdx = (pair of x elements, ey, ez).
template <class UT, class CT>
__device__ __forceinline__ Tile make_tile(const StageArgs<UT, CT>& a) {
    return make_tile_at(a, 0, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z);
}
// The same tiles in reverse order (the update kernels: they start where the
// visc kernel ended, on the data it touched last).
template <class UT, class CT>
__device__ __forceinline__ Tile make_tile_reversed(const StageArgs<UT, CT>& a) {
    return make_tile_at(a, 0, (int)(gridDim.x - 1 - blockIdx.x), (int)(gridDim.y - 1 - blockIdx.y),
                        (int)(gridDim.z - 1 - blockIdx.z));
}

__device__ __forceinline__ bool tile_valid(const Tile& t, int le) { return le ? t.v1 : t.v0; }
__device__ __forceinline__ int64_t tile_eidx(const Tile& t, int le) { return (int64_t)(le ? t.e1 : t.e0); }
// Neighbour across the side `side` face of direction D of local element le.
template <int D>
__device__ __forceinline__ int64_t tile_nb(const Tile& t, int le, int side) {
    const Elem& a = t.el0;
    const Elem& b = t.el1;
    int l0 = D == 0 ? a.xl : (D == 1 ? a.yl : a.zl), r0 = D == 0 ? a.xr : (D == 1 ? a.yr : a.zr);
    int l1 = D == 0 ? b.xl : (D == 1 ? b.yl : b.zl), r1 = D == 0 ? b.xr : (D == 1 ? b.yr : b.zr);
    int n0 = side ? r0 : l0, n1 = side ? r1 : l1;
    return (int64_t)(le ? n1 : n0);
}

// Neighbour across face f (0..5) of local element le, -1 = wall; selects on
// named fields only.
// Neighbour across face f (lane-varying) of tile element le (lane-varying):
// selects among the twelve wave-uniform indices, taken by value so that the
// selects stay in registers (a nested conditional on the struct fields
// compiles to divergent branches, a select of field references to a
// scratch-memory load).
__device__ __forceinline__ int64_t tile_nb_rt(const Tile& t, int le, int f) {
    const int xl = le ? t.el1.xl + 0 : t.el0.xl + 0, xr = le ? t.el1.xr + 0 : t.el0.xr + 0;
    const int yl = le ? t.el1.yl + 0 : t.el0.yl + 0, yr = le ? t.el1.yr + 0 : t.el0.yr + 0;
    const int zl = le ? t.el1.zl + 0 : t.el0.zl + 0, zr = le ? t.el1.zr + 0 : t.el0.zr + 0;
    const int x = (f & 1) ? xr : xl, y = (f & 1) ? yr : yl, z = (f & 1) ? zr : zl;
    return (int64_t)(f < 2 ? x : (f < 4 ? y : z));
}

// ---------------------------------------------------------------------------
// Neighbour-trace prefetch (f32 compute).  At kernel start each work-group
// copies the neighbour traces it will read straight from global memory into
// LDS with gfx950's global_load_lds_dwordx4 (no registers, one exposed
// latency), as 16-byte chunks:
//   A [0, 240)    neighbour facing traces, BR1 order, faces 0..5: 2 x 6 x 80
//   B [240, 320)  neighbour facing x traces, sequential order:   2 x 2 x 80
//                 (strict build only; the fast build has one x trace)
// The face viscous fluxes fv are read from global memory in the face pass
// instead: staging them too (another 1536 floats) limited the update kernel
// to 4.5 waves per SIMD by LDS; without them it runs 6 (a step measured
// ~6 % faster in both builds).
// The float offsets of the two groups in the buffer:
constexpr int kPfA = 0, kPfB = 960;
constexpr int kPfChunksVisc = 240;                            // visc kernel: A
constexpr int kPfChunksUpdate = kOneX ? 240 : 320;            // update kernel: A (+ B)
constexpr int kPfFloatsUpdate = 4 * kPfChunksUpdate;

// Source of chunk `id`; false for chunks of a missing second element (odd
// nx), which are not copied.
template <class UT, class CT>
__device__ __forceinline__ bool prefetch_chunk(const StageArgs<UT, CT>& a, const Tile& tl, int id, const CT*& src) {
    if (id < 240) {  // A
        const int le = id / 120, f = (id % 120) / 20, c = id % 20;
        if (!tile_valid(tl, le)) return false;
        const int64_t own = tile_eidx(tl, le), nb = tile_nb_rt(tl, le, f);
        src = a.tr_in + (nb >= 0 ? nb : own) * kTraceStride + (f ^ 1) * 80 + c * 4;
    } else {  // B
        if (kOneX) return false;  // the x Riemann inputs are group A's
        const int r = id - 240, le = r / 40, f = (r % 40) / 20, c = r % 20;
        if (!tile_valid(tl, le)) return false;
        const int64_t nb = tile_nb_rt(tl, le, f);  // x is periodic: always a neighbour
        src = a.tr_in + nb * kTraceStride + (f == 0 ? 7 : 6) * 80 + c * 4;
    }
    return true;
}

// Issue the first n_chunks chunks; waves split the 64-chunk instructions.
// The caller waits (s_waitcnt 0, then a barrier for the other wave's copies)
// before reading the buffer.
template <class UT, class CT>
__device__ __forceinline__ void prefetch_issue(const StageArgs<UT, CT>& a, const Tile& tl, float* pf, int n_chunks,
                                               int wave, int lane) {
    const int n_inst = (n_chunks + 63) / 64;
    wave = __builtin_amdgcn_readfirstlane(wave);  // wave-uniform: a scalar loop
    const int first = wave == 0 ? 0 : (n_inst + 1) / 2;
    const int last = wave == 0 ? (n_inst + 1) / 2 : n_inst;
    for (int q = first; q < last; ++q) {
        const int id = q * 64 + lane;
        const CT* src = nullptr;
        if (id < n_chunks && prefetch_chunk(a, tl, id, src)) {
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                             (__attribute__((address_space(3))) void*)(pf + q * 256), 16, 0, 0);
        }
    }
}

// BR1 jump 1/2 (u_nb + u_own) - u_own at every face point of direction D
// (accumulate_jump, rhs.cu:105-108; wall ghost rhs.cu:111-114), once per face
// point instead of once per lane of the normal line.
template <int D, bool PF, class UT, class CT>
__device__ __forceinline__ void br1_jumps(const StageArgs<UT, CT>& a, const CT (*sU)[5][64], CT (*sJmp)[6][5][16],
                                          const Tile& tl, const CT* pf, int lane) {
    const FaceTask t = face_task<D>(lane);
    if (!tile_valid(tl, t.le)) return;
    int base, stride;
    face_line(D, t.fp, base, stride);
    const Coef4<CT> c = side_coef<CT>(t.side);
    // with the prefetch the neighbour index is not needed here, only whether
    // the face is a wall (keeps 12 indices out of the registers)
    const int64_t nb = PF ? 0 : tile_nb<D>(tl, t.le, t.side);
    const bool wall = PF ? tile_wall<D>(tl, t.le, t.side) : nb < 0;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        CT own = trace4(sU[t.le][v], base, stride, c.c0, c.c1, c.c2, c.c3, br1_order(D));
        const CT sign = (v > 0 && v < 4) ? ((CT)-1.0) : ((CT)1.0);
        CT far;
        if constexpr (PF)
            far = pf[kPfA + (t.le * 6 + t.f) * 80 + v * 16 + t.fp];
        else
            far = !wall ? a.tr_in[nb * kTraceStride + (t.f ^ 1) * 80 + v * 16 + t.fp] : (CT)0;
        CT nbv = !wall ? far : own * sign;
        sJmp[t.le][t.f][v][t.fp] = ((CT)0.5) * (nbv + own) - own;
    }
}

// BR1-corrected gradient at the lane's point (rhs.cu:366-465) from the
// element's LDS tile and the face jumps; Q round trip included.
// J0..J5: the jump blocks [5][16] of faces xL, xR, yL, yR, zL, zR.
template <class UT, class CT>
__device__ __forceinline__ void lane_gradient_faces(const StageArgs<UT, CT>& a, const CT (*sU)[64], const CT* J0,
                                                    const CT* J1, const CT* J2, const CT* J3, const CT* J4,
                                                    const CT* J5, int i, int j, int k, CT* g) {
    const CoefTable<CT>& o = coef_lds<CT>();
    const CT* Di = o.D[i];
    const CT* Dj = o.D[j];
    const CT* Dk = o.D[k];
    const int xb = 4 * j + 16 * k, yb = i + 16 * k, zb = i + 4 * j;
    const int fpx = j + 4 * k, fpy = i + 4 * k, fpz = i + 4 * j;
    const CT gLi = o.g[i][0], gRi = o.g[i][1];
    const CT gLj = o.g[j][0], gRj = o.g[j][1];
    const CT gLk = o.g[k][0], gRk = o.g[k][1];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        CT dz = 0, dxi = 0, deta = 0;
#pragma unroll
        for (int m = 0; m < 4; ++m) dz += Dk[m] * sU[v][zb + 16 * m];
#pragma unroll
        for (int m = 0; m < 4; ++m) dxi += Di[m] * sU[v][xb + m];
#pragma unroll
        for (int m = 0; m < 4; ++m) deta += Dj[m] * sU[v][yb + 4 * m];
#if OD_FAST
        // one Jacobian factor per direction
        CT gx = (dxi + J0[v * 16 + fpx] * gLi + J1[v * 16 + fpx] * gRi) * a.jx;
        CT gy = (deta + J2[v * 16 + fpy] * gLj + J3[v * 16 + fpy] * gRj) * a.jy;
        CT gz = (dz + J4[v * 16 + fpz] * gLk + J5[v * 16 + fpz] * gRk) * a.jz;
#else
        CT gx = dxi * a.jx, gy = deta * a.jy, gz = dz * a.jz;
        gx += J0[v * 16 + fpx] * gLi * a.jx;
        gx += J1[v * 16 + fpx] * gRi * a.jx;
        gy += J2[v * 16 + fpy] * gLj * a.jy;
        gy += J3[v * 16 + fpy] * gRj * a.jy;
        gz += J4[v * 16 + fpz] * gLk * a.jz;
        gz += J5[v * 16 + fpz] * gRk * a.jz;
#endif
        if (a.round_q) {  // Q_REAL round trip (rhs.cu:457-461), narrows only when CT is double
            gx = (CT)(float)gx;
            gy = (CT)(float)gy;
            gz = (CT)(float)gz;
        }
        // Finish variable v before the next variable's LDS reads: left to
        // itself the compiler issues all 90 reads first and computes at the
        // first use, which needs ~140 VGPRs (and spills at 4 waves/SIMD).
        asm volatile("" : "+v"(gx), "+v"(gy), "+v"(gz)::"memory");
        g[3 * v + 0] = gx;
        g[3 * v + 1] = gy;
        g[3 * v + 2] = gz;
    }
}

template <class UT, class CT>
__device__ __forceinline__ void lane_gradient(const StageArgs<UT, CT>& a, const CT (*sU)[64], const CT (*sJmp)[5][16],
                                              int i, int j, int k, CT* g) {
    lane_gradient_faces(a, sU, &sJmp[0][0][0], &sJmp[1][0][0], &sJmp[2][0][0], &sJmp[3][0][0], &sJmp[4][0][0],
                        &sJmp[5][0][0], i, j, k, g);
}

template <bool PF, class UT, class CT>
__device__ __forceinline__ void run_br1_jumps(const StageArgs<UT, CT>& a, const CT (*sU)[5][64],
                                              CT (*sJmp)[6][5][16], const Tile& tl, const CT* pf, int wave,
                                              int lane) {
    if (wave == 0) {
        br1_jumps<0, PF>(a, sU, sJmp, tl, pf, lane);
        br1_jumps<2, PF>(a, sU, sJmp, tl, pf, lane);
    } else {
        br1_jumps<1, PF>(a, sU, sJmp, tl, pf, lane);
    }
}

// The four viscous flux components 1..4 of one point, stored together
// (fv is point-major, component fastest) so that each lane moves them with
// one 16-byte (f32) access instead of four 4-byte ones.
template <class C>
struct alignas(4 * sizeof(C)) FvQuad {
    C c[4];
};

// Viscous normal flux at the element faces of direction D (the
// dispatch_viscous_flux calls of compute_face_j / compute_face_j_wall,
// rhs.cu:284-285, 318) from the Riemann-order state trace and the gradient
// slice; only the 11 gradient components the direction reads are traced.
// One face point: element e (state tile sUe, gradient tile sGe), face side
// `side` of direction D, face point fp.
template <int D, class UT, class CT>
__device__ __forceinline__ FvQuad<CT> fv_value(const StageArgs<UT, CT>& a, const CT (*sUe)[64],
                                               const CT (*sGe)[64], int64_t e, int side, int fp) {
    int base, stride;
    face_line(D, fp, base, stride);
    const Coef4<CT> c = side_coef<CT>(side);
    CT st[5], gt[15], fvv[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) st[v] = trace4(sUe[v], base, stride, c.c0, c.c1, c.c2, c.c3, fv_order(D));
#pragma unroll
    for (int q = 0; q < 15; ++q) gt[q] = trace4_swz(sGe[q], base, stride, c.c0, c.c1, c.c2, c.c3, fv_order(D));
    CT mu = a.sutherland ? sutherland(st, a.mu_ref, a.t_ref, a.s_suth) : a.mu;
    if (a.mu_t) mu = mu + a.mu_t[e];
    viscous_flux(D, st, gt, mu, fvv);
    return FvQuad<CT>{{fvv[1], fvv[2], fvv[3], fvv[4]}};
}

template <int D, class UT, class CT>
__device__ __forceinline__ void fv_point(const StageArgs<UT, CT>& a, const CT (*sUe)[64], const CT (*sGe)[64],
                                         int64_t e, int side, int fp) {
    *reinterpret_cast<FvQuad<CT>*>(a.fv + e * kFvStride + (2 * D + side) * 64 + fp * 4) =
        fv_value<D>(a, sUe, sGe, e, side, fp);
}

template <int D, class UT, class CT>
__device__ __forceinline__ void fv_pass(const StageArgs<UT, CT>& a, const CT (*sU)[5][64], const CT (*sG)[15][64],
                                        const Tile& tl, int lane) {
    const FaceTask t = face_task<D>(lane);
    if (!tile_valid(tl, t.le)) return;
    fv_point<D>(a, sU[t.le], sG[t.le], tile_eidx(tl, t.le), t.side, t.fp);
}

// Own-side face correction j from the interface flux fr of a face point
// (compute_face_j / compute_face_j_wall, rhs.cu:273-339): own is the
// element's trace, fvo / fvn the own and the neighbour's viscous normal flux
// (component 0 zero), right = the face is the element's right face.
template <int D, bool VISC, class CT>
__device__ __forceinline__ void face_correction(const CT* fr, const CT* own, const CT* fvo, const CT* fvn, bool wall,
                                                bool right, CT* jv) {
    OD_CONTRACT_IN_EXPRESSIONS
    CT fe[5];
    euler_flux(D, own, fe);
    CT wall_p = (CT)0;
    if (D == 1) wall_p = fe[2] - dv(own[2] * own[2], own[0]);  // rhs.cu:323-326
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const CT wf = (D == 1 && v == 2) ? wall_p : (CT)0.0;
        if (VISC) {
            const CT fl = right ? fvo[v] : fvn[v];
            const CT frr = right ? fvn[v] : fvo[v];
            CT jn;
            if constexpr (kHalfDiff<CT>) {
                // F* - (F_e - F_v,own) = (F_riemann - F_e) +- (F_vL - F_vR) / 2
                const CT hd = ((CT)0.5) * (fl - frr);
                jn = (fr[v] - fe[v]) + (right ? hd : -hd);
            } else {
                CT fc = fr[v] - ((CT)0.5) * (fl + frr);
                jn = fc - (fe[v] - fvo[v]);
            }
            CT jw = (wf - fvo[v]) - (fe[v] - fvo[v]);
            jv[v] = wall ? jw : jn;
        } else {
            jv[v] = wall ? (wf - fe[v]) : (fr[v] - fe[v]);
        }
    }
}

// The same correction at an interior face whose viscous half-difference
// hd = (F_vL - F_vR) / 2 (component 0 zero) is stored (fast f32 build).
template <int D, class CT>
__device__ __forceinline__ void face_correction_hd(const CT* fr, const CT* own, const CT* hd, bool right, CT* jv) {
    OD_CONTRACT_IN_EXPRESSIONS
    CT fe[5];
    euler_flux(D, own, fe);
#pragma unroll
    for (int v = 0; v < 5; ++v) jv[v] = (fr[v] - fe[v]) + (right ? hd[v] : -hd[v]);
}

// Interface flux at the element faces of direction D -> own-side correction
// j into LDS (compute_face_j / compute_face_j_wall, rhs.cu:273-339).
template <int D, bool VISC, bool PF, class UT, class CT>
__device__ __forceinline__ void flux_pass(const StageArgs<UT, CT>& a, const CT (*sU)[5][64], CT (*sJ)[6][5][16],
                                          const Tile& tl, const CT* pf, int lane) {
    const FaceTask t = face_task<D>(lane);
    if (!tile_valid(tl, t.le)) return;
    int base, stride;
    face_line(D, t.fp, base, stride);
    const Coef4<CT> c = side_coef<CT>(t.side);
    const bool right = t.side;
    CT own[5], oth[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) own[v] = trace4(sU[t.le][v], base, stride, c.c0, c.c1, c.c2, c.c3, face_order(D));
    const int64_t nb = (PF && !VISC) ? 0 : tile_nb<D>(tl, t.le, t.side);
    const bool wall = PF ? tile_wall<D>(tl, t.le, t.side) : nb < 0;
    // the neighbour's facing trace in the same order (x: 7/6 strict, 1/0 fast;
    // y: 3/2; z: 5/4)
    const int rf = (D == 0 && !kOneX) ? (right ? 6 : 7) : (t.f ^ 1);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        CT far;
        if constexpr (PF)
            far = (D == 0 && !kOneX) ? pf[kPfB + (t.le * 2 + t.side) * 80 + v * 16 + t.fp]
                         : pf[kPfA + (t.le * 6 + t.f) * 80 + v * 16 + t.fp];
        else
            far = wall ? own[v] : a.tr_in[nb * kTraceStride + rf * 80 + v * 16 + t.fp];
        oth[v] = wall ? own[v] : far;
    }
    CT fvo[5] = {(CT)0, (CT)0, (CT)0, (CT)0, (CT)0}, fvn[5] = {(CT)0, (CT)0, (CT)0, (CT)0, (CT)0};
    if (VISC) {
        const FvQuad<CT> o =
            *reinterpret_cast<const FvQuad<CT>*>(a.fv + tile_eidx(tl, t.le) * kFvStride + t.f * 64 + t.fp * 4);
        FvQuad<CT> n = o;
        if (!wall) n = *reinterpret_cast<const FvQuad<CT>*>(a.fv + nb * kFvStride + (t.f ^ 1) * 64 + t.fp * 4);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            fvo[1 + q] = o.c[q];
            fvn[1 + q] = n.c[q];
        }
    }
    CT L[5], R[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        L[v] = right ? own[v] : oth[v];
        R[v] = right ? oth[v] : own[v];
    }
    CT fr[5], jv[5];
    interface_flux<D>(a.flux, L, R, fr);
    face_correction<D, VISC>(fr, own, fvo, fvn, wall, right, jv);
#pragma unroll
    for (int v = 0; v < 5; ++v) sJ[t.le][t.f][v][t.fp] = jv[v];
}

// Index of this wave in its work-group, as a wave-uniform (scalar) value: the
// compiler cannot tell that threadIdx.x >> 6 is the same in every lane, and
// per-wave conditions derived from it (the wave's element is valid, the
// wave's face directions) would otherwise branch on the execution mask.
__device__ __forceinline__ int wave_index() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// Orders this wave's earlier LDS writes before its later LDS reads (of
// other lanes' data): a wave's LDS operations execute in order, so this only
// has to keep the compiler from moving them.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Streaming (non-temporal) accesses of the update kernels: the new state and
// its traces are read next by the visc kernel, which is bound by instruction
// issue rather than by memory, and the previous stage's state is not read
// again in this stage.  Marked as streaming they do not displace, in the
// memory-side cache, the viscous fluxes, state and traces that the visc
// kernel has just touched and the update kernel is about to read.
template <class T>
__device__ __forceinline__ T stream_load(const T* p) { return __builtin_nontemporal_load(p); }
template <class T>
__device__ __forceinline__ void stream_store(T v, T* p) { __builtin_nontemporal_store(v, p); }

// Face traces of one element (its state tile sUe in LDS) -> dst, the
// element's kTraceStride block of tr_out.  The element's own wave writes
// them: 16 lanes per face, 4 faces per pass.  Without BR1 (Euler) the strict
// build skips the pairwise x traces, which only BR1 reads.  STREAM: streaming
// stores (the update kernels).
template <class CT, bool STREAM = false>
__device__ __forceinline__ void write_traces_own(const CT (*sUe)[64], CT* dst, int lane, bool need_pairwise_x) {
    const int fp = lane & 15;
#pragma unroll
    for (int pass = 0; pass < (kOwnFaces + 3) / 4; ++pass) {
        const int f = 4 * pass + (lane >> 4);
        if (f >= kOwnFaces || (!kOneX && !need_pairwise_x && f < 2)) continue;
        int base, stride;
        face_line(trace_face_dir(f), fp, base, stride);
        const Coef4<CT> c = side_coef<CT>(f & 1);
        const int order = trace_face_order(f);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            const CT tv = trace4(sUe[v], base, stride, c.c0, c.c1, c.c2, c.c3, order);
            if constexpr (STREAM)
                stream_store(tv, &dst[f * 80 + v * 16 + fp]);
            else
                dst[f * 80 + v * 16 + fp] = tv;
        }
    }
}

// State tile of element e into LDS, sUe[v][pt] = u[v][e][pt].  With an f32
// state and f32 compute each wave copies its element with LDS-DMA
// (global_load_lds_dword: no registers; the data lands with the
// neighbour-trace prefetch and is covered by the same wait); otherwise
// through registers with the cast to the compute type.
// Cache-policy operand of an LDS-DMA load: 0, or the non-temporal (streaming)
// bit of gfx950's cache policy, for data read once.
constexpr int kDmaDefault = 0;
constexpr int kDmaStream = 2;

template <class UT, class CT, int POLICY = kDmaDefault>
__device__ __forceinline__ void load_state_tile(const UT* u, int64_t NP, int64_t e, CT (*sUe)[64], int pt) {
    if constexpr (std::is_same<UT, float>::value && std::is_same<CT, float>::value) {
#pragma unroll
        for (int v = 0; v < 5; ++v)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(u + v * NP + e * 64 + pt),
                                             (__attribute__((address_space(3))) void*)&sUe[v][0], 4, 0, POLICY);
    } else {
#pragma unroll
        for (int v = 0; v < 5; ++v) sUe[v][pt] = (CT)u[v * NP + e * 64 + pt];
    }
}

// Divergence of the volume viscous fluxes at the lane's point (the F_v half
// of accumulate_volume_rhs_shfl / _z, rhs.cu:229-266) from the corrected
// gradient g, x then y then z, components 1..4 (F_v of the mass equation is
// zero) -> the element's fv block after its face part.  S: this wave's
// 256-float LDS staging area.  Used by the fast f32 build (kSplitVisc).
template <class UT, class CT>
__device__ __forceinline__ void visc_divergence(const StageArgs<UT, CT>& a, int64_t e, const CT (*sUe)[64],
                                                const CT* g, CT* S, int i, int j, int k, int pt) {
    CT st[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) st[v] = sUe[v][pt];
    CT mu = a.sutherland ? sutherland(st, a.mu_ref, a.t_ref, a.s_suth) : a.mu;
    if (a.mu_t) mu = mu + a.mu_t[e];
    const ViscPoint<CT> vp = viscous_point(st, g, mu, 7);
    CT acc[4] = {(CT)0, (CT)0, (CT)0, (CT)0};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        CT fvv[5];
        viscous_flux_dir(d, vp, fvv);
        if constexpr (kDppX<CT>) {
            if (d == 0) {
                const CT* Dx = coef_lds<CT>().D[i];
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] += quad_line(Dx, fvv[1 + q]) * a.jx;
                continue;
            }
        }
        if (d > (kDppX<CT> ? 1 : 0)) wave_sync();  // the previous direction's reads are done
#pragma unroll
        for (int q = 0; q < 4; ++q) S[q * 64 + pt] = fvv[1 + q];
        wave_sync();
        const int idx = d == 0 ? i : (d == 1 ? j : k);
        const int base = d == 0 ? 4 * j + 16 * k : (d == 1 ? i + 16 * k : i + 4 * j);
        const int stride = d == 0 ? 1 : (d == 1 ? 4 : 16);
        const CT jac = d == 0 ? a.jx : (d == 1 ? a.jy : a.jz);
        const CT* Dd = coef_lds<CT>().D[idx];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            CT dd = 0;
#pragma unroll
            for (int m = 0; m < 4; ++m) dd += Dd[m] * S[q * 64 + base + m * stride];
            acc[q] += dd * jac;
        }
    }
    FvQuad<CT>* dst = reinterpret_cast<FvQuad<CT>*>(a.fv + e * kFvStride + kFvFaces + pt * 4);
    *dst = FvQuad<CT>{{acc[0], acc[1], acc[2], acc[3]}};
}

// ---------------------------------------------------------------------------
// Kernels.
// ---------------------------------------------------------------------------
// LDS of one tile (two elements).  The prefetch buffer comes first so that
// its 16-byte LDS-DMA rows start on a 16-byte boundary.
// With the prefetch (f32 compute) the gradient tile sG reuses the prefetched
// traces and the BR1 jumps (960 floats each, contiguous): element 0's
// gradients take the prefetch buffer, element 1's the jumps, written only
// after both waves have finished reading them (a barrier).  Between the
// jumps and the gradient tile the fast f32 build stages the volume viscous
// fluxes of element w in pf[256 w, 256 w + 256) (visc_divergence): the
// prefetched traces are consumed by then and sG is written only after the
// next barrier.
template <class CT>
struct ViscLds {
    static constexpr bool kPF = std::is_same<CT, float>::value;
    float pf[kPF ? kPfB : 4];
    CT sJmp[kEpb][6][5][16];
    CT sU[kEpb][5][64];
    CT sG_own[kPF ? 0 : kEpb][15][64];  // zero-length when aliased
    __device__ CT (*sG())[15][64] {
        static_assert(!kPF || (kPfB == 15 * 64 && sizeof(sJmp) == kPfB * sizeof(float)), "sG aliasing layout");
        if constexpr (kPF) return reinterpret_cast<CT(*)[15][64]>(pf);
        else return sG_own;
    }
};

// With the prefetch (f32 compute) the face-correction array sJ shares the
// prefetched neighbour traces (group A has sJ's exact layout and each face
// point's lane reads its entries there before it writes them), and the
// volume-flux staging sF of element w shares the BR1 jumps of element w
// (read only by wave w, before it writes sF).
template <class CT, bool VISC>
struct UpdateLds {
    static constexpr bool kPF = std::is_same<CT, float>::value;
    static constexpr bool kJmp = VISC && !kSplitVisc<CT>;  // BR1 jumps formed in this kernel
    float pf[kPF ? kPfFloatsUpdate : 4];
    CT sU[kEpb][5][64];
    CT sJmp[kJmp ? kEpb : 0][6][5][16];
    CT sJ_own[kPF ? 0 : kEpb][6][5][16];  // zero-length when aliased
    CT sF_own[kJmp ? 0 : kEpb][5][64];
    __device__ CT (*sJ())[6][5][16] {
        if constexpr (kPF) return reinterpret_cast<CT(*)[6][5][16]>(pf + kPfA);
        else return sJ_own;
    }
    __device__ CT (*sF(int w))[64] {
        if constexpr (kJmp) return reinterpret_cast<CT(*)[64]>(&sJmp[w][0][0][0]);
        else return sF_own[w];
    }
};

// Face traces of the state -> tr_out (wave w loads element w of the tile).
template <class UT, class CT>
__device__ __forceinline__ void trace_body(const StageArgs<UT, CT>& a, const Tile& tl, CT (*sU)[5][64], int wave,
                                           int pt) {
    const CT cf = coef_load<CT>((int)threadIdx.x);
    const int64_t NP = (int64_t)a.nx * a.ny * a.nz * 64;
    if (tile_valid(tl, wave)) {
#pragma unroll
        for (int v = 0; v < 5; ++v) sU[wave][v][pt] = (CT)a.uin[v * NP + tile_eidx(tl, wave) * 64 + pt];
    }
    coef_store<CT>((int)threadIdx.x, cf);
    __syncthreads();
    if (tile_valid(tl, wave)) write_traces_own(sU[wave], a.tr_out + tile_eidx(tl, wave) * kTraceStride, pt, true);
}

// Corrected gradients of the tile's elements and the viscous fluxes at
// their own faces -> fv.
template <class UT, class CT>
__device__ __forceinline__ void visc_body(const StageArgs<UT, CT>& a, const Tile& tl, ViscLds<CT>& L, int wave,
                                          int pt) {
    constexpr bool PF = std::is_same<CT, float>::value;
    const int i = pt & 3, j = (pt >> 2) & 3, k = pt >> 4;
    const int64_t NP = (int64_t)a.nx * a.ny * a.nz * 64;
    const CT cf = coef_load<CT>((int)threadIdx.x);
    if constexpr (PF) prefetch_issue(a, tl, L.pf, kPfChunksVisc, wave, pt);
    if (tile_valid(tl, wave)) load_state_tile(a.uin, NP, tile_eidx(tl, wave), L.sU[wave], pt);
    coef_store<CT>((int)threadIdx.x, cf);
    if constexpr (PF) __builtin_amdgcn_s_waitcnt(0);  // this wave's LDS copies have landed
    __syncthreads();
    run_br1_jumps<PF>(a, L.sU, L.sJmp, tl, reinterpret_cast<const CT*>(L.pf), wave, pt);
    __syncthreads();
    CT g[15];
    if (tile_valid(tl, wave)) lane_gradient(a, L.sU[wave], L.sJmp[wave], i, j, k, g);
    if constexpr (kSplitVisc<CT>) {
        if (tile_valid(tl, wave))
            visc_divergence(a, tile_eidx(tl, wave), L.sU[wave], g, L.pf + wave * 256, i, j, k, pt);
    }
    if constexpr (PF) __syncthreads();  // sG overwrites the jumps the other wave may still read
    if (tile_valid(tl, wave)) {
#pragma unroll
        for (int q = 0; q < 15; ++q) L.sG()[wave][q][swz64(pt)] = g[q];
    }
    __syncthreads();
    if (wave == 0) {
        fv_pass<0>(a, L.sU, L.sG(), tl, pt);
        fv_pass<2>(a, L.sU, L.sG(), tl, pt);
    } else {
        fv_pass<1>(a, L.sU, L.sG(), tl, pt);
    }
}

// Full stage for the tile: volume RHS, interface fluxes, FR correction, RK
// blend, and the face traces of the new state.
template <class UT, class CT, bool VISC>
__device__ __forceinline__ void update_body(const StageArgs<UT, CT>& a, const Tile& tl, UpdateLds<CT, VISC>& L,
                                            int wave, int pt) {
    constexpr bool PF = std::is_same<CT, float>::value;
    const CT* pfc = reinterpret_cast<const CT*>(L.pf);
    const int i = pt & 3, j = (pt >> 2) & 3, k = pt >> 4;
    const bool mine = tile_valid(tl, wave);
    const int64_t NP = (int64_t)a.nx * a.ny * a.nz * 64;
    const int64_t gbase = tile_eidx(tl, wave) * 64 + pt;
    const CT cf = coef_load<CT>((int)threadIdx.x);
    if constexpr (PF) prefetch_issue(a, tl, L.pf, kPfChunksUpdate, wave, pt);

    constexpr bool SAME = std::is_same<UT, CT>::value;  // sU holds the state exactly
    UT ucur[5];
    if constexpr (SAME && PF) {  // f32 state and compute: LDS-DMA, read back from sU below
        if (mine) {
            load_state_tile(a.uin, NP, tile_eidx(tl, wave), L.sU[wave], pt);
        } else {
#pragma unroll
            for (int v = 0; v < 5; ++v) L.sU[wave][v][pt] = (CT)1;
        }
    } else {
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            ucur[v] = mine ? a.uin[v * NP + gbase] : (UT)1;
            L.sU[wave][v][pt] = (CT)ucur[v];
        }
    }
    coef_store<CT>((int)threadIdx.x, cf);
    if constexpr (PF) __builtin_amdgcn_s_waitcnt(0);  // this wave's LDS copies have landed
    __syncthreads();

    // with kSplitVisc the volume viscous terms come from the visc kernel
    constexpr bool VOLV = VISC && !kSplitVisc<CT>;
    CT g[15];
    if constexpr (VOLV) {
        run_br1_jumps<PF>(a, L.sU, L.sJmp, tl, pfc, wave, pt);
        __syncthreads();
        lane_gradient(a, L.sU[wave], L.sJmp[wave], i, j, k, g);
    }
    CT st[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) st[v] = L.sU[wave][v][pt];
    CT mu = (CT)0;
    if (VOLV) mu = a.sutherland ? sutherland(st, a.mu_ref, a.t_ref, a.s_suth) : a.mu;
    if (VOLV && a.mu_t) mu = mu + (mine ? a.mu_t[tile_eidx(tl, wave)] : (CT)0);
    CT vdiv[4] = {(CT)0, (CT)0, (CT)0, (CT)0};
    if (VISC && !VOLV && mine) {
        const FvQuad<CT> src =
            *reinterpret_cast<const FvQuad<CT>*>(a.fv + tile_eidx(tl, wave) * kFvStride + kFvFaces + pt * 4);
#pragma unroll
        for (int q = 0; q < 4; ++q) vdiv[q] = src.c[q];
    }

    // Volume flux divergence, x then y then z (rhs.cu:229-266, 490-510).
    CT rs[5] = {(CT)0, (CT)0, (CT)0, (CT)0, (CT)0};
    ViscPoint<CT> vp;
    if (VOLV) vp = viscous_point(st, g, mu, 7);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        CT fe[5], F[5];
        euler_flux(d, st, fe);
        if (VOLV) {
            CT fvv[5];
            viscous_flux_dir(d, vp, fvv);
#pragma unroll
            for (int v = 0; v < 5; ++v) F[v] = fe[v] - fvv[v];
        } else {
#pragma unroll
            for (int v = 0; v < 5; ++v) F[v] = fe[v];
        }
        if constexpr (kDppX<CT>) {
            if (d == 0) {
                const CT* Dx = coef_lds<CT>().D[i];
#pragma unroll
                for (int v = 0; v < 5; ++v) rs[v] += quad_line(Dx, F[v]) * a.jx;
                continue;
            }
        }
        if (d > (kDppX<CT> ? 1 : 0)) __syncthreads();
#pragma unroll
        for (int v = 0; v < 5; ++v) L.sF(wave)[v][pt] = F[v];
        __syncthreads();
        const int idx = d == 0 ? i : (d == 1 ? j : k);
        const int base = d == 0 ? 4 * j + 16 * k : (d == 1 ? i + 16 * k : i + 4 * j);
        const int stride = d == 0 ? 1 : (d == 1 ? 4 : 16);
        const CT jac = d == 0 ? a.jx : (d == 1 ? a.jy : a.jz);
        const CT* Dd = coef_lds<CT>().D[idx];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            CT dd = 0;
#pragma unroll
            for (int m = 0; m < 4; ++m) dd += Dd[m] * L.sF(wave)[v][base + m * stride];
            rs[v] += dd * jac;
        }
    }
    // the volume RHS waits in LDS (this wave's sF slot) during the face passes
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        CT val = -rs[v];
        if (VISC && !VOLV && v > 0) val += vdiv[v - 1];  // + div F_v from the visc kernel
        if (v == 1) val += a.bf;
        L.sF(wave)[v][pt] = a.round_rhs ? (CT)(float)val : val;  // RHS_REAL round trip (rhs.cu:509, rk.cu:52)
    }

    // Interface fluxes at all own faces, direction by direction.
    if (wave == 0) {
        flux_pass<0, VISC, PF>(a, L.sU, L.sJ(), tl, pfc, pt);
        flux_pass<2, VISC, PF>(a, L.sU, L.sJ(), tl, pfc, pt);
    } else {
        flux_pass<1, VISC, PF>(a, L.sU, L.sJ(), tl, pfc, pt);
    }
    __syncthreads();

    // FR correction and stage update (rk.cu:12-25, 63-82).
    const CoefTable<CT>& o = coef_lds<CT>();
#if OD_FAST
    // correction slopes with the Jacobian folded in
    const CT gLi = o.g[i][0] * a.jx, gRi = o.g[i][1] * a.jx;
    const CT gLj = o.g[j][0] * a.jy, gRj = o.g[j][1] * a.jy;
    const CT gLk = o.g[k][0] * a.jz, gRk = o.g[k][1] * a.jz;
#else
    const CT gLi = o.g[i][0], gRi = o.g[i][1];
    const CT gLj = o.g[j][0], gRj = o.g[j][1];
    const CT gLk = o.g[k][0], gRk = o.g[k][1];
#endif
    const int fpx = j + 4 * k, fpy = i + 4 * k, fpz = i + 4 * j;
    UT uprv[5], ucr[5];
    CT r[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        ucr[v] = SAME ? (UT)L.sU[wave][v][pt] : ucur[v];
        r[v] = L.sF(wave)[v][pt];
    }
    if (a.stage1) {  // u^n is the state this stage read (uprev == uin): no second load
#pragma unroll
        for (int v = 0; v < 5; ++v) uprv[v] = ucr[v];
    } else {
#pragma unroll
        for (int v = 0; v < 5; ++v) uprv[v] = mine ? stream_load(&a.uprev[v * NP + gbase]) : (UT)1;
    }
    CT rts[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const CT(*J)[5][16] = L.sJ()[wave];
#if OD_FAST
        CT cx = -(J[0][v][fpx] * gLi + J[1][v][fpx] * gRi);
        CT cy = -(J[2][v][fpy] * gLj + J[3][v][fpy] * gRj);
        CT cz = -(J[4][v][fpz] * gLk + J[5][v][fpz] * gRk);
#else
        CT cx = -(J[0][v][fpx] * gLi * a.jx) - (J[1][v][fpx] * gRi * a.jx);
        CT cy = -(J[2][v][fpy] * gLj * a.jy) - (J[3][v][fpy] * gRj * a.jy);
        CT cz = -(J[4][v][fpz] * gLk * a.jz) - (J[5][v][fpz] * gRk * a.jz);
#endif
        rts[v] = r[v] + (cx + cy + cz);
    }
    if (a.mode == 1) {  // the stage's RHS into dbg, state unchanged
#pragma unroll
        for (int v = 0; v < 5; ++v)
            if (mine) a.dbg[v * NP + gbase] = rts[v];
        return;
    }
    UT unew[5];
    if constexpr (std::is_same<UT, CT>::value) {
        // one type: the compute-type and the state-type blends differ only
        // in their coefficients, chosen once; uniform branch on the stage
        const CT kA = a.accum_state ? a.ua : a.ca;
        const CT kB = a.accum_state ? a.ub : a.cb;
        const CT kC = a.accum_state ? a.uc : a.cc;
        if (a.stage1) {
#pragma unroll
            for (int v = 0; v < 5; ++v) unew[v] = rk_blend1(kA, kC, uprv[v], rts[v]);
        } else {
#pragma unroll
            for (int v = 0; v < 5; ++v) unew[v] = rk_blend(kA, kB, kC, uprv[v], ucr[v], rts[v]);
        }
    } else {
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            const CT rt = rts[v];
            if (!a.accum_state) {
                CT up = (CT)uprv[v], uc = (CT)ucr[v];
                CT val = a.stage1 ? (a.ca * up + a.cc * rt) : (a.ca * up + a.cb * uc + a.cc * rt);
                unew[v] = (UT)val;
            } else {
                UT rr = (UT)rt;
                unew[v] =
                    a.stage1 ? (UT)(a.ua * uprv[v] + a.uc * rr) : (UT)(a.ua * uprv[v] + a.ub * ucr[v] + a.uc * rr);
            }
        }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v)
        if (mine) stream_store(unew[v], &a.uout[v * NP + gbase]);

    // Face traces of the new state for the next stage, each wave those of its
    // own element: from the RK phase on only wave w reads sU[w], so no
    // work-group barrier is needed.
#pragma unroll
    for (int v = 0; v < 5; ++v) L.sU[wave][v][pt] = (CT)unew[v];
    wave_sync();
    if (mine) write_traces_own<CT, true>(L.sU[wave], a.tr_out + tile_eidx(tl, wave) * kTraceStride, pt, VISC);
}

template <class UT, class CT>
__global__ __launch_bounds__(128) void od_trace_kernel(StageArgs<UT, CT> a) {
    __shared__ CT sU[kEpb][5][64];
    trace_body(a, make_tile(a), sU, wave_index(), threadIdx.x & 63);
}

template <class UT, class CT>
__global__ __launch_bounds__(128) void od_visc_kernel(StageArgs<UT, CT> a) {
    __shared__ ViscLds<CT> L;
    visc_body(a, make_tile(a), L, wave_index(), threadIdx.x & 63);
}

// Occupancy target of the update kernels.  The strict f32 build needs ~78
// VGPRs (6 waves per SIMD).  The fast all-f32 single-box kernel and the batch
// kernel's single-unit form, which need no gradient there (kSplitVisc), are
// held to 64 (8 waves; they need 61 with no spills; their LDS, 9.1 KB per
// work-group, allows 8).  The strict build, f64 states and the persistent batch kernel
// keep their natural allocation (the 4-wave cap does not bind); f64 compute
// needs ~145 VGPRs.
template <class CT>
constexpr int kMinWaves = sizeof(CT) == 4 ? 4 : 1;
template <class UT, class CT>
constexpr int kMinWavesUpdate = (OD_FAST && sizeof(UT) == 4 && sizeof(CT) == 4) ? 8 : kMinWaves<CT>;

template <class UT, class CT, bool VISC>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(kMinWavesUpdate<UT, CT>))) void od_update_kernel(
    StageArgs<UT, CT> a) {
    __shared__ UpdateLds<CT, VISC> L;
    update_body<UT, CT, VISC>(a, make_tile_reversed(a), L, wave_index(), threadIdx.x & 63);
}

#if OD_FAST
// ---------------------------------------------------------------------------
// Fast f32 build, one box with even nx, ny, nz: the update kernel on blocks
// of 2x2x2 elements (512 threads; wave w owns local element w = lx + 2 ly +
// 4 lz in the point-wise phases).  Three of an element's six faces face
// another element of the block.  Their neighbour traces come from the
// block's LDS state tiles instead of HBM, and the interface flux of such an
// internal face is evaluated once for both of its elements.  Every value is
// formed from the same operands in the same order as in update_body, so each
// element gets the same bits as in the two-element-tile kernels (which the
// ensemble kernels and meshes with an odd dimension use).
// ---------------------------------------------------------------------------
constexpr int kBlk = 8;

struct Block {
    int32_t ex0, ey0, ez0;  // coordinates of local element 0 (even)
    int32_t e0;             // its element index
};

// REVERSE: the work-group -> block map runs backwards (the last block of the
// mesh first).  The visc kernel walks the mesh forwards and the update kernel
// backwards, so each kernel starts where the previous one ended, on the data
// that kernel touched last (its neighbour traces, viscous fluxes, state),
// which is the most likely to still be in the memory-side cache.
template <bool REVERSE, class UT, class CT>
__device__ __forceinline__ Block make_block(const StageArgs<UT, CT>& a) {
    Block b;
    const int bx = REVERSE ? (int)(gridDim.x - 1 - blockIdx.x) : (int)blockIdx.x;
    const int by = REVERSE ? (int)(gridDim.y - 1 - blockIdx.y) : (int)blockIdx.y;
    const int bz = REVERSE ? (int)(gridDim.z - 1 - blockIdx.z) : (int)blockIdx.z;
    b.ex0 = 2 * bx;
    b.ey0 = 2 * by;
    b.ez0 = 2 * bz;
    b.e0 = b.ex0 + a.nx * (b.ey0 + a.ny * b.ez0);
    return b;
}

// Bit d of a local element index: its coordinate along d in the block, and
// the side (0 left, 1 right) of its external face normal to d.
__device__ __forceinline__ int blk_bit(int le, int d) { return (le >> d) & 1; }

template <class UT, class CT>
__device__ __forceinline__ int blk_elem(const StageArgs<UT, CT>& a, const Block& b, int le) {
    return b.e0 + (le & 1) + a.nx * (((le >> 1) & 1) + a.ny * (le >> 2));
}

// Neighbour across the external face normal to d of local element le (as
// make_elem forms it); -1 = slip wall.  Branch-free: the offsets of the
// block's left (bit d = 0) and right (bit d = 1) external neighbours are
// block-uniform, each lane selects by its element's bit.
template <class UT, class CT>
__device__ __forceinline__ int blk_ext_nb(const StageArgs<UT, CT>& a, const Block& b, int le, int d) {
    const int e = blk_elem(a, b, le);
    const int s = blk_bit(le, d);
    const int n = d == 0 ? a.nx : (d == 1 ? a.ny : a.nz);
    const int c0 = d == 0 ? b.ex0 : (d == 1 ? b.ey0 : b.ez0);  // the block's left coordinate along d
    const int stride = d == 0 ? 1 : (d == 1 ? a.nx : a.nx * a.ny);
    const int off_l = (c0 == 0 ? n - 1 : -1) * stride;
    const int off_r = (c0 + 1 == n - 1 ? 1 - n : 1) * stride;
    const bool wall_l = d == 1 && a.wall && c0 == 0;
    const bool wall_r = d == 1 && a.wall && c0 + 1 == n - 1;
    return (s ? wall_r : wall_l) ? -1 : e + (s ? off_r : off_l);
}

// Source of 16-byte chunk `chunk` (0..59) of the external-face traces of
// local element le, as the LDS prefetch lays them out ([d][5][16]): the
// neighbour's facing trace (its own at a slip wall, where it is not used).
// le is wave-uniform, so the three neighbours are scalar work and each lane
// only selects its direction's.
template <class UT, class CT>
__device__ __forceinline__ const UT* blk_ext_trace_src(const StageArgs<UT, CT>& a, const Block& b, int le,
                                                       int chunk) {
    const int nb0 = blk_ext_nb(a, b, le, 0), nb1 = blk_ext_nb(a, b, le, 1), nb2 = blk_ext_nb(a, b, le, 2);
    const int d = chunk >= 40 ? 2 : (chunk >= 20 ? 1 : 0), c = chunk - 20 * d;
    const int nb = d == 0 ? nb0 : (d == 1 ? nb1 : nb2);
    const int src_e = nb >= 0 ? nb : blk_elem(a, b, le);
    const int f = 2 * d + blk_bit(le, d);
    return a.tr_in + (int64_t)src_e * kTraceStride + (f ^ 1) * 80 + c * 4;
}

// Traces of the element's three external faces (face 2d + bit d of le, one
// face point per lane, lanes 48..63 idle), as write_traces_own forms them.
__device__ __forceinline__ void blk_write_ext_traces(const float (*sUe)[64], float* dst, int le, int lane) {
    const int d = lane >> 4, fp = lane & 15;
    if (d > 2) return;
    const int f = 2 * d + blk_bit(le, d);
    int base, stride;
    face_line(d, fp, base, stride);
    const Coef4<float> c = side_coef<float>(f & 1);
#pragma unroll
    for (int v = 0; v < 5; ++v)
        stream_store(trace4(sUe[v], base, stride, c.c0, c.c1, c.c2, c.c3, trace_face_order(f)),
                     &dst[f * 80 + v * 16 + fp]);
}

// State tiles of a block's elements, one every 5 * 64 + 16 floats: the pad
// puts consecutive elements' tiles 16 banks apart, so the two elements whose
// face points one half-wave reads in a face pass do not share banks.
struct BlockTiles {
    float t[kBlk][5 * 64 + 16];
    __device__ float (*operator[](int le))[64] { return reinterpret_cast<float(*)[64]>(t[le]); }
};

// LDS of a block.  sJe[le][d]: the external face of direction d of element
// le -- first the neighbour's facing trace (LDS-DMA, chunk id = 4 x its
// float offset), then the element's correction j there; sJi[le][d]: j at
// the internal face of direction d.
struct BlockUpdateLds {
    alignas(16) float sJe[kBlk][3][5][16];
    float sJi[kBlk][3][5][16];
    BlockTiles sU;
    float sF[kBlk][5][64];  // volume-flux line staging, then the volume RHS
};

// External faces of direction D of the four elements 4h .. 4h+3: one face
// point per lane.  The arithmetic is flux_pass's.
template <int D, bool VISC>
__device__ __forceinline__ void blk_ext_pass(const StageArgs<float, float>& a, const Block& b, BlockUpdateLds& L,
                                             int h, int lane) {
    const int le = 4 * h + (lane >> 4), fp = lane & 15;
    const int side = blk_bit(le, D);
    const bool right = side;
    int base, stride;
    face_line(D, fp, base, stride);
    const Coef4<float> c = side_coef<float>(side);
    float own[5], oth[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) own[v] = trace4(L.sU[le][v], base, stride, c.c0, c.c1, c.c2, c.c3, face_order(D));
    const int nb = blk_ext_nb(a, b, le, D);
    const bool wall = D == 1 && nb < 0;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const float far = L.sJe[le][D][v][fp];
        oth[v] = wall ? own[v] : far;
    }
    float fvo[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, fvn[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    if (VISC) {
        const int f = 2 * D + side;
        const FvQuad<float> o =
            *reinterpret_cast<const FvQuad<float>*>(a.fv + (int64_t)blk_elem(a, b, le) * kFvStride + f * 64 + fp * 4);
        FvQuad<float> n = o;
        if (!wall) n = *reinterpret_cast<const FvQuad<float>*>(a.fv + (int64_t)nb * kFvStride + (f ^ 1) * 64 + fp * 4);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            fvo[1 + q] = o.c[q];
            fvn[1 + q] = n.c[q];
        }
    }
    float Lq[5], Rq[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        Lq[v] = right ? own[v] : oth[v];
        Rq[v] = right ? oth[v] : own[v];
    }
    float fr[5], jv[5];
    interface_flux<D>(a.flux, Lq, Rq, fr);
    face_correction<D, VISC>(fr, own, fvo, fvn, wall, right, jv);
#pragma unroll
    for (int v = 0; v < 5; ++v) L.sJe[le][D][v][fp] = jv[v];
}

// Internal faces of direction D: pair p = lane / 16 (left element lL, right
// element lL + 2^D), one face point per lane; one interface flux for both
// sides, each side's correction as flux_pass forms it.
template <int D, bool VISC>
__device__ __forceinline__ void blk_int_pass(const StageArgs<float, float>& a, const Block& b, BlockUpdateLds& L,
                                             int lane) {
    const int p = lane >> 4, fp = lane & 15;
    const int lL = D == 0 ? 2 * p : (D == 1 ? ((p & 1) | ((p >> 1) << 2)) : p);
    const int lR = lL + (1 << D);
    int base, stride;
    face_line(D, fp, base, stride);
    const Coef4<float> cr = side_coef<float>(true), cl = side_coef<float>(false);
    float tl[5], tr[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) tl[v] = trace4(L.sU[lL][v], base, stride, cr.c0, cr.c1, cr.c2, cr.c3, face_order(D));
#pragma unroll
    for (int v = 0; v < 5; ++v) tr[v] = trace4(L.sU[lR][v], base, stride, cl.c0, cl.c1, cl.c2, cl.c3, face_order(D));
    float fr[5], jl[5], jr[5];
    interface_flux<D>(a.flux, tl, tr, fr);
    if constexpr (VISC) {
        // the half-difference the block visc kernel stored in the left
        // element's slot of the face; component 0 as the tile kernels form it
        const FvQuad<float> q = *reinterpret_cast<const FvQuad<float>*>(
            a.fv + (int64_t)blk_elem(a, b, lL) * kFvStride + (2 * D + 1) * 64 + fp * 4);
        const float hd[5] = {0.5f * (0.f - 0.f), q.c[0], q.c[1], q.c[2], q.c[3]};
        face_correction_hd<D>(fr, tl, hd, true, jl);
        face_correction_hd<D>(fr, tr, hd, false, jr);
    } else {
        const float zero[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
        face_correction<D, VISC>(fr, tl, zero, zero, false, true, jl);
        face_correction<D, VISC>(fr, tr, zero, zero, false, false, jr);
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        L.sJi[lL][D][v][fp] = jl[v];
        L.sJi[lR][D][v][fp] = jr[v];
    }
}

// EXT: write the new state's traces of the external faces only (the block
// visc kernel forms the internal ones from its LDS state tiles).
template <bool VISC, bool EXT>
__device__ __forceinline__ void update_block_body(const StageArgs<float, float>& a, const Block& b) {
    __shared__ BlockUpdateLds L;
    const int wave = wave_index(), pt = threadIdx.x & 63;
    const int i = pt & 3, j = (pt >> 2) & 3, k = pt >> 4;
    const int64_t NP = (int64_t)a.nx * a.ny * a.nz * 64;
    const int64_t e = blk_elem(a, b, wave);
    const int64_t gbase = e * 64 + pt;
    const float cf = coef_load<float>((int)threadIdx.x);
    // the element's three external neighbour traces (60 16-byte LDS-DMA
    // chunks) and its state, both read for the last time in this stage:
    // streaming, like the update kernel's stores
    if (pt < 60) {
        const float* src = blk_ext_trace_src(a, b, wave, pt);
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                         (__attribute__((address_space(3))) void*)&L.sJe[wave][0][0][0], 16, 0,
                                         kDmaStream);
    }
    load_state_tile<float, float, kDmaStream>(a.uin, NP, e, L.sU[wave], pt);
    coef_store<float>((int)threadIdx.x, cf);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's LDS copies have landed
    __syncthreads();

    // Volume RHS of the wave's element (update_body's fast f32 path).
    float st[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) st[v] = L.sU[wave][v][pt];
    float vdiv[4] = {0.f, 0.f, 0.f, 0.f};
    if (VISC) {
        const FvQuad<float> src = *reinterpret_cast<const FvQuad<float>*>(a.fv + e * kFvStride + kFvFaces + pt * 4);
#pragma unroll
        for (int q = 0; q < 4; ++q) vdiv[q] = src.c[q];
    }
    float rs[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float F[5];
        euler_flux(d, st, F);
        if (d == 0) {
            const float* Dx = coef_lds<float>().D[i];
#pragma unroll
            for (int v = 0; v < 5; ++v) rs[v] += quad_line(Dx, F[v]) * a.jx;
            continue;
        }
        if (d > 1) wave_sync();  // the y lines have been read
#pragma unroll
        for (int v = 0; v < 5; ++v) L.sF[wave][v][pt] = F[v];
        wave_sync();
        const int idx = d == 1 ? j : k;
        const int base = d == 1 ? i + 16 * k : i + 4 * j;
        const int stride = d == 1 ? 4 : 16;
        const float jac = d == 1 ? a.jy : a.jz;
        const float* Dd = coef_lds<float>().D[idx];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            float dd = 0;
#pragma unroll
            for (int m = 0; m < 4; ++m) dd += Dd[m] * L.sF[wave][v][base + m * stride];
            rs[v] += dd * jac;
        }
    }
    wave_sync();
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        float val = -rs[v];
        if (VISC && v > 0) val += vdiv[v - 1];
        if (v == 1) val += a.bf;
        L.sF[wave][v][pt] = val;
    }

    // Interface fluxes: 6 external passes (4 elements x 16 points each) and
    // 3 internal ones (4 pairs x 16 points), 9 wave passes over 8 waves.
    switch (wave) {
        case 0: blk_ext_pass<0, VISC>(a, b, L, 0, pt); blk_int_pass<2, VISC>(a, b, L, pt); break;
        case 1: blk_ext_pass<0, VISC>(a, b, L, 1, pt); break;
        case 2: blk_ext_pass<1, VISC>(a, b, L, 0, pt); break;
        case 3: blk_ext_pass<1, VISC>(a, b, L, 1, pt); break;
        case 4: blk_ext_pass<2, VISC>(a, b, L, 0, pt); break;
        case 5: blk_ext_pass<2, VISC>(a, b, L, 1, pt); break;
        case 6: blk_int_pass<0, VISC>(a, b, L, pt); break;
        default: blk_int_pass<1, VISC>(a, b, L, pt); break;
    }
    __syncthreads();

    // FR correction and stage update of the wave's element (rk.cu:12-25,
    // 63-82; update_body's fast f32 path).
    const CoefTable<float>& o = coef_lds<float>();
    const float gLi = o.g[i][0] * a.jx, gRi = o.g[i][1] * a.jx;
    const float gLj = o.g[j][0] * a.jy, gRj = o.g[j][1] * a.jy;
    const float gLk = o.g[k][0] * a.jz, gRk = o.g[k][1] * a.jz;
    const int fpx = j + 4 * k, fpy = i + 4 * k, fpz = i + 4 * j;
    const int lx = wave & 1, ly = (wave >> 1) & 1, lz = wave >> 2;
    const float* Jext = &L.sJe[wave][0][0][0];
    const float* Jint = &L.sJi[wave][0][0][0];
    const float* Jx0 = (lx == 0 ? Jext : Jint);
    const float* Jx1 = (lx == 1 ? Jext : Jint);
    const float* Jy0 = (ly == 0 ? Jext : Jint) + 80;
    const float* Jy1 = (ly == 1 ? Jext : Jint) + 80;
    const float* Jz0 = (lz == 0 ? Jext : Jint) + 160;
    const float* Jz1 = (lz == 1 ? Jext : Jint) + 160;
    float uprv[5], ucr[5], r[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        ucr[v] = L.sU[wave][v][pt];
        r[v] = L.sF[wave][v][pt];
    }
    if (a.stage1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) uprv[v] = ucr[v];
    } else {
#pragma unroll
        for (int v = 0; v < 5; ++v) uprv[v] = stream_load(&a.uprev[v * NP + gbase]);
    }
    float rt[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        float cx = -(Jx0[v * 16 + fpx] * gLi + Jx1[v * 16 + fpx] * gRi);
        float cy = -(Jy0[v * 16 + fpy] * gLj + Jy1[v * 16 + fpy] * gRj);
        float cz = -(Jz0[v * 16 + fpz] * gLk + Jz1[v * 16 + fpz] * gRk);
        rt[v] = r[v] + (cx + cy + cz);
    }
    if (a.mode == 1) {  // the stage's RHS into dbg, state unchanged
#pragma unroll
        for (int v = 0; v < 5; ++v) a.dbg[v * NP + gbase] = rt[v];
        return;
    }
    // the blend in the compute or the state type: both float here, so the
    // two differ only in their coefficients, chosen once (scalar selects)
    const float kA = a.accum_state ? a.ua : a.ca;
    const float kB = a.accum_state ? a.ub : a.cb;
    const float kC = a.accum_state ? a.uc : a.cc;
    float unew[5];
    if (a.stage1) {  // uniform branch: each form computes only its own blend
#pragma unroll
        for (int v = 0; v < 5; ++v) unew[v] = rk_blend1(kA, kC, uprv[v], rt[v]);
    } else {
#pragma unroll
        for (int v = 0; v < 5; ++v) unew[v] = rk_blend(kA, kB, kC, uprv[v], ucr[v], rt[v]);
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) stream_store(unew[v], &a.uout[v * NP + gbase]);

    // Face traces of the new state.
#pragma unroll
    for (int v = 0; v < 5; ++v) L.sU[wave][v][pt] = unew[v];
    wave_sync();
    if constexpr (EXT) {
        blk_write_ext_traces(L.sU[wave], a.tr_out + e * kTraceStride, wave, pt);
    } else {
        write_traces_own<float, true>(L.sU[wave], a.tr_out + e * kTraceStride, pt, VISC);
    }
}

// ---------------------------------------------------------------------------
// The visc kernel on the same 2x2x2 blocks: the BR1 jumps of the internal
// faces come from the block's LDS state tiles (one task per internal face
// point forms both sides' jumps), those of the external faces from the
// prefetched neighbour traces; gradient, volume viscous divergence and the
// face viscous fluxes are visc_body's, element by element.  LDS: the jumps
// [le][ext d0..d2, int d0..d2][5][16] (the external part first holds the
// prefetched traces) share one buffer with the gradient tile [le][15][64]
// (written after a barrier), and the state tiles.
// ---------------------------------------------------------------------------
struct BlockViscLds {
    alignas(16) float buf[kBlk * 15 * 64];
    BlockTiles sU;
    // 480 jumps per element, 496 apart (the tiles' bank offset again)
    __device__ float* jumps(int le) { return buf + le * 496; }
    __device__ float (*sG())[15][64] { return reinterpret_cast<float(*)[15][64]>(buf); }
};

// BR1 jump at the external face of direction D of elements 4h .. 4h+3 (as br1_jumps).
template <int D>
__device__ __forceinline__ void blk_br1_ext(const StageArgs<float, float>& a, const Block& b, BlockViscLds& L, int h,
                                            int lane) {
    const int le = 4 * h + (lane >> 4), fp = lane & 15;
    const int side = blk_bit(le, D);
    int base, stride;
    face_line(D, fp, base, stride);
    const Coef4<float> c = side_coef<float>(side);
    const bool wall = D == 1 && blk_ext_nb(a, b, le, D) < 0;
    float* J = L.jumps(le) + D * 80;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        float own = trace4(L.sU[le][v], base, stride, c.c0, c.c1, c.c2, c.c3, br1_order(D));
        const float sign = (v > 0 && v < 4) ? -1.0f : 1.0f;
        const float far = J[v * 16 + fp];
        float nbv = !wall ? far : own * sign;
        J[v * 16 + fp] = 0.5f * (nbv + own) - own;
    }
}

// BR1 jumps of both sides of the internal faces of direction D (pairs as in
// blk_int_pass); each side's jump as br1_jumps forms it with the other
// side's trace as the neighbour value.
template <int D>
__device__ __forceinline__ void blk_br1_int(BlockViscLds& L, int lane) {
    const int p = lane >> 4, fp = lane & 15;
    const int lL = D == 0 ? 2 * p : (D == 1 ? ((p & 1) | ((p >> 1) << 2)) : p);
    const int lR = lL + (1 << D);
    int base, stride;
    face_line(D, fp, base, stride);
    const Coef4<float> cr = side_coef<float>(true), cl = side_coef<float>(false);
    float* JL = L.jumps(lL) + (3 + D) * 80;
    float* JR = L.jumps(lR) + (3 + D) * 80;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const float tl = trace4(L.sU[lL][v], base, stride, cr.c0, cr.c1, cr.c2, cr.c3, br1_order(D));
        const float tr = trace4(L.sU[lR][v], base, stride, cl.c0, cl.c1, cl.c2, cl.c3, br1_order(D));
        JL[v * 16 + fp] = 0.5f * (tr + tl) - tl;
        JR[v * 16 + fp] = 0.5f * (tl + tr) - tr;
    }
}

// Viscous normal fluxes at the faces of direction D of the q-th element pair
// along D (left lL, right lL + 2^D): lane = (element of the pair, side,
// face point).  External faces store their flux as the tile kernels do; at
// the pair's interior face the left element's lane (16..31) takes the right
// element's flux from lane + 16 (32..47) and stores the half-difference
// (F_vL - F_vR) / 2 in the left element's slot, which is all the update
// kernel reads there.
template <int D>
__device__ __forceinline__ void blk_fv_pass(const StageArgs<float, float>& a, const Block& b, BlockViscLds& L, int q,
                                            int pt) {
    const int lL = D == 0 ? 2 * q : (D == 1 ? ((q & 1) | ((q >> 1) << 2)) : q);
    const int second = pt >> 5, side = (pt >> 4) & 1, fp = pt & 15;
    const int le = second ? lL + (1 << D) : lL;
    const int64_t e = blk_elem(a, b, le);
    const FvQuad<float> f = fv_value<D>(a, L.sU[le], L.sG()[le], e, side, fp);
    FvQuad<float> g;
    const int from = (pt ^ 48) << 2;  // lane 16 + fp <-> 32 + fp (byte address of the source lane)
#pragma unroll
    for (int c = 0; c < 4; ++c) g.c[c] = __int_as_float(__builtin_amdgcn_ds_bpermute(from, __float_as_int(f.c[c])));
    FvQuad<float>* dst = reinterpret_cast<FvQuad<float>*>(a.fv + e * kFvStride + (2 * D + side) * 64 + fp * 4);
    if (second == side) {
        *dst = f;  // external face
    } else if (!second) {
        FvQuad<float> h;
#pragma unroll
        for (int c = 0; c < 4; ++c) h.c[c] = 0.5f * (f.c[c] - g.c[c]);
        *dst = h;
    }
}

__device__ __forceinline__ void visc_block_body(const StageArgs<float, float>& a, const Block& b) {
    __shared__ BlockViscLds L;
    const int wave = wave_index(), pt = threadIdx.x & 63;
    const int i = pt & 3, j = (pt >> 2) & 3, k = pt >> 4;
    const int64_t NP = (int64_t)a.nx * a.ny * a.nz * 64;
    const int64_t e = blk_elem(a, b, wave);
    const float cf = coef_load<float>((int)threadIdx.x);
    if (pt < 60) {  // the element's three external neighbour traces: 60 16-byte LDS-DMA chunks
        const float* src = blk_ext_trace_src(a, b, wave, pt);
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                         (__attribute__((address_space(3))) void*)L.jumps(wave), 16, 0, 0);
    }
    load_state_tile(a.uin, NP, e, L.sU[wave], pt);
    coef_store<float>((int)threadIdx.x, cf);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    switch (wave) {
        case 0: blk_br1_ext<0>(a, b, L, 0, pt); blk_br1_int<2>(L, pt); break;
        case 1: blk_br1_ext<0>(a, b, L, 1, pt); break;
        case 2: blk_br1_ext<1>(a, b, L, 0, pt); break;
        case 3: blk_br1_ext<1>(a, b, L, 1, pt); break;
        case 4: blk_br1_ext<2>(a, b, L, 0, pt); break;
        case 5: blk_br1_ext<2>(a, b, L, 1, pt); break;
        case 6: blk_br1_int<0>(L, pt); break;
        default: blk_br1_int<1>(L, pt); break;
    }
    __syncthreads();
    // the wave's element: face f = 2d + s has its jumps in the external slot
    // d when s is the element's coordinate bit d, else in the internal slot
    const float* Jw = L.jumps(wave);
    const int lx = wave & 1, ly = (wave >> 1) & 1, lz = wave >> 2;
    float g[15];
    lane_gradient_faces(a, L.sU[wave], Jw + (lx == 0 ? 0 : 240), Jw + (lx == 1 ? 0 : 240), Jw + (ly == 0 ? 80 : 320),
                        Jw + (ly == 1 ? 80 : 320), Jw + (lz == 0 ? 160 : 400), Jw + (lz == 1 ? 160 : 400), i, j, k, g);
    // staging of the volume viscous-flux lines in the element's own jumps,
    // which only this wave has read
    visc_divergence(a, e, L.sU[wave], g, L.jumps(wave), i, j, k, pt);
    __syncthreads();  // the gradient tile overwrites jumps other waves may still read
#pragma unroll
    for (int q = 0; q < 15; ++q) L.sG()[wave][q][swz64(pt)] = g[q];
    __syncthreads();
    // face viscous fluxes: pass (D, q) = the q-th pair of elements along D,
    // both faces of direction D (waves 0..3: x then z, waves 4..7: y)
    if (wave < 4) {
        blk_fv_pass<0>(a, b, L, wave, pt);
        blk_fv_pass<2>(a, b, L, wave, pt);
    } else {
        blk_fv_pass<1>(a, b, L, wave - 4, pt);
    }
}

// One box: the update kernel walks the blocks backwards, the visc kernel
// forwards (make_block).
template <bool VISC, bool EXT>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8))) void od_update_block_kernel(
    StageArgs<float, float> a) {
    update_block_body<VISC, EXT>(a, make_block<true>(a));
}
__global__ __launch_bounds__(512) void od_visc_block_kernel(StageArgs<float, float> a) {
    visc_block_body(a, make_block<false>(a));
}
#endif

// ---------------------------------------------------------------------------
// Ensembles: the tiles of every box of a batch in one launch per phase.  A
// work unit is `PAIRS` consecutive tiles (one 128-thread tile each); work-
// groups walk the units in a fixed order, so every unit is done exactly once
// and the result of each element does not depend on the schedule.
// ---------------------------------------------------------------------------

// Unit handled by this work-group at iteration `it`, or -1 when it is done.
// order 0: round robin over all units.  order 1: work-groups are dispatched
// round robin over the 8 XCDs, so XCD x = blockIdx.x % 8 takes the x-th
// contiguous eighth of the units (its neighbours' traces stay in its L2).
__device__ __forceinline__ int sched_unit(int order, int n_units, int it) {
    if (order == 0) {
        const int u = (int)blockIdx.x + it * (int)gridDim.x;
        return u < n_units ? u : -1;
    }
    const int m = (int)gridDim.x >> 3, x = (int)blockIdx.x & 7, s = (int)blockIdx.x >> 3;
    const int chunk = (n_units + 7) >> 3;
    const int r0 = x * chunk, r1 = min(n_units, r0 + chunk);
    const int u = r0 + s + it * m;
    return u < r1 ? u : -1;
}

// Wave-uniform copies (SGPRs) of values loaded from the box table: the
// loads sit after stores of earlier units, so the compiler cannot use
// scalar loads by itself.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uni(float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); }
__device__ __forceinline__ int64_t uni(int64_t x) {
    const uint64_t u = (uint64_t)x;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double uni(double x) { return __longlong_as_double(uni((int64_t)__double_as_longlong(x))); }
__device__ __forceinline__ Elem uni(const Elem& e) {
    return Elem{uni(e.xl), uni(e.xr), uni(e.yl), uni(e.yr), uni(e.zl), uni(e.zr)};
}

// Stage arguments of one box of a batch: the batch's arguments with the
// box's dimensions, Jacobians, stage coefficients and buffer offsets.
template <class UT, class CT>
__device__ __forceinline__ void batch_box_args(const StageArgs<UT, CT>& a, const BoxDev<UT, CT>& bx, int stage,
                                               StageArgs<UT, CT>& ab) {
    ab = a;
    ab.nx = uni(bx.nx);
    ab.ny = uni(bx.ny);
    ab.nz = uni(bx.nz);
    ab.jx = uni(bx.jx);
    ab.jy = uni(bx.jy);
    ab.jz = uni(bx.jz);
    ab.cc = uni(bx.cc[stage]);
    ab.uc = uni(bx.uc[stage]);
    const int64_t e_off = uni(bx.e_off);
    const int64_t so = e_off * 320;  // box state [5][E_box][64] starts at 5 * 64 * e_off
    ab.uin = a.uin + so;
    ab.uprev = a.uprev + so;
    ab.uout = a.uout + so;
    ab.dbg = a.dbg ? a.dbg + so : a.dbg;
    ab.tr_in = a.tr_in + e_off * kTraceStride;
    ab.tr_out = a.tr_out + e_off * kTraceStride;
    ab.fv = a.fv ? a.fv + e_off * kFvStride : a.fv;
    ab.mu_t = a.mu_t ? a.mu_t + e_off : a.mu_t;
}

// Tile lt of the shell of a box whose core runs on 2x2x2 blocks: the
// elements with x = nx - 1 (odd nx), then y = ny - 1 (odd ny), then z = nz - 1
// (odd nz), each element in the first of these slabs it lies in.  The x slab
// pairs its elements along y, the others along x.  A tile's elements and
// neighbours come from make_tile_at as for any tile, so each face is formed
// as in a box run on tiles alone.
template <class UT, class CT>
__device__ __forceinline__ Tile shell_tile(const StageArgs<UT, CT>& a, int lt) {
    const int hx = a.nx >> 1, hy = (a.ny + 1) >> 1;
    const int nA = (a.nx & 1) ? hy * a.nz : 0;
    const int nB = (a.ny & 1) ? hx * a.nz : 0;
    if (lt < nA) return make_tile_at(a, 1, a.nx - 1, lt % hy, lt / hy);
    lt -= nA;
    if (lt < nB) return make_tile_at(a, 0, lt % hx, a.ny - 1, lt / hx);
    lt -= nB;
    return make_tile_at(a, 0, lt % hx, lt / hx, a.nz - 1);
}

// Stage arguments and tile of batch tile t (t >= n_tiles: an empty tile).
template <class UT, class CT>
__device__ __forceinline__ void batch_tile(const StageArgs<UT, CT>& a, const BatchArgs<UT, CT>& b, int t,
                                           StageArgs<UT, CT>& ab, Tile& tl) {
    const bool ok = t < b.n_tiles;
    const int tt = ok ? t : 0;
    int lo = 0, hi = b.n_boxes - 1;
    if (b.n_boxes <= kInlineBoxes) {  // kernel-argument copy: scalar loads, no memory latency chain
        while (lo < hi) {  // last box whose first tile is <= tt
            const int mid = (lo + hi + 1) >> 1;
            if (b.tile_off[mid] <= tt)
                lo = mid;
            else
                hi = mid - 1;
        }
    } else {
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (uni(b.boxes[mid].tile_off) <= tt)
                lo = mid;
            else
                hi = mid - 1;
        }
    }
    const BoxDev<UT, CT>& bx = b.boxes[lo];
    batch_box_args(a, bx, b.stage, ab);
    const int tile_off = uni(bx.tile_off);
    const int lt = tt - tile_off;
    if (uni(bx.shell)) {
        tl = shell_tile(ab, lt);
    } else {
        const int d0 = b.orient == 0 ? (ab.nx + kEpb - 1) / kEpb : ab.nx;
        const int d1 = b.orient == 1 ? (ab.ny + kEpb - 1) / kEpb : ab.ny;
        const int t0 = uni(lt % d0), r = uni(lt / d0);
        tl = make_tile_at(ab, b.orient, t0, uni(r % d1), uni(r / d1));
    }
    tl.e0 = uni(tl.e0);
    tl.e1 = uni(tl.e1);
    tl.el0 = uni(tl.el0);
    tl.el1 = uni(tl.el1);
    tl.walls = uni(tl.walls);
    tl.v0 = ok;
    tl.v1 = ok && tl.v1;
}

#if OD_FAST
// Ensembles on 2x2x2-element blocks: the core of every box (the box less the
// last layer along each odd dimension; the whole box when all are even) runs
// the block kernels, one launch per phase for all boxes, and the shell the
// tile kernels (shell_tile).  The blocks' neighbour map (blk_ext_nb) needs
// no change: a core block's right neighbour is the next element unless the
// block ends the box, and its left one across x = 0 (y = 0, z = 0) is the
// last element of the box, a shell element when that dimension is odd.
// Batch block t -> its box (the last whose first block is
// <= t; a box without blocks shares the next box's first block, so it is
// never the last such box) -> the box's stage arguments and the block within
// the box, x fastest.  The update kernel walks the blocks backwards, as the
// single-box one does.
template <bool REVERSE>
__device__ __forceinline__ void batch_block(const StageArgs<float, float>& a, const BlockBatch& bb,
                                            StageArgs<float, float>& ab, Block& b) {
    const int t = REVERSE ? bb.n_blocks - 1 - (int)blockIdx.x : (int)blockIdx.x;
    int lo = 0, hi = bb.n_boxes - 1, off;
    if (bb.n_boxes <= kInlineBoxes) {  // kernel-argument copy: scalar loads, no memory latency chain
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (bb.blk_off[mid] <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        off = bb.blk_off[lo];
    } else {
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (uni(bb.boxes[mid].blk_off) <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        off = uni(bb.boxes[lo].blk_off);
    }
    batch_box_args(a, bb.boxes[lo], bb.stage, ab);
    const int lb = t - off;
    const int gx = ab.nx >> 1, gy = ab.ny >> 1;
    b.ex0 = 2 * (lb % gx);
    b.ey0 = 2 * ((lb / gx) % gy);
    b.ez0 = 2 * (lb / (gx * gy));
    b.e0 = b.ex0 + ab.nx * (b.ey0 + ab.ny * b.ez0);
}

template <bool VISC>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8))) void od_batch_update_block_kernel(
    StageArgs<float, float> a, BlockBatch bb) {
    StageArgs<float, float> ab;
    Block b;
    batch_block<true>(a, bb, ab, b);
    update_block_body<VISC, true>(ab, b);
}

__global__ __launch_bounds__(512) void od_batch_visc_block_kernel(StageArgs<float, float> a, BlockBatch bb) {
    StageArgs<float, float> ab;
    Block b;
    batch_block<false>(a, bb, ab, b);
    visc_block_body(ab, b);
}
#endif

// LOOP = false: one unit per work-group (grid >= units, order 0); the
// persistent loop is a separate instantiation, so that its loop-carried
// state does not raise the register count of the single-unit kernel.
template <class UT, class CT, int PAIRS, bool LOOP>
__global__ __launch_bounds__(128 * PAIRS) void od_batch_visc_kernel(StageArgs<UT, CT> a, BatchArgs<UT, CT> b) {
    __shared__ ViscLds<CT> L[PAIRS];
    const int pair = PAIRS == 1 ? 0 : (int)(threadIdx.x >> 7);
    const int wave = wave_index() & 1, pt = threadIdx.x & 63;
    if constexpr (!LOOP) {
        StageArgs<UT, CT> ab;
        Tile tl;
        batch_tile(a, b, (int)blockIdx.x * PAIRS + pair, ab, tl);
        visc_body(ab, tl, L[pair], wave, pt);
        return;
    }
    for (int it = 0;; ++it) {
        const int unit = sched_unit(b.order, b.n_units, it);
        if (unit < 0) break;
        // A fresh copy of the lane index each unit keeps the compiler from
        // hoisting lane-dependent values out of the loop (they would stay
        // live in registers across the whole kernel).
        int wv = wave, pv = pt;
        asm volatile("" : "+v"(wv), "+v"(pv));
        StageArgs<UT, CT> ab;
        Tile tl;
        batch_tile(a, b, unit * PAIRS + pair, ab, tl);
        visc_body(ab, tl, L[pair], wv, pv);
        __syncthreads();  // LDS is reused by the next unit
    }
}

template <class UT, class CT, bool VISC, int PAIRS, bool LOOP>
__global__ __launch_bounds__(128 * PAIRS) __attribute__((amdgpu_waves_per_eu(LOOP ? kMinWaves<CT> : kMinWavesUpdate<UT, CT>))) void od_batch_update_kernel(
    StageArgs<UT, CT> a, BatchArgs<UT, CT> b) {
    __shared__ UpdateLds<CT, VISC> L[PAIRS];
    const int pair = PAIRS == 1 ? 0 : (int)(threadIdx.x >> 7);
    const int wave = wave_index() & 1, pt = threadIdx.x & 63;
    if constexpr (!LOOP) {  // units in reverse order, as the single-box update kernels
        StageArgs<UT, CT> ab;
        Tile tl;
        batch_tile(a, b, ((int)gridDim.x - 1 - (int)blockIdx.x) * PAIRS + pair, ab, tl);
        update_body<UT, CT, VISC>(ab, tl, L[pair], wave, pt);
        return;
    }
    for (int it = 0;; ++it) {
        const int unit = sched_unit(b.order, b.n_units, it);
        if (unit < 0) break;
        // A fresh copy of the lane index each unit keeps the compiler from
        // hoisting lane-dependent values out of the loop (they would stay
        // live in registers across the whole kernel).
        int wv = wave, pv = pt;
        asm volatile("" : "+v"(wv), "+v"(pv));
        StageArgs<UT, CT> ab;
        Tile tl;
        batch_tile(a, b, unit * PAIRS + pair, ab, tl);
        update_body<UT, CT, VISC>(ab, tl, L[pair], wv, pv);
        __syncthreads();  // LDS is reused by the next unit
    }
}

// Conservation sums: per-element GL-weighted sums (sequential over points),
// then one block reduces them in a fixed order -> deterministic.
struct W3 {
    double w[64];
};

template <class UT>
__global__ __launch_bounds__(256) void od_sums_partial_kernel(const UT* u, double* partial, int64_t E, W3 w3) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int64_t NP = E * 64;
    for (int v = 0; v < 5; ++v) {
        double s = 0.0;
        for (int pt = 0; pt < 64; ++pt) s += w3.w[pt] * (double)u[v * NP + e * 64 + pt];
        partial[v * E + e] = s;
    }
}

__global__ __launch_bounds__(256) void od_sums_final_kernel(const double* partial, double* out, int64_t E, double jac) {
    __shared__ double red[256];
    const int t = threadIdx.x;
    for (int v = 0; v < 5; ++v) {
        double s = 0.0;
        for (int64_t e = t; e < E; e += 256) s += partial[v * E + e];
        red[t] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (t < w) red[t] = red[t] + red[t + w];
            __syncthreads();
        }
        if (t == 0) out[v] = red[0] * jac;
        __syncthreads();
    }
}

inline int check_launch() { return hipGetLastError() == hipSuccess ? 0 : 1; }

template <class UT, class CT>
dim3 grid_of(const StageArgs<UT, CT>& a) {
    return dim3((unsigned)((a.nx + kEpb - 1) / kEpb), (unsigned)a.ny, (unsigned)a.nz);
}

}  // namespace

template <class UT, class CT>
int launch_trace(const StageArgs<UT, CT>& a, int64_t n_elem, void* stream) {
    hipLaunchKernelGGL((od_trace_kernel<UT, CT>), grid_of(a), dim3(64 * kEpb), 0, (hipStream_t)stream, a);
    return check_launch();
}

template <class UT, class CT>
int launch_visc(const StageArgs<UT, CT>& a, int64_t n_elem, void* stream) {
    hipLaunchKernelGGL((od_visc_kernel<UT, CT>), grid_of(a), dim3(64 * kEpb), 0, (hipStream_t)stream, a);
    return check_launch();
}

template <class UT, class CT, bool VISC>
int launch_update(const StageArgs<UT, CT>& a, int64_t n_elem, void* stream) {
    hipLaunchKernelGGL((od_update_kernel<UT, CT, VISC>), grid_of(a), dim3(64 * kEpb), 0, (hipStream_t)stream, a);
    return check_launch();
}

#if OD_FAST
inline dim3 block_grid(const StageArgs<float, float>& a) {
    return dim3((unsigned)(a.nx / 2), (unsigned)(a.ny / 2), (unsigned)(a.nz / 2));
}

template <bool VISC>
int launch_update_block(const StageArgs<float, float>& a, bool ext_traces, void* stream) {
    if (a.nx % 2 || a.ny % 2 || a.nz % 2) return 1;  // blocks of 2x2x2 elements tile the box exactly
    if (ext_traces)
        hipLaunchKernelGGL((od_update_block_kernel<VISC, true>), block_grid(a), dim3(64 * kBlk), 0,
                           (hipStream_t)stream, a);
    else
        hipLaunchKernelGGL((od_update_block_kernel<VISC, false>), block_grid(a), dim3(64 * kBlk), 0,
                           (hipStream_t)stream, a);
    return check_launch();
}
template int launch_update_block<true>(const StageArgs<float, float>&, bool, void*);
template int launch_update_block<false>(const StageArgs<float, float>&, bool, void*);

int launch_visc_block(const StageArgs<float, float>& a, void* stream) {
    if (a.nx % 2 || a.ny % 2 || a.nz % 2) return 1;
    hipLaunchKernelGGL(od_visc_block_kernel, block_grid(a), dim3(64 * kBlk), 0, (hipStream_t)stream, a);
    return check_launch();
}
template <bool VISC>
int launch_batch_update_block(const StageArgs<float, float>& a, const BlockBatch& bb, void* stream) {
    if (bb.n_blocks <= 0) return 0;
    hipLaunchKernelGGL((od_batch_update_block_kernel<VISC>), dim3((unsigned)bb.n_blocks), dim3(64 * kBlk), 0,
                       (hipStream_t)stream, a, bb);
    return check_launch();
}
int launch_batch_visc_block(const StageArgs<float, float>& a, const BlockBatch& bb, void* stream) {
    if (bb.n_blocks <= 0) return 0;
    hipLaunchKernelGGL(od_batch_visc_block_kernel, dim3((unsigned)bb.n_blocks), dim3(64 * kBlk), 0,
                       (hipStream_t)stream, a, bb);
    return check_launch();
}
template int launch_batch_update_block<true>(const StageArgs<float, float>&, const BlockBatch&, void*);
template int launch_batch_update_block<false>(const StageArgs<float, float>&, const BlockBatch&, void*);
#endif

template <class UT, class CT>
int launch_batch_visc(const StageArgs<UT, CT>& a, const BatchArgs<UT, CT>& b, int pairs, int groups, void* stream) {
    const bool single = groups >= b.n_units && b.order == 0;  // one unit per work-group
    const hipStream_t st = (hipStream_t)stream;
    if (pairs == 2) {
        if (single)
            hipLaunchKernelGGL((od_batch_visc_kernel<UT, CT, 2, false>), dim3(groups), dim3(256), 0, st, a, b);
        else
            hipLaunchKernelGGL((od_batch_visc_kernel<UT, CT, 2, true>), dim3(groups), dim3(256), 0, st, a, b);
    } else {
        if (single)
            hipLaunchKernelGGL((od_batch_visc_kernel<UT, CT, 1, false>), dim3(groups), dim3(128), 0, st, a, b);
        else
            hipLaunchKernelGGL((od_batch_visc_kernel<UT, CT, 1, true>), dim3(groups), dim3(128), 0, st, a, b);
    }
    return check_launch();
}

template <class UT, class CT, bool VISC>
int launch_batch_update(const StageArgs<UT, CT>& a, const BatchArgs<UT, CT>& b, int pairs, int groups, void* stream) {
    const bool single = groups >= b.n_units && b.order == 0;  // one unit per work-group
    const hipStream_t st = (hipStream_t)stream;
    if (pairs == 2) {
        if (single)
            hipLaunchKernelGGL((od_batch_update_kernel<UT, CT, VISC, 2, false>), dim3(groups), dim3(256), 0, st, a, b);
        else
            hipLaunchKernelGGL((od_batch_update_kernel<UT, CT, VISC, 2, true>), dim3(groups), dim3(256), 0, st, a, b);
    } else {
        if (single)
            hipLaunchKernelGGL((od_batch_update_kernel<UT, CT, VISC, 1, false>), dim3(groups), dim3(128), 0, st, a, b);
        else
            hipLaunchKernelGGL((od_batch_update_kernel<UT, CT, VISC, 1, true>), dim3(groups), dim3(128), 0, st, a, b);
    }
    return check_launch();
}

template <class UT, class CT>
int batch_occupancy(int kind, int pairs, bool visc, int* groups_per_cu) {
    int n = 0;
    hipError_t e;
    // the persistent (loop) kernels: the grid this sizes is a persistent one
    if (kind == 0)
        e = pairs == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, od_batch_visc_kernel<UT, CT, 2, true>, 256, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, od_batch_visc_kernel<UT, CT, 1, true>, 128, 0);
    else if (visc)
        e = pairs == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                             &n, od_batch_update_kernel<UT, CT, true, 2, true>, 256, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                             &n, od_batch_update_kernel<UT, CT, true, 1, true>, 128, 0);
    else
        e = pairs == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                             &n, od_batch_update_kernel<UT, CT, false, 2, true>, 256, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                             &n, od_batch_update_kernel<UT, CT, false, 1, true>, 128, 0);
    *groups_per_cu = n;
    return e == hipSuccess ? 0 : 1;
}

template <class UT>
int launch_sums(const UT* u, double* partial, double* out, int64_t n_elem, const double w3[64], double jac,
                void* stream) {
    W3 w;
    for (int q = 0; q < 64; ++q) w.w[q] = w3[q];
    unsigned blocks = (unsigned)((n_elem + 255) / 256);
    hipLaunchKernelGGL((od_sums_partial_kernel<UT>), dim3(blocks), dim3(256), 0, (hipStream_t)stream, u, partial,
                       n_elem, w);
    if (check_launch()) return 1;
    hipLaunchKernelGGL(od_sums_final_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, partial, out, n_elem, jac);
    return check_launch();
}

#if !OD_FAST
const char* kernel_name(int kind) {
    switch (kind) {
        case 0: return "od_visc_kernel";
        case 1: return "od_update_kernel";
        case 2: return "od_trace_kernel";
        default: return "?";
    }
}
#endif

#define OD_INST(UT, CT)                                                                                      \
    template int launch_trace<UT, CT>(const StageArgs<UT, CT>&, int64_t, void*);                             \
    template int launch_visc<UT, CT>(const StageArgs<UT, CT>&, int64_t, void*);                              \
    template int launch_update<UT, CT, true>(const StageArgs<UT, CT>&, int64_t, void*);                      \
    template int launch_update<UT, CT, false>(const StageArgs<UT, CT>&, int64_t, void*);                     \
    template int launch_batch_visc<UT, CT>(const StageArgs<UT, CT>&, const BatchArgs<UT, CT>&, int, int, void*); \
    template int launch_batch_update<UT, CT, true>(const StageArgs<UT, CT>&, const BatchArgs<UT, CT>&, int, int, \
                                                   void*);                                                   \
    template int launch_batch_update<UT, CT, false>(const StageArgs<UT, CT>&, const BatchArgs<UT, CT>&, int, int, \
                                                    void*);                                                  \
    template int batch_occupancy<UT, CT>(int, int, bool, int*);
OD_INST(float, float)
OD_INST(double, float)
#if !OD_FAST
OD_INST(double, double)
OD_INST(float, double)
#endif
#undef OD_INST
#if !OD_FAST
template int launch_sums<float>(const float*, double*, double*, int64_t, const double*, double, void*);
template int launch_sums<double>(const double*, double*, double*, int64_t, const double*, double, void*);
#endif

int upload_ops(const Ops<float>& f32, const Ops<double>& f64) {
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_ops_f32), &f32, sizeof(f32)) != hipSuccess) return 1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_ops_f64), &f64, sizeof(f64)) != hipSuccess) return 1;
    return 0;
}

#if OD_FAST
}  // namespace fastmath
#endif
}  // namespace od
