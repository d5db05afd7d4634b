// CPU oracle: op-for-op restatement of drunkseal/opendynamo's FR P=3 step.
// TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with -ffp-contract=off.
//
// Layouts follow the reference exactly:
//   state / RHS: [var 5][elem E][pt 64], elem = ex + nx*ey + nx*ny*ez,
//                pt = i + 4j + 16k                       (rhs.cu:352-361)
//   j arrays:    [var][elem][fp 16], fp_x = j+4k, fp_y = i+4k, fp_z = i+4j
//   slices:      [elem][fp][12], components 3*var+dir ascending with a
//                per-direction skip set, slot 11 = 0      (rhs.cu:116-217)
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// FR operators, fr_math.rs:3-142 (f64, Rust evaluation order).
// ---------------------------------------------------------------------------
double legendre_poly(int p, double x) {  // fr_math.rs:3-21
    if (p == 0) return 1.0;
    if (p == 1) return x;
    double prev2 = 1.0, prev = x, value = 0.0;
    for (int k = 2; k <= p; ++k) {
        double kf = (double)k;
        value = ((2.0 * kf - 1.0) * x * prev - (kf - 1.0) * prev2) / kf;
        prev2 = prev;
        prev = value;
    }
    return value;
}

double legendre_poly_deriv(int p, double x) {  // fr_math.rs:23-29
    if (p == 0) return 0.0;
    double pf = (double)p;
    return pf / (x * x - 1.0) * (x * legendre_poly(p, x) - legendre_poly(p - 1, x));
}

struct FrOps {
    double points[4], d[16], el[4], er[4], gl[4], gr[4], w[4];
};

FrOps make_ops() {
    FrOps o{};
    const int P = 3, n = 4;
    // gauss_legendre_points, fr_math.rs:31-53: Newton from cos(pi(i+.75)/(n+.5)).
    double roots[4];
    for (int i = 0; i < n; ++i) {
        double x = std::cos(M_PI * ((double)i + 0.75) / ((double)n + 0.5));
        for (;;) {
            double lv = legendre_poly(n, x);
            double ld = legendre_poly_deriv(n, x);
            double step = lv / ld;
            x -= step;
            if (std::fabs(step) < 1e-14) break;
        }
        roots[i] = x;
    }
    std::sort(roots, roots + n);
    for (int i = 0; i < n; ++i) o.points[i] = roots[i];
    // lagrange_deriv_matrix, fr_math.rs:65-96.
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            if (i != j) {
                double prod = 1.0;
                for (int k = 0; k < n; ++k)
                    if (k != i && k != j) prod *= o.points[i] - o.points[k];
                double den = 1.0;
                for (int k = 0; k < n; ++k)
                    if (k != j) den *= o.points[j] - o.points[k];
                o.d[i * n + j] = prod / den;
            } else {
                double sum = 0.0;
                for (int k = 0; k < n; ++k)
                    if (k != i) sum += 1.0 / (o.points[i] - o.points[k]);
                o.d[i * n + i] = sum;
            }
        }
    // lagrange_poly at -1 / +1, fr_math.rs:55-63, 112-117.
    for (int i = 0; i < n; ++i) {
        double l = 1.0, r = 1.0;
        for (int j = 0; j < n; ++j)
            if (i != j) {
                l *= (-1.0 - o.points[j]) / (o.points[i] - o.points[j]);
                r *= (1.0 - o.points[j]) / (o.points[i] - o.points[j]);
            }
        o.el[i] = l;
        o.er[i] = r;
    }
    // Radau correction derivatives, fr_math.rs:119-132 (P odd -> sign -1).
    const double sign = (P % 2 == 0) ? 1.0 : -1.0;
    for (int i = 0; i < n; ++i) {
        double dp = legendre_poly_deriv(P, o.points[i]);
        double dp1 = legendre_poly_deriv(P + 1, o.points[i]);
        o.gr[i] = 0.5 * (dp + dp1);
        o.gl[i] = 0.5 * sign * (dp - dp1);
    }
    // Gauss-Legendre weights 2 / ((1 - x^2) P_n'(x)^2) (textbook; used only by
    // the conservation sums, which have no reference counterpart).
    for (int i = 0; i < n; ++i) {
        double x = o.points[i];
        double dl = legendre_poly_deriv(n, x);
        o.w[i] = 2.0 / ((1.0 - x * x) * dl * dl);
    }
    return o;
}

const FrOps& ops() {
    static const FrOps o = make_ops();
    return o;
}

// ---------------------------------------------------------------------------
// Point physics, device_functions.cuh:1-257, in the compute type C.
// `mu` replaces the reference's compile-time MU (constants.cuh:42).
// ---------------------------------------------------------------------------
template <class C>
struct Phys {
    static C gam() { return (C)1.4; }
    static C pr() { return (C)0.72; }

    static C pressure(C rho, C ru, C rv, C rw, C e) {  // :1-5
        C ke = ((C)0.5) * (ru * ru + rv * rv + rw * rw) / rho;
        C p = (gam() - ((C)1.0)) * (e - ke);
        return std::fmax(((C)1e-7), p);
    }

    // dir 0/1/2 = x/y/z, :7-21
    static void euler(int dir, const C* s, C* f) {
        C rho = s[0], ru = s[1], rv = s[2], rw = s[3], e = s[4];
        C p = pressure(rho, ru, rv, rw, e);
        if (dir == 0) {
            C u = ru / rho;
            f[0] = ru; f[1] = ru * u + p; f[2] = rv * u; f[3] = rw * u; f[4] = u * (e + p);
        } else if (dir == 1) {
            C v = rv / rho;
            f[0] = rv; f[1] = ru * v; f[2] = rv * v + p; f[3] = rw * v; f[4] = v * (e + p);
        } else {
            C w = rw / rho;
            f[0] = rw; f[1] = ru * w; f[2] = rv * w; f[3] = rw * w + p; f[4] = w * (e + p);
        }
    }

    // :23-102.  g[3*var + dir] are the conservative-variable gradients.
    static void viscous(int dir, const C* s, const C* g, C mu, C* f) {
        C rho = s[0], ru = s[1], rv = s[2], rw = s[3], e = s[4];
        C u = ru / rho, v = rv / rho, w = rw / rho;
        const C two = (C)2.0, three = (C)3.0, one = (C)1.0, half = (C)0.5;
        if (dir == 0) {
            C du_dx = (g[1 * 3 + 0] - u * g[0 * 3 + 0]) / rho;
            C dv_dx = (g[2 * 3 + 0] - v * g[0 * 3 + 0]) / rho;
            C dw_dx = (g[3 * 3 + 0] - w * g[0 * 3 + 0]) / rho;
            C dv_dy = (g[2 * 3 + 1] - v * g[0 * 3 + 1]) / rho;
            C dw_dz = (g[3 * 3 + 2] - w * g[0 * 3 + 2]) / rho;
            C du_dy = (g[1 * 3 + 1] - u * g[0 * 3 + 1]) / rho;
            C du_dz = (g[1 * 3 + 2] - u * g[0 * 3 + 2]) / rho;
            C div = du_dx + dv_dy + dw_dz;
            C txx = mu * (two * du_dx - two / three * div);
            C txy = mu * (du_dy + dv_dx);
            C txz = mu * (du_dz + dw_dx);
            C ke = half * (ru * ru + rv * rv + rw * rw) / rho;
            C p = (gam() - one) * (e - ke);
            C dke = (ru * g[1 * 3 + 0] + rv * g[2 * 3 + 0] + rw * g[3 * 3 + 0]) / rho - ke * g[0 * 3 + 0] / rho;
            C dp = (gam() - one) * (g[4 * 3 + 0] - dke);
            C coeff = mu * gam() / (pr() * (gam() - one));
            C q = -coeff * (dp / rho - p * g[0 * 3 + 0] / (rho * rho));
            f[0] = (C)0.0; f[1] = txx; f[2] = txy; f[3] = txz;
            f[4] = u * txx + v * txy + w * txz - q;
        } else if (dir == 1) {
            C du_dx = (g[1 * 3 + 0] - u * g[0 * 3 + 0]) / rho;
            C dv_dx = (g[2 * 3 + 0] - v * g[0 * 3 + 0]) / rho;
            C dw_dy = (g[3 * 3 + 1] - w * g[0 * 3 + 1]) / rho;
            C dv_dy = (g[2 * 3 + 1] - v * g[0 * 3 + 1]) / rho;
            C dw_dz = (g[3 * 3 + 2] - w * g[0 * 3 + 2]) / rho;
            C du_dy = (g[1 * 3 + 1] - u * g[0 * 3 + 1]) / rho;
            C dv_dz = (g[2 * 3 + 2] - v * g[0 * 3 + 2]) / rho;
            C div = du_dx + dv_dy + dw_dz;
            C tyy = mu * (two * dv_dy - two / three * div);
            C txy = mu * (du_dy + dv_dx);
            C tyz = mu * (dv_dz + dw_dy);
            C ke = half * (ru * ru + rv * rv + rw * rw) / rho;
            C p = (gam() - one) * (e - ke);
            C dke = (ru * g[1 * 3 + 1] + rv * g[2 * 3 + 1] + rw * g[3 * 3 + 1]) / rho - ke * g[0 * 3 + 1] / rho;
            C dp = (gam() - one) * (g[4 * 3 + 1] - dke);
            C coeff = mu * gam() / (pr() * (gam() - one));
            C q = -coeff * (dp / rho - p * g[0 * 3 + 1] / (rho * rho));
            f[0] = (C)0.0; f[1] = txy; f[2] = tyy; f[3] = tyz;
            f[4] = u * txy + v * tyy + w * tyz - q;
        } else {
            C du_dx = (g[1 * 3 + 0] - u * g[0 * 3 + 0]) / rho;
            C dv_dy = (g[2 * 3 + 1] - v * g[0 * 3 + 1]) / rho;
            C dw_dz = (g[3 * 3 + 2] - w * g[0 * 3 + 2]) / rho;
            C du_dz = (g[1 * 3 + 2] - u * g[0 * 3 + 2]) / rho;
            C dv_dz = (g[2 * 3 + 2] - v * g[0 * 3 + 2]) / rho;
            C dw_dx = (g[3 * 3 + 0] - w * g[0 * 3 + 0]) / rho;
            C dw_dy = (g[3 * 3 + 1] - w * g[0 * 3 + 1]) / rho;
            C div = du_dx + dv_dy + dw_dz;
            C tzz = mu * (two * dw_dz - two / three * div);
            C txz = mu * (du_dz + dw_dx);
            C tyz = mu * (dv_dz + dw_dy);
            C ke = half * (ru * ru + rv * rv + rw * rw) / rho;
            C p = (gam() - one) * (e - ke);
            C dke = (ru * g[1 * 3 + 2] + rv * g[2 * 3 + 2] + rw * g[3 * 3 + 2]) / rho - ke * g[0 * 3 + 2] / rho;
            C dp = (gam() - one) * (g[4 * 3 + 2] - dke);
            C coeff = mu * gam() / (pr() * (gam() - one));
            C q = -coeff * (dp / rho - p * g[0 * 3 + 2] / (rho * rho));
            f[0] = (C)0.0; f[1] = txz; f[2] = tyz; f[3] = tzz;
            f[4] = u * txz + v * tyz + w * tzz - q;
        }
    }

    static C sound(C p, C rho) {
        return std::sqrt(std::fmax(((C)0.0), gam() * p / std::fmax(((C)1e-7), rho)));
    }

    static void rusanov(const C* l, const C* r, const C* n, C* f) {  // :104-128
        C rl = l[0], rul = l[1], rvl = l[2], rwl = l[3], el = l[4];
        C rr = r[0], rur = r[1], rvr = r[2], rwr = r[3], er = r[4];
        C ul = rul / rl, vl = rvl / rl, wl = rwl / rl;
        C ur = rur / rr, vr = rvr / rr, wr = rwr / rr;
        C pl = pressure(rl, rul, rvl, rwl, el);
        C pr_ = pressure(rr, rur, rvr, rwr, er);
        C vnl = ul * n[0] + vl * n[1] + wl * n[2];
        C vnr = ur * n[0] + vr * n[1] + wr * n[2];
        C cl = sound(pl, rl), cr = sound(pr_, rr);
        C lam = std::fmax(std::fabs(vnl) + cl, std::fabs(vnr) + cr);
        C fl[5], fr[5];
        fl[0] = rl * vnl; fl[1] = rul * vnl + pl * n[0]; fl[2] = rvl * vnl + pl * n[1];
        fl[3] = rwl * vnl + pl * n[2]; fl[4] = vnl * (el + pl);
        fr[0] = rr * vnr; fr[1] = rur * vnr + pr_ * n[0]; fr[2] = rvr * vnr + pr_ * n[1];
        fr[3] = rwr * vnr + pr_ * n[2]; fr[4] = vnr * (er + pr_);
        for (int v = 0; v < 5; ++v) f[v] = ((C)0.5) * (fl[v] + fr[v] - lam * (r[v] - l[v]));
    }

    static void ausm(const C* l, const C* r, const C* n, C* f) {  // :130-188
        C rl = l[0], rul = l[1], rvl = l[2], rwl = l[3], el = l[4];
        C rr = r[0], rur = r[1], rvr = r[2], rwr = r[3], er = r[4];
        C ul = rul / rl, vl = rvl / rl, wl = rwl / rl;
        C ur = rur / rr, vr = rvr / rr, wr = rwr / rr;
        C pl = pressure(rl, rul, rvl, rwl, el);
        C pr_ = pressure(rr, rur, rvr, rwr, er);
        C vnl = ul * n[0] + vl * n[1] + wl * n[2];
        C vnr = ur * n[0] + vr * n[1] + wr * n[2];
        C hl = (el + pl) / rl, hr = (er + pr_) / rr;
        C cl = sound(pl, rl), cr = sound(pr_, rr);
        C cmax = std::fmax(cl, cr);
        C ml = vnl / cmax, mr = vnr / cmax;
        C mp, mm, pp, pm;
        const C alpha = (C)0.1875, beta = (C)0.125, one = (C)1.0, q = (C)0.25;
        if (std::fabs(ml) >= one) {
            mp = ((C)0.5) * (ml + std::fabs(ml));
            pp = (ml > ((C)0.0)) ? one : ((C)0.0);
        } else {
            mp = q * (ml + one) * (ml + one) + beta * (ml * ml - one) * (ml * ml - one);
            pp = q * (ml + one) * (ml + one) * (((C)2.0) - ml) + alpha * ml * (ml * ml - one) * (ml * ml - one);
        }
        if (std::fabs(mr) >= one) {
            mm = ((C)0.5) * (mr - std::fabs(mr));
            pm = (mr < ((C)0.0)) ? one : ((C)0.0);
        } else {
            mm = -q * (mr - one) * (mr - one) - beta * (mr * mr - one) * (mr * mr - one);
            pm = q * (mr - one) * (mr - one) * (((C)2.0) + mr) - alpha * mr * (mr * mr - one) * (mr * mr - one);
        }
        C mach = mp + mm;
        C p = pp * pl + pm * pr_;
        C mf = mach * cmax * (mach > ((C)0.0) ? rl : rr);
        if (mach > ((C)0.0)) {
            f[0] = mf; f[1] = mf * ul + p * n[0]; f[2] = mf * vl + p * n[1];
            f[3] = mf * wl + p * n[2]; f[4] = mf * hl;
        } else {
            f[0] = mf; f[1] = mf * ur + p * n[0]; f[2] = mf * vr + p * n[1];
            f[3] = mf * wr + p * n[2]; f[4] = mf * hr;
        }
    }

    static void slau2(const C* l, const C* r, const C* n, C* f) {  // :190-257
        C rl = l[0], rul = l[1], rvl = l[2], rwl = l[3], el = l[4];
        C rr = r[0], rur = r[1], rvr = r[2], rwr = r[3], er = r[4];
        C ul = rul / rl, vl = rvl / rl, wl = rwl / rl;
        C ur = rur / rr, vr = rvr / rr, wr = rwr / rr;
        C pl = pressure(rl, rul, rvl, rwl, el);
        C pr_ = pressure(rr, rur, rvr, rwr, er);
        C vnl = ul * n[0] + vl * n[1] + wl * n[2];
        C vnr = ur * n[0] + vr * n[1] + wr * n[2];
        C hl = (el + pl) / rl, hr = (er + pr_) / rr;
        C sql = ul * ul + vl * vl + wl * wl;
        C sqr = ur * ur + vr * vr + wr * wr;
        C cl = sound(pl, rl), cr = sound(pr_, rr);
        const C half = (C)0.5, one = (C)1.0, zero = (C)0.0;
        C cf = half * (cl + cr);
        C ml = vnl / cf, mr = vnr / cf;
        C mt = std::fmin(one, (one / cf) * std::sqrt(half * (sql + sqr)));
        C chi = one - mt;
        chi = chi * chi;
        C g = -std::fmax(std::fmin(ml, zero), -one) * std::fmin(std::fmax(mr, zero), one);
        C vavg = (rl * std::fabs(vnl) + rr * std::fabs(vnr)) / (rl + rr);
        C val = (one - g) * vavg + g * std::fabs(vnl);
        C var = (one - g) * vavg + g * std::fabs(vnr);
        C mf = half * (rl * (vnl + val) + rr * (vnr - var) - (chi / cf) * (pr_ - pl));
        C bl, br;
        if (std::fabs(ml) < one)
            bl = ((C)0.25) * (((C)2.0) - ml) * (ml + one) * (ml + one);
        else
            bl = (ml >= zero) ? one : zero;
        if (std::fabs(mr) < one)
            br = ((C)0.25) * (((C)2.0) + mr) * (mr - one) * (mr - one);
        else
            br = (mr >= zero) ? zero : one;
        C pf = half * (pl + pr_) + half * (bl - br) * (pl - pr_);
        pf += std::sqrt(half * (sql + sqr)) * (bl + br - one) * cf * half * (rl + rr);
        if (mf > zero) {
            f[0] = mf; f[1] = mf * ul + pf * n[0]; f[2] = mf * vl + pf * n[1];
            f[3] = mf * wl + pf * n[2]; f[4] = mf * hl;
        } else {
            f[0] = mf; f[1] = mf * ur + pf * n[0]; f[2] = mf * vr + pf * n[1];
            f[3] = mf * wr + pf * n[2]; f[4] = mf * hr;
        }
    }

    static void riemann(int scheme, const C* l, const C* r, const C* n, C* f) {  // :259-265
        if (scheme == 0) slau2(l, r, n, f);
        else if (scheme == 1) ausm(l, r, n, f);
        else rusanov(l, r, n, f);
    }

    // Sutherland law with the non-dimensional temperature ratio
    // theta = gamma p / rho (c = 1 at the reference state):
    //   mu = ((mu_ref * (theta*sqrt(theta))) * (t_ref + s)) / (t_ref*theta + s)
    static C sutherland(const C* s, C mu_ref, C t_ref, C s_suth) {
        C theta = gam() * pressure(s[0], s[1], s[2], s[3], s[4]) / s[0];
        return ((mu_ref * (theta * std::sqrt(theta))) * (t_ref + s_suth)) / (t_ref * theta + s_suth);
    }
};

// ---------------------------------------------------------------------------
// Small thread pool helper: run fn(begin, end) over [0, n) in contiguous chunks.
// ---------------------------------------------------------------------------
void parallel_for(int64_t n, int n_threads, const std::function<void(int64_t, int64_t)>& fn) {
    int t = n_threads > 0 ? n_threads : (int)std::max(1u, std::thread::hardware_concurrency());
    if (t > n) t = (int)std::max<int64_t>(1, n);
    if (t <= 1) { fn(0, n); return; }
    std::vector<std::thread> th;
    int64_t chunk = (n + t - 1) / t;
    for (int q = 0; q < t; ++q) {
        int64_t b = q * chunk, e = std::min<int64_t>(n, b + chunk);
        if (b >= e) break;
        th.emplace_back(fn, b, e);
    }
    for (auto& x : th) x.join();
}

// Slice skip sets (rhs.cu:129, 151, 169) and the 12-slot row order.
const int kSkip[3][4] = {{8, 10, 13, 14}, {5, 9, 12, 14}, {4, 6, 12, 13}};
inline bool skipped(int dir, int c) {
    for (int q = 0; q < 4; ++q)
        if (kSkip[dir][q] == c) return true;
    return false;
}

// ---------------------------------------------------------------------------
// The solver, templated like the reference's pipeline <U, Q, RHS, COMPUTE>.
// ---------------------------------------------------------------------------
template <class U, class Q, class R, class C>
struct Solver {
    using P = Phys<C>;
    orc_params prm;
    int nx, ny, nz;
    int64_t E, NP, NF;  // elements, points, face points (E*16)
    C pts[4], D[16], eL[4], eR[4], gL[4], gR[4];
    C lx, ly, lz, jx, jy, jz, mu_fixed, mu_ref, t_ref, s_suth, bf;
    std::vector<R> rhs;
    std::vector<C> sl[6];  // slices xL,xR,yL,yR,zL,zR
    std::vector<C> jarr[6];  // j arrays xL,xR,yL,yR,zL,zR

    explicit Solver(const orc_params& p) : prm(p) {
        nx = p.nx; ny = p.ny; nz = p.nz;
        E = (int64_t)nx * ny * nz;
        NP = E * 64;
        NF = E * 16;
        const FrOps& o = ops();
        for (int i = 0; i < 4; ++i) {
            pts[i] = (C)o.points[i]; eL[i] = (C)o.el[i]; eR[i] = (C)o.er[i];
            gL[i] = (C)o.gl[i]; gR[i] = (C)o.gr[i];
        }
        for (int i = 0; i < 16; ++i) D[i] = (C)o.d[i];
        lx = (C)p.lx; ly = (C)p.ly; lz = (C)p.lz;
        // jac = 2 / (L / n) in the compute type (rhs.cu:366-369, rk.cu:93-95)
        jx = ((C)2.0) / (lx / ((C)nx));
        jy = ((C)2.0) / (ly / ((C)ny));
        jz = ((C)2.0) / (lz / ((C)nz));
        mu_fixed = (C)p.mu_ref;
        mu_ref = (C)p.mu_ref; t_ref = (C)p.t_ref; s_suth = (C)p.s_suth;
        bf = (C)p.body_force_x;
        rhs.assign(5 * NP, (R)0);
        for (int f = 0; f < 6; ++f) {
            if (p.viscous) sl[f].assign(NF * 12, (C)0);
            jarr[f].assign(5 * NF, (C)0);
        }
    }

    int64_t eidx(int ex, int ey, int ez) const { return ex + (int64_t)nx * (ey + (int64_t)ny * ez); }
    // Per-element eddy viscosity added to the molecular one (no reference
    // counterpart; empty = plain FR): mu = mu_mol + mu_t[e] in the compute type.
    std::vector<C> mut;
    C mu_at(const C* s, int64_t e) const {
        C m = prm.viscosity_model == 1 ? P::sutherland(s, mu_ref, t_ref, s_suth) : mu_fixed;
        if (!mut.empty()) m = m + mut[e];
        return m;
    }
    static C pw(C t0, C t1, C t2, C t3) { return (t0 + t1) + (t2 + t3); }  // shfl butterfly
    static C seq0(C t0, C t1, C t2, C t3) {                              // acc = 0; acc += t
        C a = 0;
        a += t0; a += t1; a += t2; a += t3;
        return a;
    }
    static C seq1(C t0, C t1, C t2, C t3) { return t0 + t1 + t2 + t3; }  // from the first product

    // ---- compute_rhs_volume_kernel, rhs.cu:344-511 -------------------------
    void volume(const U* u, int64_t e) {
        const int64_t TP = NP;
        int ex = (int)(e % nx), ey = (int)((e / nx) % ny), ez = (int)(e / ((int64_t)nx * ny));
        C s[5][64];
        for (int v = 0; v < 5; ++v)
            for (int pt = 0; pt < 64; ++pt) s[v][pt] = (C)u[v * TP + e * 64 + pt];
        auto U_at = [&](int v, int64_t el, int pt) -> C { return (C)u[v * TP + el * 64 + pt]; };

        int64_t exl = eidx((ex - 1 + nx) % nx, ey, ez), exr = eidx((ex + 1) % nx, ey, ez);
        int64_t ezl = eidx(ex, ey, (ez - 1 + nz) % nz), ezr = eidx(ex, ey, (ez + 1) % nz);
        int64_t eyl, eyr;
        if (prm.y_wall) {
            eyl = ey > 0 ? eidx(ex, ey - 1, ez) : -1;
            eyr = ey < ny - 1 ? eidx(ex, ey + 1, ez) : -1;
        } else {
            eyl = eidx(ex, (ey - 1 + ny) % ny, ez);
            eyr = eidx(ex, (ey + 1) % ny, ez);
        }

        C g[15][64];
        if (prm.viscous) {
            for (int pt = 0; pt < 64; ++pt) {
                int i = pt % 4, j = (pt / 4) % 4, k = pt / 16;
                C gx[5], gy[5], gz[5];
                for (int v = 0; v < 5; ++v) {
                    C dz = 0, dxi = 0, deta = 0;
                    for (int m = 0; m < 4; ++m) dz += D[k * 4 + m] * s[v][i + j * 4 + m * 16];
                    for (int m = 0; m < 4; ++m) dxi += D[i * 4 + m] * s[v][m + j * 4 + k * 16];
                    for (int m = 0; m < 4; ++m) deta += D[j * 4 + m] * s[v][i + m * 4 + k * 16];
                    gx[v] = dxi * jx;
                    gy[v] = deta * jy;
                    gz[v] = dz * jz;
                }
                auto jump = [](C& gc, C own, C nb, C gp, C jac) {  // accumulate_jump, rhs.cu:105-108
                    C jmp = ((C)0.5) * (nb + own) - own;
                    gc += jmp * gp * jac;
                };
                // x faces: own and neighbour traces by shuffle butterfly (pairwise)
                C ul[5], ur[5];
                for (int v = 0; v < 5; ++v) {
                    ul[v] = pw(s[v][0 + j * 4 + k * 16] * eL[0], s[v][1 + j * 4 + k * 16] * eL[1],
                               s[v][2 + j * 4 + k * 16] * eL[2], s[v][3 + j * 4 + k * 16] * eL[3]);
                    ur[v] = pw(s[v][0 + j * 4 + k * 16] * eR[0], s[v][1 + j * 4 + k * 16] * eR[1],
                               s[v][2 + j * 4 + k * 16] * eR[2], s[v][3 + j * 4 + k * 16] * eR[3]);
                }
                for (int v = 0; v < 5; ++v) {
                    C nb = pw(U_at(v, exl, 0 + j * 4 + k * 16) * eR[0], U_at(v, exl, 1 + j * 4 + k * 16) * eR[1],
                              U_at(v, exl, 2 + j * 4 + k * 16) * eR[2], U_at(v, exl, 3 + j * 4 + k * 16) * eR[3]);
                    jump(gx[v], ul[v], nb, gL[i], jx);
                }
                for (int v = 0; v < 5; ++v) {
                    C nb = pw(U_at(v, exr, 0 + j * 4 + k * 16) * eL[0], U_at(v, exr, 1 + j * 4 + k * 16) * eL[1],
                              U_at(v, exr, 2 + j * 4 + k * 16) * eL[2], U_at(v, exr, 3 + j * 4 + k * 16) * eL[3]);
                    jump(gx[v], ur[v], nb, gR[i], jx);
                }
                // y faces (pairwise), slip-wall ghost negates the momenta (rhs.cu:111-114)
                for (int v = 0; v < 5; ++v) {
                    ul[v] = pw(s[v][i + 0 * 4 + k * 16] * eL[0], s[v][i + 1 * 4 + k * 16] * eL[1],
                               s[v][i + 2 * 4 + k * 16] * eL[2], s[v][i + 3 * 4 + k * 16] * eL[3]);
                    ur[v] = pw(s[v][i + 0 * 4 + k * 16] * eR[0], s[v][i + 1 * 4 + k * 16] * eR[1],
                               s[v][i + 2 * 4 + k * 16] * eR[2], s[v][i + 3 * 4 + k * 16] * eR[3]);
                }
                auto reflect = [](int v, C value) {
                    C sign = (v > 0 && v < 4) ? ((C)-1.0) : ((C)1.0);
                    return value * sign;
                };
                if (eyl < 0) {
                    for (int v = 0; v < 5; ++v) jump(gy[v], ul[v], reflect(v, ul[v]), gL[j], jy);
                } else {
                    for (int v = 0; v < 5; ++v) {
                        C nb = pw(U_at(v, eyl, i + 0 * 4 + k * 16) * eR[0], U_at(v, eyl, i + 1 * 4 + k * 16) * eR[1],
                                  U_at(v, eyl, i + 2 * 4 + k * 16) * eR[2], U_at(v, eyl, i + 3 * 4 + k * 16) * eR[3]);
                        jump(gy[v], ul[v], nb, gL[j], jy);
                    }
                }
                if (eyr < 0) {
                    for (int v = 0; v < 5; ++v) jump(gy[v], ur[v], reflect(v, ur[v]), gR[j], jy);
                } else {
                    for (int v = 0; v < 5; ++v) {
                        C nb = pw(U_at(v, eyr, i + 0 * 4 + k * 16) * eL[0], U_at(v, eyr, i + 1 * 4 + k * 16) * eL[1],
                                  U_at(v, eyr, i + 2 * 4 + k * 16) * eL[2], U_at(v, eyr, i + 3 * 4 + k * 16) * eL[3]);
                        jump(gy[v], ur[v], nb, gR[j], jy);
                    }
                }
                // z faces: sequential m-loops from 0 (rhs.cu:89-101)
                for (int v = 0; v < 5; ++v) {
                    ul[v] = seq0(s[v][i + j * 4 + 0] * eL[0], s[v][i + j * 4 + 16] * eL[1],
                                 s[v][i + j * 4 + 32] * eL[2], s[v][i + j * 4 + 48] * eL[3]);
                    ur[v] = seq0(s[v][i + j * 4 + 0] * eR[0], s[v][i + j * 4 + 16] * eR[1],
                                 s[v][i + j * 4 + 32] * eR[2], s[v][i + j * 4 + 48] * eR[3]);
                }
                for (int v = 0; v < 5; ++v) {
                    C nb = seq0(U_at(v, ezl, i + j * 4 + 0) * eR[0], U_at(v, ezl, i + j * 4 + 16) * eR[1],
                                U_at(v, ezl, i + j * 4 + 32) * eR[2], U_at(v, ezl, i + j * 4 + 48) * eR[3]);
                    jump(gz[v], ul[v], nb, gL[k], jz);
                }
                for (int v = 0; v < 5; ++v) {
                    C nb = seq0(U_at(v, ezr, i + j * 4 + 0) * eL[0], U_at(v, ezr, i + j * 4 + 16) * eL[1],
                                U_at(v, ezr, i + j * 4 + 32) * eL[2], U_at(v, ezr, i + j * 4 + 48) * eL[3]);
                    jump(gz[v], ur[v], nb, gR[k], jz);
                }
                // Q_REAL round trip (rhs.cu:457-465)
                for (int v = 0; v < 5; ++v) {
                    g[v * 3 + 0][pt] = (C)(Q)gx[v];
                    g[v * 3 + 1][pt] = (C)(Q)gy[v];
                    g[v * 3 + 2][pt] = (C)(Q)gz[v];
                }
            }
            // gradient face slices (rhs.cu:471-488, 126-180)
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) {
                    // x: i == 0 lanes, (j,k) = (a,b), sequential from the first product
                    {
                        int fp = a + 4 * b;
                        C* rl = &sl[0][(e * 16 + fp) * 12];
                        C* rr = &sl[1][(e * 16 + fp) * 12];
                        int slot = 0;
                        for (int c = 0; c < 15; ++c) {
                            if (skipped(0, c)) continue;
                            const C* col = &g[c][a * 4 + b * 16];
                            C el = col[0] * eL[0];
                            el += col[1] * eL[1];
                            el += col[2] * eL[2];
                            el += col[3] * eL[3];
                            C er = col[0] * eR[0];
                            er += col[1] * eR[1];
                            er += col[2] * eR[2];
                            er += col[3] * eR[3];
                            rl[slot] = el; rr[slot] = er; ++slot;
                        }
                        rl[11] = 0; rr[11] = 0;
                    }
                    // y: j == 0 lanes, (i,k) = (a,b), pairwise of rounded products
                    {
                        int fp = a + 4 * b;
                        C* rl = &sl[2][(e * 16 + fp) * 12];
                        C* rr = &sl[3][(e * 16 + fp) * 12];
                        int slot = 0;
                        for (int c = 0; c < 15; ++c) {
                            if (skipped(1, c)) continue;
                            const C* col = &g[c][a + b * 16];
                            C s0 = col[0], s1 = col[4], s2 = col[8], s3 = col[12];
                            rl[slot] = pw(s0 * eL[0], s1 * eL[1], s2 * eL[2], s3 * eL[3]);
                            rr[slot] = pw(s0 * eR[0], s1 * eR[1], s2 * eR[2], s3 * eR[3]);
                            ++slot;
                        }
                        rl[11] = 0; rr[11] = 0;
                    }
                    // z: k == 0 lanes, (i,j) = (a,b), sequential from 0
                    {
                        int fp = a + 4 * b;
                        C* rl = &sl[4][(e * 16 + fp) * 12];
                        C* rr = &sl[5][(e * 16 + fp) * 12];
                        int slot = 0;
                        for (int c = 0; c < 15; ++c) {
                            if (skipped(2, c)) continue;
                            const C* col = &g[c][a + b * 4];
                            C el = 0, er = 0;
                            for (int m = 0; m < 4; ++m) {
                                C smp = col[m * 16];
                                el += smp * eL[m];
                                er += smp * eR[m];
                            }
                            rl[slot] = el; rr[slot] = er; ++slot;
                        }
                        rl[11] = 0; rr[11] = 0;
                    }
                }
        }

        // volume flux divergence (rhs.cu:229-266, 490-510)
        C rs[5][64];
        for (int v = 0; v < 5; ++v)
            for (int pt = 0; pt < 64; ++pt) rs[v][pt] = 0;
        C F[5][64];
        for (int dir = 0; dir < 3; ++dir) {
            C jac = dir == 0 ? jx : (dir == 1 ? jy : jz);
            for (int pt = 0; pt < 64; ++pt) {
                C st[5], gr[15], fe[5], fv[5];
                for (int v = 0; v < 5; ++v) st[v] = s[v][pt];
                P::euler(dir, st, fe);
                if (prm.viscous) {
                    for (int c = 0; c < 15; ++c) gr[c] = g[c][pt];
                    P::viscous(dir, st, gr, mu_at(st, e), fv);
                    for (int v = 0; v < 5; ++v) F[v][pt] = fe[v] - fv[v];
                } else {
                    for (int v = 0; v < 5; ++v) F[v][pt] = fe[v];
                }
            }
            for (int pt = 0; pt < 64; ++pt) {
                int i = pt % 4, j = (pt / 4) % 4, k = pt / 16;
                int idx = dir == 0 ? i : (dir == 1 ? j : k);
                for (int v = 0; v < 5; ++v) {
                    C d = 0;
                    for (int m = 0; m < 4; ++m) {
                        int q = dir == 0 ? (m + j * 4 + k * 16) : (dir == 1 ? (i + m * 4 + k * 16) : (i + j * 4 + m * 16));
                        d += D[idx * 4 + m] * F[v][q];
                    }
                    rs[v][pt] += d * jac;
                }
            }
        }
        for (int pt = 0; pt < 64; ++pt)
            for (int v = 0; v < 5; ++v) {
                C val = -rs[v][pt];
                if (v == 1) val += bf;
                rhs[v * TP + e * 64 + pt] = (R)val;
            }
    }

    // compute_face_j, rhs.cu:273-297 (inviscid mode drops the viscous terms)
    void face_j(int dir, const C* ul, const C* ur, const C* gl_, const C* gr_, C* jl_arr, C* jr_arr,
                int64_t el, int64_t er, int fp) {
        C n[3] = {(C)0.0, (C)0.0, (C)0.0};
        n[dir] = (C)1.0;
        C fi[5], fvl[5], fvr[5], fel[5], fer[5];
        P::riemann(prm.flux_scheme, ul, ur, n, fi);
        if (prm.viscous) {
            P::viscous(dir, ul, gl_, mu_at(ul, el), fvl);
            P::viscous(dir, ur, gr_, mu_at(ur, er), fvr);
        }
        P::euler(dir, ul, fel);
        P::euler(dir, ur, fer);
        for (int v = 0; v < 5; ++v) {
            C jl, jr;
            if (prm.viscous) {
                C fc = fi[v] - ((C)0.5) * (fvl[v] + fvr[v]);
                jl = fc - (fel[v] - fvl[v]);
                jr = fc - (fer[v] - fvr[v]);
            } else {
                jl = fi[v] - fel[v];
                jr = fi[v] - fer[v];
            }
            jl_arr[v * NF + el * 16 + fp] = jl;  // the left element's +side slot
            jr_arr[v * NF + er * 16 + fp] = jr;  // the right element's -side slot
        }
    }

    // read_gradient_slice, rhs.cu:201-217
    void read_slice(int dir, const C* row, C* grad) const {
        int slot = 0;
        for (int c = 0; c < 15; ++c) {
            if (skipped(dir, c)) { grad[c] = 0; continue; }
            grad[c] = row[slot++];
        }
    }

    // compute_face_flux_x_kernel, rhs.cu:513-556
    void face_x(const U* u, int64_t e) {
        int ex = (int)(e % nx), ey = (int)((e / nx) % ny), ez = (int)(e / ((int64_t)nx * ny));
        int64_t el = eidx((ex - 1 + nx) % nx, ey, ez), er = e;
        for (int fp = 0; fp < 16; ++fp) {
            C ul[5], ur[5], gl_[15] = {0}, gr_[15] = {0};
            for (int v = 0; v < 5; ++v) {
                const U* bl = u + v * NP + el * 64 + 4 * fp;
                const U* br = u + v * NP + er * 64 + 4 * fp;
                ul[v] = seq1((C)bl[0] * eR[0], (C)bl[1] * eR[1], (C)bl[2] * eR[2], (C)bl[3] * eR[3]);
                ur[v] = seq1((C)br[0] * eL[0], (C)br[1] * eL[1], (C)br[2] * eL[2], (C)br[3] * eL[3]);
            }
            if (prm.viscous) {
                read_slice(0, &sl[1][(el * 16 + fp) * 12], gl_);
                read_slice(0, &sl[0][(er * 16 + fp) * 12], gr_);
            }
            face_j(0, ul, ur, gl_, gr_, jarr[1].data(), jarr[0].data(), el, er, fp);
        }
    }

    // compute_face_flux_y_kernel, rhs.cu:615-687 (ey = index of the right element)
    void face_y(const U* u, int64_t e) {
        int ex = (int)(e % nx), ey = (int)((e / nx) % ny), ez = (int)(e / ((int64_t)nx * ny));
        if (prm.y_wall && ey == 0) return;  // wall face: face_y_wall
        int yl = prm.y_wall ? ey - 1 : (ey - 1 + ny) % ny;
        int64_t el = eidx(ex, yl, ez), er = e;
        for (int fp = 0; fp < 16; ++fp) {
            int i = fp % 4, k = fp / 4;
            C ul[5], ur[5], gl_[15] = {0}, gr_[15] = {0};
            for (int v = 0; v < 5; ++v) {
                C tl[4], tr[4];
                for (int m = 0; m < 4; ++m) {
                    tl[m] = (C)u[v * NP + el * 64 + i + m * 4 + k * 16] * eR[m];
                    tr[m] = (C)u[v * NP + er * 64 + i + m * 4 + k * 16] * eL[m];
                }
                ul[v] = pw(tl[0], tl[1], tl[2], tl[3]);
                ur[v] = pw(tr[0], tr[1], tr[2], tr[3]);
            }
            if (prm.viscous) {
                read_slice(1, &sl[3][(el * 16 + fp) * 12], gl_);
                read_slice(1, &sl[2][(er * 16 + fp) * 12], gr_);
            }
            face_j(1, ul, ur, gl_, gr_, jarr[3].data(), jarr[2].data(), el, er, fp);
        }
    }

    // compute_face_flux_y_wall_kernel + compute_face_j_wall, rhs.cu:572-613, 310-339
    void face_y_wall(const U* u, int64_t col) {  // col = ex + nx*ez
        int ex = (int)(col % nx), ez = (int)(col / nx);
        for (int plane = 0; plane < 2; ++plane) {
            bool upper = plane == 1;
            int64_t er = upper ? eidx(ex, ny - 1, ez) : eidx(ex, 0, ez);
            for (int fp = 0; fp < 16; ++fp) {
                int i = fp % 4, k = fp / 4;
                C st[5], grd[15] = {0};
                for (int v = 0; v < 5; ++v) {
                    C t[4];
                    for (int m = 0; m < 4; ++m)
                        t[m] = (C)u[v * NP + er * 64 + i + m * 4 + k * 16] * (upper ? eR[m] : eL[m]);
                    st[v] = pw(t[0], t[1], t[2], t[3]);
                }
                if (prm.viscous) read_slice(1, &sl[upper ? 3 : 2][(er * 16 + fp) * 12], grd);
                C fv[5] = {0, 0, 0, 0, 0}, fe[5];
                if (prm.viscous) P::viscous(1, st, grd, mu_at(st, er), fv);
                P::euler(1, st, fe);
                C v2r = st[2] * st[2] / st[0];
                C wall[5] = {(C)0.0, (C)0.0, fe[2] - v2r, (C)0.0, (C)0.0};
                C* dst = upper ? jarr[3].data() : jarr[2].data();
                for (int v = 0; v < 5; ++v) {
                    C jv;
                    if (prm.viscous) {
                        C fc = wall[v] - fv[v];
                        jv = fc - (fe[v] - fv[v]);
                    } else {
                        jv = wall[v] - fe[v];
                    }
                    dst[v * NF + er * 16 + fp] = jv;
                }
            }
        }
    }

    // compute_face_flux_z_kernel, rhs.cu:689-739
    void face_z(const U* u, int64_t e) {
        int ex = (int)(e % nx), ey = (int)((e / nx) % ny), ez = (int)(e / ((int64_t)nx * ny));
        int64_t el = eidx(ex, ey, (ez - 1 + nz) % nz), er = e;
        for (int fp = 0; fp < 16; ++fp) {
            int i = fp % 4, j = fp / 4;
            C ul[5] = {0, 0, 0, 0, 0}, ur[5] = {0, 0, 0, 0, 0}, gl_[15] = {0}, gr_[15] = {0};
            for (int m = 0; m < 4; ++m)
                for (int v = 0; v < 5; ++v) {
                    ur[v] += (C)u[v * NP + er * 64 + i + j * 4 + m * 16] * eL[m];
                    ul[v] += (C)u[v * NP + el * 64 + i + j * 4 + m * 16] * eR[m];
                }
            if (prm.viscous) {
                read_slice(2, &sl[5][(el * 16 + fp) * 12], gl_);
                read_slice(2, &sl[4][(er * 16 + fp) * 12], gr_);
            }
            face_j(2, ul, ur, gl_, gr_, jarr[5].data(), jarr[4].data(), el, er, fp);
        }
    }

    // One stage's PDE work (solver.rs:610-694): volume, x, y, [wall], z.
    void rhs_stage(const U* u) {
        int nt = prm.n_threads;
        parallel_for(E, nt, [&](int64_t b, int64_t en) { for (int64_t e = b; e < en; ++e) volume(u, e); });
        parallel_for(E, nt, [&](int64_t b, int64_t en) {
            for (int64_t e = b; e < en; ++e) { face_x(u, e); face_y(u, e); face_z(u, e); }
        });
        if (prm.y_wall)
            parallel_for((int64_t)nx * nz, nt, [&](int64_t b, int64_t en) {
                for (int64_t c = b; c < en; ++c) face_y_wall(u, c);
            });
    }

    // rhs_with_j_corrections4, rk.cu:12-25
    C rhs_total(int64_t v, int64_t e, int pt) const {
        int i = pt % 4, j = (pt / 4) % 4, k = pt / 16;
        int64_t base = v * NF + e * 16;
        C jxl = jarr[0][base + j + 4 * k], jxr = jarr[1][base + j + 4 * k];
        C jyl = jarr[2][base + i + 4 * k], jyr = jarr[3][base + i + 4 * k];
        C jzl = jarr[4][base + i + 4 * j], jzr = jarr[5][base + i + 4 * j];
        C cx = -(jxl * gL[i] * jx) - (jxr * gR[i] * jx);
        C cy = -(jyl * gL[j] * jy) - (jyr * gR[j] * jy);
        C cz = -(jzl * gL[k] * jz) - (jzr * gR[k] * jz);
        return (C)rhs[v * NP + e * 64 + pt] + (cx + cy + cz);
    }

};

// One RK3 step, solver.rs:262-317, with the stage update of rk.cu:27-115:
//   stage 1: u1 = 1*u + dt*R(u)
//   stage 2: u1 = 0.75*u + 0.25*u1 + (0.25*dt)*R(u1)
//   stage 3: u  = (1/3)*u + (2/3)*u1 + ((2/3)*dt)*R(u1)
// Coefficients are formed in the compute type from the f32 dt (solver.rs:256,
// 275-316); the blend is evaluated in the compute type and cast to the state
// type (rk.cu:63-82).  rk_accum = 1 instead blends in the state type.
template <class U, class Q, class R, class C>
void rk3_step(Solver<U, Q, R, C>& s, U* u, U* u1, double dt) {
    auto blend = [&](U* dst, const U* prev, const U* cur, bool stage1, double a, double b, double cm) {
        const C ac = (C)a, bc = (C)b, cc = (C)cm * (C)dt;
        const U au = (U)a, bu = (U)b, cu = (U)cm * (U)dt;
        parallel_for(s.E, s.prm.n_threads, [&](int64_t b0, int64_t e0) {
            for (int64_t e = b0; e < e0; ++e)
                for (int v = 0; v < 5; ++v)
                    for (int pt = 0; pt < 64; ++pt) {
                        int64_t idx = v * s.NP + e * 64 + pt;
                        C rt = s.rhs_total(v, e, pt);
                        if (s.prm.rk_accum == 0) {
                            C up = (C)prev[idx];
                            if (stage1) {
                                dst[idx] = (U)(ac * up + cc * rt);
                            } else {
                                C uu = (C)cur[idx];
                                dst[idx] = (U)(ac * up + bc * uu + cc * rt);
                            }
                        } else {
                            U up = prev[idx];
                            U r = (U)rt;
                            if (stage1) {
                                dst[idx] = (U)(au * up + cu * r);
                            } else {
                                U uu = cur[idx];
                                dst[idx] = (U)(au * up + bu * uu + cu * r);
                            }
                        }
                    }
        });
    };
    s.rhs_stage(u);
    blend(u1, u, nullptr, true, 1.0, 0.0, 1.0);          // stage 1 (rk.cu:101-115)
    s.rhs_stage(u1);
    blend(u1, u, u1, false, 0.75, 0.25, 0.25);           // stage 2
    s.rhs_stage(u1);
    blend(u, u, u1, false, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0);  // stage 3
}

template <class U, class Q, class R, class C>
int run_step(const orc_params& p, void* state, double dt, int n, const double* mu_t) {
    Solver<U, Q, R, C> s(p);
    if (mu_t) {
        s.mut.resize(s.E);
        for (int64_t e = 0; e < s.E; ++e) s.mut[e] = (C)mu_t[e];
    }
    U* u = (U*)state;
    std::vector<U> u1(5 * s.NP);
    for (int t = 0; t < n; ++t) rk3_step(s, u, u1.data(), dt);
    return 0;
}

template <class U, class Q, class R, class C>
int run_rhs_total(const orc_params& p, const void* state, void* out) {
    Solver<U, Q, R, C> s(p);
    s.rhs_stage((const U*)state);
    C* o = (C*)out;
    for (int v = 0; v < 5; ++v)
        for (int64_t e = 0; e < s.E; ++e)
            for (int pt = 0; pt < 64; ++pt) o[v * s.NP + e * 64 + pt] = s.rhs_total(v, e, pt);
    return 0;
}

template <class U, class Q, class R, class C>
int run_debug(const orc_params& p, const void* state, int which, void* out) {
    Solver<U, Q, R, C> s(p);
    s.rhs_stage((const U*)state);
    if (which == 0) {
        std::memcpy(out, s.rhs.data(), s.rhs.size() * sizeof(R));
    } else if (which >= 1 && which <= 6) {
        if (!p.viscous) return 2;
        std::memcpy(out, s.sl[which - 1].data(), s.sl[which - 1].size() * sizeof(C));
    } else if (which >= 7 && which <= 12) {
        std::memcpy(out, s.jarr[which - 7].data(), s.jarr[which - 7].size() * sizeof(C));
    } else {
        return 1;
    }
    return 0;
}

// Runtime precision dispatch (solver.rs:841-895 / main.rs:19-25).
#define OD_DISPATCH(FN, p, ...)                                                          \
    [&]() -> int {                                                                       \
        int uu = (p).u_real, qq = (p).q_real, rr = (p).rhs_real, cc = (p).compute_real; \
        auto ok = [](int x) { return x == 4 || x == 8; };                                \
        if (!ok(uu) || !ok(qq) || !ok(rr) || !ok(cc)) return -1;                        \
        int code = (uu == 8) * 8 + (qq == 8) * 4 + (rr == 8) * 2 + (cc == 8);            \
        switch (code) {                                                                  \
            case 0: return FN<float, float, float, float>(__VA_ARGS__);                  \
            case 1: return FN<float, float, float, double>(__VA_ARGS__);                 \
            case 2: return FN<float, float, double, float>(__VA_ARGS__);                 \
            case 3: return FN<float, float, double, double>(__VA_ARGS__);                \
            case 4: return FN<float, double, float, float>(__VA_ARGS__);                 \
            case 5: return FN<float, double, float, double>(__VA_ARGS__);                \
            case 6: return FN<float, double, double, float>(__VA_ARGS__);                \
            case 7: return FN<float, double, double, double>(__VA_ARGS__);               \
            case 8: return FN<double, float, float, float>(__VA_ARGS__);                 \
            case 9: return FN<double, float, float, double>(__VA_ARGS__);                \
            case 10: return FN<double, float, double, float>(__VA_ARGS__);               \
            case 11: return FN<double, float, double, double>(__VA_ARGS__);              \
            case 12: return FN<double, double, float, float>(__VA_ARGS__);               \
            case 13: return FN<double, double, float, double>(__VA_ARGS__);              \
            case 14: return FN<double, double, double, float>(__VA_ARGS__);              \
            default: return FN<double, double, double, double>(__VA_ARGS__);            \
        }                                                                                \
    }()

bool valid(const orc_params* p) {
    return p && p->nx > 0 && p->ny > 0 && p->nz > 0 && p->flux_scheme >= 0 && p->flux_scheme <= 2;
}

template <class C>
int point_impl(int kind, const double* in, double* out) {
    using P = Phys<C>;
    C a[32], o[5];
    for (int q = 0; q < 32; ++q) a[q] = (C)in[q];
    if (kind == 0) {
        out[0] = (double)P::pressure(a[0], a[1], a[2], a[3], a[4]);
        return 0;
    }
    if (kind >= 1 && kind <= 3) {
        P::euler(kind - 1, a, o);
    } else if (kind >= 4 && kind <= 6) {
        P::viscous(kind - 4, a, a + 5, a[20], o);
    } else if (kind >= 7 && kind <= 9) {
        P::riemann(kind - 7, a, a + 5, a + 10, o);
    } else if (kind == 10) {
        out[0] = (double)P::sutherland(a, a[5], a[6], a[7]);
        return 0;
    } else {
        return 1;
    }
    for (int v = 0; v < 5; ++v) out[v] = (double)o[v];
    return 0;
}

template <class U>
int sums_impl(const orc_params& p, const void* state, double out[5]) {
    const FrOps& o = ops();
    const U* u = (const U*)state;
    int64_t E = (int64_t)p.nx * p.ny * p.nz, NP = E * 64;
    double jac = ((double)p.lx / p.nx) * ((double)p.ly / p.ny) * ((double)p.lz / p.nz) / 8.0;
    for (int v = 0; v < 5; ++v) {
        long double acc = 0.0L;
        for (int64_t e = 0; e < E; ++e)
            for (int pt = 0; pt < 64; ++pt) {
                int i = pt % 4, j = (pt / 4) % 4, k = pt / 16;
                acc += (long double)(o.w[i] * o.w[j] * o.w[k]) * (long double)u[v * NP + e * 64 + pt];
            }
        out[v] = (double)(acc * (long double)jac);
    }
    return 0;
}

template <class U>
double cfl_impl(const orc_params& p, const void* state, double cfl) {
    const U* u = (const U*)state;
    int64_t NP = (int64_t)p.nx * p.ny * p.nz * 64;
    double dx = p.lx / p.nx, dy = p.ly / p.ny, dz = p.lz / p.nz;
    double lam = 0.0;
    for (int64_t q = 0; q < NP; ++q) {
        double r = (double)u[q], ru = (double)u[NP + q], rv = (double)u[2 * NP + q];
        double rw = (double)u[3 * NP + q], e = (double)u[4 * NP + q];
        double ke = 0.5 * (ru * ru + rv * rv + rw * rw) / r;
        double pp = std::fmax(1e-7, 0.4 * (e - ke));
        double c = std::sqrt(1.4 * pp / r);
        double l = (std::fabs(ru / r) + c) / dx + (std::fabs(rv / r) + c) / dy + (std::fabs(rw / r) + c) / dz;
        lam = std::fmax(lam, l);
    }
    double dt = cfl / (7.0 * lam);
    return (double)(float)dt;
}

}  // namespace

extern "C" {

void orc_fr_operators(double points[4], double d_matrix[16], double extrap_left[4], double extrap_right[4],
                      double g_left_prime[4], double g_right_prime[4], double weights[4]) {
    const FrOps& o = ops();
    for (int i = 0; i < 4; ++i) {
        if (points) points[i] = o.points[i];
        if (extrap_left) extrap_left[i] = o.el[i];
        if (extrap_right) extrap_right[i] = o.er[i];
        if (g_left_prime) g_left_prime[i] = o.gl[i];
        if (g_right_prime) g_right_prime[i] = o.gr[i];
        if (weights) weights[i] = o.w[i];
    }
    if (d_matrix)
        for (int i = 0; i < 16; ++i) d_matrix[i] = o.d[i];
}

int orc_step(const orc_params* p, void* state, double dt, int n_steps) {
    if (!valid(p) || !state || n_steps < 0) return 1;
    return OD_DISPATCH(run_step, *p, *p, state, dt, n_steps, (const double*)nullptr);
}

int orc_step_mut(const orc_params* p, void* state, const double* mu_t, double dt, int n_steps) {
    if (!valid(p) || !state || !mu_t || n_steps < 0) return 1;
    return OD_DISPATCH(run_step, *p, *p, state, dt, n_steps, mu_t);
}

int orc_rhs_total(const orc_params* p, const void* state, void* out) {
    if (!valid(p) || !state || !out) return 1;
    return OD_DISPATCH(run_rhs_total, *p, *p, state, out);
}

int orc_debug_array(const orc_params* p, const void* state, int which, void* out) {
    if (!valid(p) || !state || !out) return 1;
    return OD_DISPATCH(run_debug, *p, *p, state, which, out);
}

int orc_point(int real, int kind, const double* in, double* out) {
    if (real == 4) return point_impl<float>(kind, in, out);
    if (real == 8) return point_impl<double>(kind, in, out);
    return 1;
}

int orc_sums(const orc_params* p, const void* state, double out[5]) {
    if (!valid(p) || !state) return 1;
    return p->u_real == 4 ? sums_impl<float>(*p, state, out) : sums_impl<double>(*p, state, out);
}

double orc_cfl_dt(const orc_params* p, const void* state, double cfl) {
    if (!valid(p) || !state) return -1.0;
    return p->u_real == 4 ? cfl_impl<float>(*p, state, cfl) : cfl_impl<double>(*p, state, cfl);
}

}  // extern "C"
