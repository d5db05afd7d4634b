// Learned closure (no reference counterpart): element features, the MLP
// (BF16 weights and activations on MFMA, FP32 accumulation; hand-written
// forward and backward), Adam, and the eddy-viscosity refresh of a solver
// context with its training-range and kinetic-energy-budget fallbacks.
//
// MFMA used: v_mfma_f32_16x16x16_bf16, D[16x16] = A[16x16] B[16x16] + C with,
// for lane l (r = l % 16, g = l / 16):
//   A: row r, k = 4g .. 4g+3        B: column r, k = 4g .. 4g+3
//   D: rows 4g .. 4g+3, column r
// A tile is 16 elements; a layer Z = H W^T (H [16][K], W [64][K]) is 4
// column blocks x K/16 k-steps.  The transposes between the D layout
// (elements along the 4 registers) and the A layout (elements along lanes)
// go through LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "od_ctx.hpp"

using namespace od::host;

namespace {

constexpr int NF = OD_MLP_FEATURES;    // 12 input features
constexpr int KP = 16;                 // padded input width
constexpr int NH = OD_MLP_HIDDEN;      // 64 hidden units
constexpr int FS = OD_FEATURE_STRIDE;  // floats per element from the feature kernel
constexpr int oW1 = 0, oB1 = oW1 + NH * NF, oW2 = oB1 + NH, oB2 = oW2 + NH * NH, oW3 = oB2 + NH,
              oB3 = oW3 + NH * NH, oW4 = oB3 + NH, oB4 = oW4 + NH, NPARAM = oB4 + 1;
static_assert(NPARAM == OD_MLP_PARAMS, "parameter layout");
constexpr int LDH = NH + 8;  // LDS row stride (bf16) of a 16 x 64 activation tile

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// f32 -> bf16, round to nearest even (finite inputs).
__host__ __device__ inline unsigned short f2bf(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short b) { return __uint_as_float((uint32_t)b << 16); }

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// tanh-form GELU and its derivative
__device__ inline float gelu(float x) {
    const float a = 0.7978845608028654f, b = 0.044715f;
    return 0.5f * x * (1.0f + tanhf(a * (x + b * x * x * x)));
}
__device__ inline float gelu_d(float x) {
    const float a = 0.7978845608028654f, b = 0.044715f;
    const float t = tanhf(a * (x + b * x * x * x));
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * a * (1.0f + 3.0f * b * x * x);
}
__device__ inline float softplus(float z) { return z > 20.0f ? z : log1pf(expf(z)); }
__device__ inline float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

__device__ inline float wave_sum(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
// sum over the 16 lanes that share l / 16
__device__ inline float group16_sum(float v) {
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
// sum over the 4 lane groups (same l % 16)
__device__ inline float across_groups(float v) {
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

// ---------------------------------------------------------------------------
// Feature kernel: one wave per element, lane = solution point.
// ---------------------------------------------------------------------------
struct FeatConsts {
    float D[16];    // D[i][m] = l_m'(x_i)
    float L[16];    // L[a][i] = (2a+1)/2 w_i P_a(x_i): nodal -> Legendre modal
    float w[4];     // GL weights
    float nrm[4];   // 2 / (2a+1)
    float jx, jy, jz;      // 2 / h_d
    float vol8;            // element volume / 8 (dV = vol8 dxi deta dzeta)
    float delta;           // (h_x h_y h_z)^(1/3) / 4
    float mu_mol;
};

template <class UT>
__global__ __launch_bounds__(256) void od_features_kernel(const UT* __restrict__ u, int64_t E, FeatConsts fc,
                                                          float* __restrict__ out) {
    __shared__ float sv[4][3][64];
    __shared__ float sm[4][3][64];
    const int wave = threadIdx.x >> 6, pt = threadIdx.x & 63;
    const int i = pt & 3, j = (pt >> 2) & 3, k = pt >> 4;
    const int64_t e = (int64_t)blockIdx.x * 4 + wave;
    const bool ok = e < E;
    const int64_t NP = E * 64;
    float rho = 1.0f, mx = 0.0f, my = 0.0f, mz = 0.0f, en = 2.5f;
    if (ok) {
        const int64_t q = e * 64 + pt;
        rho = (float)u[q];
        mx = (float)u[NP + q];
        my = (float)u[2 * NP + q];
        mz = (float)u[3 * NP + q];
        en = (float)u[4 * NP + q];
    }
    const float ux = mx / rho, uy = my / rho, uz = mz / rho;
    const float u2 = ux * ux + uy * uy + uz * uz;
    const float ke = 0.5f * rho * u2;
    const float p = 0.4f * (en - ke);
    const float c = sqrtf(fmaxf(1.4f * p / rho, 0.0f));
    sv[wave][0][pt] = ux;
    sv[wave][1][pt] = uy;
    sv[wave][2][pt] = uz;
    __syncthreads();
    // velocity gradient A[a][d] = du_a / dx_d
    float A[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int idx = d == 0 ? i : (d == 1 ? j : k);
        const int base = d == 0 ? 4 * j + 16 * k : (d == 1 ? i + 16 * k : i + 4 * j);
        const int stride = d == 0 ? 1 : (d == 1 ? 4 : 16);
        const float jac = d == 0 ? fc.jx : (d == 1 ? fc.jy : fc.jz);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float s = 0.0f;
#pragma unroll
            for (int m = 0; m < 4; ++m) s += fc.D[idx * 4 + m] * sv[wave][a][base + m * stride];
            A[a][d] = s * jac;
        }
    }
    // Legendre modal coefficients of the velocity: x, then y, then z
    float t[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float s = 0.0f;
#pragma unroll
        for (int m = 0; m < 4; ++m) s += fc.L[i * 4 + m] * sv[wave][a][m + 4 * j + 16 * k];
        t[a] = s;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) sm[wave][a][pt] = t[a];
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float s = 0.0f;
#pragma unroll
        for (int m = 0; m < 4; ++m) s += fc.L[j * 4 + m] * sm[wave][a][i + 4 * m + 16 * k];
        t[a] = s;
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 3; ++a) sv[wave][a][pt] = t[a];
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float s = 0.0f;
#pragma unroll
        for (int m = 0; m < 4; ++m) s += fc.L[k * 4 + m] * sv[wave][a][i + 4 * j + 16 * m];
        t[a] = s;
    }
    const float emode = (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) * fc.nrm[i] * fc.nrm[j] * fc.nrm[k];
    const int top = max(i, max(j, k));

    // invariants at the point
    const float div = A[0][0] + A[1][1] + A[2][2];
    float ss = 0.0f, oo = 0.0f, tr3 = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float sab = 0.5f * (A[a][b] + A[b][a]) - (a == b ? div / 3.0f : 0.0f);
            const float oab = 0.5f * (A[a][b] - A[b][a]);
            ss += sab * sab;
            oo += oab * oab;
#pragma unroll
            for (int q = 0; q < 3; ++q) tr3 += A[a][b] * A[b][q] * A[q][a];
        }
    const float smag = sqrtf(2.0f * ss), omag = sqrtf(2.0f * oo);
    float sfull = 0.0f;  // S_ij S_ij of the full strain for Q
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float sab = 0.5f * (A[a][b] + A[b][a]);
            sfull += sab * sab;
        }
    const float Q = 0.5f * (oo - sfull);
    const float R = -tr3 / 3.0f;

    const float w = fc.w[i] * fc.w[j] * fc.w[k];
    const float inv8 = 0.125f;  // 1 / sum of the weights
    const float s_rho = wave_sum(w * rho), s_p = wave_sum(w * p), s_c = wave_sum(w * c);
    const float s_S = wave_sum(w * smag), s_O = wave_sum(w * omag), s_Q = wave_sum(w * Q), s_R = wave_sum(w * R);
    const float s_div = wave_sum(w * div);
    const float s_ux = wave_sum(w * ux), s_uy = wave_sum(w * uy), s_uz = wave_sum(w * uz);
    const float s_ke = wave_sum(w * ke), s_SS = wave_sum(w * 2.0f * ss);
    // second moments about the element means (two passes, no cancellation)
    const float dr = rho - s_rho * inv8, dp = p - s_p * inv8;
    const float dux = ux - s_ux * inv8, duy = uy - s_uy * inv8, duz = uz - s_uz * inv8;
    const float v_rho = wave_sum(w * dr * dr), v_p = wave_sum(w * dp * dp);
    const float v_u = wave_sum(w * (dux * dux + duy * duy + duz * duz));
    const float e1 = wave_sum(top == 1 ? emode : 0.0f), e2 = wave_sum(top == 2 ? emode : 0.0f),
                e3 = wave_sum(top == 3 ? emode : 0.0f);
    if (!ok) return;
    const float rb = s_rho * inv8, pb = s_p * inv8, cb = s_c * inv8, dl = fc.delta;
    const float S = fmaxf(s_S * inv8, 1e-6f * cb / dl);
    const float etot = e1 + e2 + e3 + 1e-30f;
    float f[FS];
    f[0] = logf(rb * dl * dl * S / fc.mu_mol);
    f[1] = s_O * inv8 / S;
    f[2] = s_Q * inv8 / (S * S);
    f[3] = s_R * inv8 / (S * S * S);
    f[4] = s_div * inv8 / S;
    f[5] = e1 / etot;
    f[6] = e2 / etot;
    f[7] = e3 / etot;
    f[8] = sqrtf(v_rho * inv8) / rb;
    f[9] = sqrtf(v_p * inv8) / pb;
    f[10] = dl * S / cb;
    f[11] = sqrtf(v_u * inv8) / (dl * S);
    f[12] = rb;
    f[13] = dl * dl * S;
    f[14] = s_ke * fc.vol8;
    f[15] = s_SS * fc.vol8;
    float val = 0.0f;
#pragma unroll
    for (int q = 0; q < FS; ++q) val = pt == q ? f[q] : val;
    if (pt < FS) out[e * FS + pt] = val;
}

// ---------------------------------------------------------------------------
// MLP device weights: bf16 copies of W1 (padded to [64][16]), W2, W3, w4 and
// f32 biases, refreshed from the f32 master vector after each update.
// ---------------------------------------------------------------------------
struct DevNet {
    unsigned short* w1;  // [64][16]
    unsigned short* w2;  // [64][64]
    unsigned short* w3;  // [64][64]
    unsigned short* w4;  // [64]
    const float* params; // f32 master (biases read from it)
};

struct Scaling {
    float shift[NF], scale[NF], lo[NF], hi[NF];
    float out_scale;  // C = out_scale * softplus(z)
};

__global__ void od_mlp_cast_kernel(const float* __restrict__ p, unsigned short* w1, unsigned short* w2,
                                   unsigned short* w3, unsigned short* w4) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < NH * KP) {
        const int n = t / KP, kk = t % KP;
        w1[t] = kk < NF ? f2bf(p[oW1 + n * NF + kk]) : (unsigned short)0;
    }
    if (t < NH * NH) {
        w2[t] = f2bf(p[oW2 + t]);
        w3[t] = f2bf(p[oW3 + t]);
    }
    if (t < NH) w4[t] = f2bf(p[oW4 + t]);
}

// Weights of one work-group in LDS, and the transposed copies the backward
// pass multiplies by.
struct NetLds {
    unsigned short w1[NH * KP];
    unsigned short w2[NH * NH];
    unsigned short w3[NH * NH];
    float w4[NH];
    float b1[NH], b2[NH], b3[NH];
};
struct BackLds {
    unsigned short w2t[NH * NH];  // w2t[in][out] = W2[out][in]
    unsigned short w3t[NH * NH];
};

__device__ void load_net(const DevNet& net, NetLds& L, BackLds* T) {
    for (int q = threadIdx.x; q < NH * NH; q += blockDim.x) {
        const unsigned short a = net.w2[q], b = net.w3[q];
        L.w2[q] = a;
        L.w3[q] = b;
        if (T) {
            const int n = q / NH, kk = q % NH;
            T->w2t[kk * NH + n] = a;
            T->w3t[kk * NH + n] = b;
        }
    }
    for (int q = threadIdx.x; q < NH * KP; q += blockDim.x) L.w1[q] = net.w1[q];
    for (int q = threadIdx.x; q < NH; q += blockDim.x) {
        L.w4[q] = bf2f(net.w4[q]);
        L.b1[q] = net.params[oB1 + q];
        L.b2[q] = net.params[oB2 + q];
        L.b3[q] = net.params[oB3 + q];
    }
    __syncthreads();
}

__device__ inline bf16x4 lds4(const unsigned short* p) {
    bf16x4 r;
    r[0] = (short)p[0];
    r[1] = (short)p[1];
    r[2] = (short)p[2];
    r[3] = (short)p[3];
    return r;
}
__device__ inline bf16x4 lds4s(const unsigned short* p, int stride) {
    bf16x4 r;
    r[0] = (short)p[0];
    r[1] = (short)p[stride];
    r[2] = (short)p[2 * stride];
    r[3] = (short)p[3 * stride];
    return r;
}

// Z[16][64] = H[16][K] W^T + b with W [64][K] (row stride K) in LDS; z[nb][q]
// = Z[4g + q][16 nb + r].
template <int K>
__device__ inline void dense(const unsigned short* H, int ldh, const unsigned short* W, const float* b, int lane,
                             f32x4 z[4]) {
    const int r = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        const float bb = b[nb * 16 + r];
        z[nb] = f32x4{bb, bb, bb, bb};
    }
#pragma unroll
    for (int kb = 0; kb < K / 16; ++kb) {
        const bf16x4 a = lds4(H + r * ldh + kb * 16 + kq);
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) z[nb] = mfma(a, lds4(W + (nb * 16 + r) * K + kb * 16 + kq), z[nb]);
    }
}

// h = GELU(z) rounded to bf16 into Hout[16][LDH].
__device__ inline void act_store(const f32x4 z[4], unsigned short* Hout, int lane) {
    const int r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
        for (int q = 0; q < 4; ++q) Hout[(4 * g + q) * LDH + nb * 16 + r] = f2bf(gelu(z[nb][q]));
}

// Standardised input tile (bf16 [16][16]) of elements e0 .. e0+15 (rows
// beyond n are zero).  x is [n][xs] raw features, idx an optional gather.
__device__ inline void load_input(const float* x, int xs, const int* idx, int64_t e0, int64_t n, const Scaling& sc,
                                  unsigned short* H0, int lane) {
    for (int q = lane; q < 16 * KP; q += 64) {
        const int row = q / KP, col = q % KP;
        const int64_t e = e0 + row;
        float v = 0.0f;
        if (e < n && col < NF) {
            const int64_t src = idx ? (int64_t)idx[e] : e;
            v = (x[src * xs + col] - sc.shift[col]) * sc.scale[col];
        }
        H0[row * KP + col] = f2bf(v);
    }
}

// Output layer: z4[q] for element 4g + q (all lanes of the group hold it).
__device__ inline void out_layer(const unsigned short* H3, const float* w4, float b4, int lane, float z4[4]) {
    const int r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float s = 0.0f;
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) s += bf2f(H3[(4 * g + q) * LDH + nb * 16 + r]) * w4[nb * 16 + r];
        z4[q] = group16_sum(s) + b4;
    }
}

struct WaveLds {
    unsigned short h0[16 * KP];
    unsigned short h1[16 * LDH];
    unsigned short h2[16 * LDH];
    unsigned short h3[16 * LDH];
    unsigned short gz[16 * LDH];  // bf16 dZ of the current layer
};

// ---- inference: mu_t per element of a solver context -----------------------
template <class CT>
__global__ __launch_bounds__(256) void od_mlp_infer_kernel(DevNet net, Scaling sc, const float* __restrict__ feat,
                                                           int64_t E, CT* __restrict__ mu_t,
                                                           float* __restrict__ eps, float* __restrict__ out_of_range) {
    __shared__ NetLds L;
    __shared__ WaveLds W[4];
    load_net(net, L, nullptr);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float b4 = net.params[oB4];
    const int64_t tiles = (E + 15) / 16;
    for (int64_t tile = (int64_t)blockIdx.x * 4 + wave; tile < tiles; tile += (int64_t)gridDim.x * 4) {
        const int64_t e0 = tile * 16;
        WaveLds& S = W[wave];
        load_input(feat, FS, nullptr, e0, E, sc, S.h0, lane);
        f32x4 z[4];
        dense<KP>(S.h0, KP, L.w1, L.b1, lane, z);
        act_store(z, S.h1, lane);
        dense<NH>(S.h1, LDH, L.w2, L.b2, lane, z);
        act_store(z, S.h2, lane);
        dense<NH>(S.h2, LDH, L.w3, L.b3, lane, z);
        act_store(z, S.h3, lane);
        float z4[4];
        out_layer(S.h3, L.w4, b4, lane, z4);
        const int r = lane & 15, g = lane >> 4;
        if (r == 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t e = e0 + 4 * g + q;
                if (e >= E) continue;
                const float* fe = feat + e * FS;
                bool in = true;
#pragma unroll
                for (int c = 0; c < NF; ++c) in = in && fe[c] >= sc.lo[c] && fe[c] <= sc.hi[c];
                const float m = in ? fe[12] * fe[13] * (sc.out_scale * softplus(z4[q])) : 0.0f;
                mu_t[e] = (CT)m;
                eps[e] = m * fe[15];
                out_of_range[e] = in ? 0.0f : 1.0f;
            }
        }
    }
}

// ---- forward on host-given features (testing / evaluation) ----------------
__global__ __launch_bounds__(256) void od_mlp_forward_kernel(DevNet net, Scaling sc, const float* __restrict__ x,
                                                             int64_t n, float* __restrict__ y) {
    __shared__ NetLds L;
    __shared__ WaveLds W[4];
    load_net(net, L, nullptr);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float b4 = net.params[oB4];
    const int64_t tiles = (n + 15) / 16;
    for (int64_t tile = (int64_t)blockIdx.x * 4 + wave; tile < tiles; tile += (int64_t)gridDim.x * 4) {
        const int64_t e0 = tile * 16;
        WaveLds& S = W[wave];
        load_input(x, NF, nullptr, e0, n, sc, S.h0, lane);
        f32x4 z[4];
        dense<KP>(S.h0, KP, L.w1, L.b1, lane, z);
        act_store(z, S.h1, lane);
        dense<NH>(S.h1, LDH, L.w2, L.b2, lane, z);
        act_store(z, S.h2, lane);
        dense<NH>(S.h2, LDH, L.w3, L.b3, lane, z);
        act_store(z, S.h3, lane);
        float z4[4];
        out_layer(S.h3, L.w4, b4, lane, z4);
        const int r = lane & 15, g = lane >> 4;
        if (r == 0)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (e0 + 4 * g + q < n) y[e0 + 4 * g + q] = sc.out_scale * softplus(z4[q]);
    }
}

// ---- backward ----------------------------------------------------------------
// dZ (D layout, f32) -> bf16 [16][LDH] into gz; bias gradient accumulated.
__device__ inline void store_dz(const f32x4 dz[4], unsigned short* G, int lane, float db[4]) {
    const int r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        float s = 0.0f;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            G[(4 * g + q) * LDH + nb * 16 + r] = f2bf(dz[nb][q]);
            s += dz[nb][q];
        }
        db[nb] += across_groups(s);
    }
}

// dW[out][in] += sum_e dZ[e][out] H[e][in] over the tile's 16 elements:
// A = dZ^T (row out, k e), B = H (k e, column in).
template <int NIN>
__device__ inline void weight_grad(const unsigned short* G, const unsigned short* Hin, int ldh, int lane,
                                   f32x4 acc[4][NIN / 16]) {
    const int r = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
        const bf16x4 a = lds4s(G + kq * LDH + rb * 16 + r, LDH);
#pragma unroll
        for (int cb = 0; cb < NIN / 16; ++cb) acc[rb][cb] = mfma(a, lds4s(Hin + kq * ldh + cb * 16 + r, ldh), acc[rb][cb]);
    }
}

// dH[16][NIN] = dZ[16][64] W (W [64][NIN]) using WT [NIN][64]; D layout out.
template <int NIN>
__device__ inline void back_dense(const unsigned short* G, const unsigned short* WT, int lane, f32x4 dh[NIN / 16]) {
    const int r = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
    for (int nb = 0; nb < NIN / 16; ++nb) dh[nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        const bf16x4 a = lds4(G + r * LDH + kb * 16 + kq);
#pragma unroll
        for (int nb = 0; nb < NIN / 16; ++nb) dh[nb] = mfma(a, lds4(WT + (nb * 16 + r) * NH + kb * 16 + kq), dh[nb]);
    }
}

// Per-wave partial gradients over a contiguous range of 16-element tiles of
// the (gathered) batch; partial[wave][NPARAM + 1] (last entry: sum of the
// squared errors).  Loss = mean over the batch of (C - t)^2.
__global__ __launch_bounds__(128) void od_mlp_grad_kernel(DevNet net, Scaling sc, const float* __restrict__ x,
                                                          const float* __restrict__ tgt, const int* __restrict__ idx,
                                                          int64_t n, int tiles_per_wave, float* __restrict__ partial) {
    __shared__ NetLds L;
    __shared__ BackLds T;
    __shared__ WaveLds W[2];
    load_net(net, L, &T);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    const int64_t gw = (int64_t)blockIdx.x * 2 + wave;
    const float b4 = net.params[oB4];
    const float inv_n = 1.0f / (float)n;
    f32x4 aW1[4][1], aW2[4][4], aW3[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        aW1[a][0] = f32x4{0, 0, 0, 0};
#pragma unroll
        for (int b = 0; b < 4; ++b) aW2[a][b] = aW3[a][b] = f32x4{0, 0, 0, 0};
    }
    float db1[4] = {0, 0, 0, 0}, db2[4] = {0, 0, 0, 0}, db3[4] = {0, 0, 0, 0}, dw4[4] = {0, 0, 0, 0};
    float db4 = 0.0f, sse = 0.0f;
    WaveLds& S = W[wave];
    for (int tt = 0; tt < tiles_per_wave; ++tt) {
        const int64_t e0 = (gw * tiles_per_wave + tt) * 16;
        if (e0 >= n) break;
        load_input(x, NF, idx, e0, n, sc, S.h0, lane);
        f32x4 z1[4], z2[4], z3[4];
        dense<KP>(S.h0, KP, L.w1, L.b1, lane, z1);
        act_store(z1, S.h1, lane);
        dense<NH>(S.h1, LDH, L.w2, L.b2, lane, z2);
        act_store(z2, S.h2, lane);
        dense<NH>(S.h2, LDH, L.w3, L.b3, lane, z3);
        act_store(z3, S.h3, lane);
        float z4[4];
        out_layer(S.h3, L.w4, b4, lane, z4);
        // dL/dz4 per element of the group
        float dz4[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t e = e0 + 4 * g + q;
            float d = 0.0f;
            if (e < n) {
                const int64_t src = idx ? (int64_t)idx[e] : e;
                const float y = sc.out_scale * softplus(z4[q]), err = y - tgt[src];
                if (r == 0) sse += err * err;
                d = 2.0f * err * inv_n * sc.out_scale * sigmoid(z4[q]);
            }
            dz4[q] = d;
        }
        // output layer gradients
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            float s = 0.0f;
#pragma unroll
            for (int q = 0; q < 4; ++q) s += dz4[q] * bf2f(S.h3[(4 * g + q) * LDH + nb * 16 + r]);
            dw4[nb] += across_groups(s);
        }
        {
            float s = 0.0f;
#pragma unroll
            for (int q = 0; q < 4; ++q) s += dz4[q];
            db4 += across_groups(r == 0 ? s : 0.0f);
        }
        // layer 3
        f32x4 dz[4];
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
            for (int q = 0; q < 4; ++q) dz[nb][q] = dz4[q] * L.w4[nb * 16 + r] * gelu_d(z3[nb][q]);
        store_dz(dz, S.gz, lane, db3);
        weight_grad<NH>(S.gz, S.h2, LDH, lane, aW3);
        f32x4 dh[4];
        back_dense<NH>(S.gz, T.w3t, lane, dh);
        // layer 2
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
            for (int q = 0; q < 4; ++q) dz[nb][q] = dh[nb][q] * gelu_d(z2[nb][q]);
        store_dz(dz, S.gz, lane, db2);
        weight_grad<NH>(S.gz, S.h1, LDH, lane, aW2);
        back_dense<NH>(S.gz, T.w2t, lane, dh);
        // layer 1
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
            for (int q = 0; q < 4; ++q) dz[nb][q] = dh[nb][q] * gelu_d(z1[nb][q]);
        store_dz(dz, S.gz, lane, db1);
        weight_grad<KP>(S.gz, S.h0, KP, lane, aW1);
    }
    float* out = partial + gw * (NPARAM + 1);
    // weight matrices: acc[rb][cb][q] = dW[out = 16 rb + 4g + q][in = 16 cb + r]
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int o = 16 * rb + 4 * g + q;
            if (r < NF) out[oW1 + o * NF + r] = aW1[rb][0][q];
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
                out[oW2 + o * NH + 16 * cb + r] = aW2[rb][cb][q];
                out[oW3 + o * NH + 16 * cb + r] = aW3[rb][cb][q];
            }
        }
    if (g == 0) {
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            out[oB1 + nb * 16 + r] = db1[nb];
            out[oB2 + nb * 16 + r] = db2[nb];
            out[oB3 + nb * 16 + r] = db3[nb];
            out[oW4 + nb * 16 + r] = dw4[nb];
        }
    }
    const float tot = wave_sum(sse);
    if (lane == 0) {
        out[oB4] = db4;
        out[NPARAM] = tot;
    }
}

// grad[p] = sum over waves of partial[w][p], in wave order (deterministic).
__global__ void od_mlp_reduce_kernel(const float* __restrict__ partial, int waves, float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > NPARAM) return;
    float s = 0.0f;
    for (int w = 0; w < waves; ++w) s += partial[(int64_t)w * (NPARAM + 1) + p];
    grad[p] = s;
}

__global__ void od_mlp_adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                   float* __restrict__ v, float lr, float b1t, float b2t) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NPARAM) return;
    const float gg = g[q];
    const float mm = 0.9f * m[q] + 0.1f * gg;
    const float vv = 0.999f * v[q] + 0.001f * gg * gg;
    m[q] = mm;
    v[q] = vv;
    const float mh = mm / (1.0f - b1t), vh = vv / (1.0f - b2t);
    p[q] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

// Deterministic sums of three float arrays of length n (one block of 256,
// fixed order): out[0..2].
__global__ __launch_bounds__(256) void od_sum3_kernel(const float* __restrict__ a, int as, const float* __restrict__ b,
                                                      const float* __restrict__ c, int64_t n, double* __restrict__ out) {
    __shared__ double ra[256], rb[256], rc[256];
    double sa = 0.0, sb = 0.0, sc = 0.0;
    for (int64_t q = threadIdx.x; q < n; q += 256) {
        sa += (double)a[q * as];
        sb += (double)b[q];
        sc += (double)c[q];
    }
    ra[threadIdx.x] = sa;
    rb[threadIdx.x] = sb;
    rc[threadIdx.x] = sc;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h) {
            ra[threadIdx.x] += ra[threadIdx.x + h];
            rb[threadIdx.x] += rb[threadIdx.x + h];
            rc[threadIdx.x] += rc[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = ra[0];
        out[1] = rb[0];
        out[2] = rc[0];
    }
}

template <class CT>
__global__ void od_fill_kernel(CT* p, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = (CT)0;
}

int check() { return hipGetLastError() == hipSuccess ? 0 : 1; }

}  // namespace

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
struct od_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float* params = nullptr;  // f32 master
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* grad = nullptr;    // NPARAM + 1
    unsigned short* wb = nullptr;  // bf16 copies: w1 | w2 | w3 | w4
    Scaling sc{};
    std::vector<float> host_params;
    uint64_t adam_t = 0;
};

namespace {

DevNet dev_net(const od_mlp* m) {
    DevNet n;
    n.w1 = m->wb;
    n.w2 = m->wb + NH * KP;
    n.w3 = n.w2 + NH * NH;
    n.w4 = n.w3 + NH * NH;
    n.params = m->params;
    return n;
}

int cast_weights(od_mlp* m) {
    DevNet n = dev_net(m);
    hipLaunchKernelGGL(od_mlp_cast_kernel, dim3((NH * NH + 255) / 256), dim3(256), 0, m->stream, m->params, n.w1,
                       n.w2, n.w3, n.w4);
    return check();
}

struct DevBuf {  // scoped device allocation
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// Gradient of the mean squared error over n gathered samples; grad and the
// sum of squared errors in m->grad (device).
od_status run_grad(od_mlp* m, const float* dx, const float* dt, const int* didx, int64_t n, float** partial,
                   size_t* partial_cap) {
    const int64_t tiles = (n + 15) / 16;
    const int tpw = (int)std::max<int64_t>(1, (tiles + 1023) / 1024);
    const int64_t waves = (tiles + tpw - 1) / tpw;
    const int64_t blocks = (waves + 1) / 2;  // two waves per work-group
    const size_t need = (size_t)blocks * 2 * (NPARAM + 1) * sizeof(float);
    if (need > *partial_cap) {
        if (*partial) (void)hipFree(*partial);
        *partial = nullptr;
        *partial_cap = 0;
        HIP_TRY(hipMalloc((void**)partial, need));
        *partial_cap = need;
    }
    // every wave (also one past the data) writes its whole partial row
    hipLaunchKernelGGL(od_mlp_grad_kernel, dim3((unsigned)blocks), dim3(128), 0, m->stream, dev_net(m), m->sc, dx, dt,
                       didx, n, tpw, *partial);
    if (check()) return fail(OD_ERR_HIP, "od_mlp_grad_kernel launch failed");
    hipLaunchKernelGGL(od_mlp_reduce_kernel, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, m->stream, *partial,
                       (int)(blocks * 2), m->grad);
    if (check()) return fail(OD_ERR_HIP, "od_mlp_reduce_kernel launch failed");
    return OD_OK;
}

// Host f64 FR operators for the feature kernel.
FeatConsts feat_consts(const od_ctx* c) {
    FeatConsts f{};
    const HostOps& o = c->ops;
    for (int q = 0; q < 16; ++q) f.D[q] = (float)o.D[q];
    for (int a = 0; a < 4; ++a) {
        for (int i = 0; i < 4; ++i) {
            // Legendre P_a at x_i
            const double x = o.x[i];
            double p0 = 1.0, p1 = x, pa = a == 0 ? 1.0 : x;
            for (int kk = 2; kk <= a; ++kk) {
                const double p2 = ((2.0 * kk - 1.0) * x * p1 - (kk - 1.0) * p0) / kk;
                p0 = p1;
                p1 = p2;
                pa = p2;
            }
            f.L[a * 4 + i] = (float)((2.0 * a + 1.0) / 2.0 * o.w[i] * pa);
        }
        f.w[a] = (float)o.w[a];
        f.nrm[a] = (float)(2.0 / (2.0 * a + 1.0));
    }
    const double hx = c->p.lx / c->p.nx, hy = c->p.ly / c->p.ny, hz = c->p.lz / c->p.nz;
    f.jx = (float)(2.0 / hx);
    f.jy = (float)(2.0 / hy);
    f.jz = (float)(2.0 / hz);
    f.vol8 = (float)(hx * hy * hz / 8.0);
    f.delta = (float)(std::cbrt(hx * hy * hz) / 4.0);
    f.mu_mol = (float)(c->p.mu_ref > 0 ? c->p.mu_ref : 1e-30);
    return f;
}

od_status ensure_closure_buffers(od_ctx* c) {
    if (!c->feat) {
        if (hipMalloc(&c->feat, (size_t)c->E * (FS + 2) * sizeof(float)) != hipSuccess)
            return fail(OD_ERR_NOMEM, "closure buffers: allocation failed");
        c->device_bytes += (size_t)c->E * (FS + 2) * sizeof(float);
    }
    if (!c->red) {
        HIP_TRY(hipMalloc((void**)&c->red, 4 * sizeof(double)));
        HIP_TRY(hipHostMalloc((void**)&c->red_host, 4 * sizeof(double), hipHostMallocDefault));
    }
    return OD_OK;
}

od_status ensure_mu_t(od_ctx* c) {
    if (c->mu_t) return OD_OK;
    if (hipMalloc(&c->mu_t, (size_t)c->E * c->cbytes) != hipSuccess)
        return fail(OD_ERR_NOMEM, "eddy viscosity buffer: allocation failed");
    c->device_bytes += (size_t)c->E * c->cbytes;
    const unsigned blocks = (unsigned)((c->E + 255) / 256);
    if (c->cbytes == 4)
        hipLaunchKernelGGL(od_fill_kernel<float>, dim3(blocks), dim3(256), 0, c->stream, (float*)c->mu_t, c->E);
    else
        hipLaunchKernelGGL(od_fill_kernel<double>, dim3(blocks), dim3(256), 0, c->stream, (double*)c->mu_t, c->E);
    return check() ? fail(OD_ERR_HIP, "fill kernel launch failed") : OD_OK;
}

int launch_features(od_ctx* c) {
    const FeatConsts fc = feat_consts(c);
    const unsigned blocks = (unsigned)((c->E + 3) / 4);
    if (c->p.u_real == OD_F32)
        hipLaunchKernelGGL(od_features_kernel<float>, dim3(blocks), dim3(256), 0, c->stream, (const float*)c->u[0],
                           c->E, fc, (float*)c->feat);
    else
        hipLaunchKernelGGL(od_features_kernel<double>, dim3(blocks), dim3(256), 0, c->stream,
                           (const double*)c->u[0], c->E, fc, (float*)c->feat);
    return check();
}

}  // namespace

// Refresh mu_t from the current state (called by od_step every
// update_interval steps while a network is attached).  Defined here, used by
// od_api.hip.
od_status od_closure_update_internal(od_ctx* c, double dt_interval) {
    od_status st = ensure_closure_buffers(c);
    if (st == OD_OK) st = ensure_mu_t(c);
    if (st != OD_OK) return st;
    if (launch_features(c)) return fail(OD_ERR_HIP, "od_features_kernel launch failed");
    float* eps = (float*)c->feat + (size_t)c->E * FS;
    float* oor = eps + c->E;
    const unsigned blocks = (unsigned)std::min<int64_t>(1024, (c->E + 63) / 64);
    const DevNet net = dev_net(c->mlp);
    if (c->cbytes == 4)
        hipLaunchKernelGGL(od_mlp_infer_kernel<float>, dim3(blocks), dim3(256), 0, c->stream, net, c->mlp->sc,
                           (const float*)c->feat, c->E, (float*)c->mu_t, eps, oor);
    else
        hipLaunchKernelGGL(od_mlp_infer_kernel<double>, dim3(blocks), dim3(256), 0, c->stream, net, c->mlp->sc,
                           (const float*)c->feat, c->E, (double*)c->mu_t, eps, oor);
    if (check()) return fail(OD_ERR_HIP, "od_mlp_infer_kernel launch failed");
    // kinetic energy, model dissipation rate and out-of-range count (exact in f64)
    hipLaunchKernelGGL(od_sum3_kernel, dim3(1), dim3(256), 0, c->stream, (const float*)c->feat + 14, FS,
                       (const float*)eps, (const float*)oor, c->E, c->red);
    if (check()) return fail(OD_ERR_HIP, "od_sum3_kernel launch failed");
    HIP_TRY(hipMemcpyAsync(c->red_host, c->red, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const double ke = c->red_host[0], eps_now = c->red_host[1];
    const int64_t n_oor = (int64_t)c->red_host[2];
    // budget check over the interval that just ended
    bool active = true;
    if (c->closure.budget_kappa > 0.0 && c->ke_prev >= 0.0 && c->closure_active) {
        // The model over-dissipates when, over an interval in which the kinetic
        // energy decayed, the energy it removed exceeds kappa times the whole
        // observed decay.  An interval in which the energy grew is not held
        // against the model (it only removes energy).
        const double dke = c->ke_prev - ke;               // observed decay
        const double work = c->eps_prev * dt_interval;    // energy the model removed
        if (dke > 0.0 && work > c->closure.budget_kappa * dke) active = false;
    }
    if (!active) {
        const unsigned fb = (unsigned)((c->E + 255) / 256);
        if (c->cbytes == 4)
            hipLaunchKernelGGL(od_fill_kernel<float>, dim3(fb), dim3(256), 0, c->stream, (float*)c->mu_t, c->E);
        else
            hipLaunchKernelGGL(od_fill_kernel<double>, dim3(fb), dim3(256), 0, c->stream, (double*)c->mu_t, c->E);
        if (check()) return fail(OD_ERR_HIP, "fill kernel launch failed");
        c->closure_stats[1] += 1.0;
    }
    c->closure_active = active;
    c->ke_prev = ke;
    c->eps_prev = active ? eps_now : 0.0;
    c->closure_stats[0] += 1.0;
    c->closure_stats[2] += (double)n_oor;
    c->closure_stats[3] = ke;
    c->closure_stats[4] = eps_now;
    c->closure_stats[5] = active ? 1.0 : 0.0;
    return OD_OK;
}

void od_closure_release_internal(od_ctx* c) {
    if (c->feat) (void)hipFree(c->feat);
    if (c->red) (void)hipFree(c->red);
    if (c->red_host) (void)hipHostFree(c->red_host);
    if (c->mu_t) (void)hipFree(c->mu_t);
    c->feat = nullptr;
    c->red = nullptr;
    c->red_host = nullptr;
    c->mu_t = nullptr;
}

extern "C" {

od_status od_mlp_create(int device, od_mlp** out) {
    if (!out) return fail(OD_ERR_INVALID, "od_mlp_create: null argument");
    *out = nullptr;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(OD_ERR_INVALID, "no such HIP device");
    HIP_TRY(hipSetDevice(device));
    od_mlp* m = new od_mlp();
    m->device = device;
    auto bad = [&](const char* what) {
        od_mlp_destroy(m);
        return fail(OD_ERR_NOMEM, what);
    };
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) return bad("stream creation failed");
    if (hipMalloc((void**)&m->params, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&m->adam_m, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&m->adam_v, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&m->grad, (NPARAM + 1) * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&m->wb, (NH * KP + 2 * NH * NH + NH) * sizeof(unsigned short)) != hipSuccess)
        return bad("device allocation failed");
    // Default: Glorot-uniform weights, zero biases (seed 0); identity scaling, open range.
    std::vector<float> p(NPARAM, 0.0f);
    std::mt19937_64 rng(0);
    auto fill = [&](int off, int fan_out, int fan_in) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        std::uniform_real_distribution<double> U(-a, a);
        for (int q = 0; q < fan_out * fan_in; ++q) p[off + q] = (float)U(rng);
    };
    fill(oW1, NH, NF);
    fill(oW2, NH, NH);
    fill(oW3, NH, NH);
    fill(oW4, 1, NH);
    for (int q = 0; q < NF; ++q) {
        m->sc.shift[q] = 0.0f;
        m->sc.scale[q] = 1.0f;
        m->sc.lo[q] = -INFINITY;
        m->sc.hi[q] = INFINITY;
    }
    m->sc.out_scale = 1.0f;
    od_status s = od_mlp_set_params(m, p.data(), p.size());
    if (s != OD_OK) {
        od_mlp_destroy(m);
        return s;
    }
    *out = m;
    return OD_OK;
}

void od_mlp_destroy(od_mlp* m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void* q : {(void*)m->params, (void*)m->adam_m, (void*)m->adam_v, (void*)m->grad, (void*)m->wb})
        if (q) (void)hipFree(q);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

od_status od_mlp_set_params(od_mlp* m, const float* params, size_t n) {
    if (!m || !params) return fail(OD_ERR_INVALID, "od_mlp_set_params: null argument");
    if (n != (size_t)NPARAM) return fail(OD_ERR_INVALID, "od_mlp_set_params: expected OD_MLP_PARAMS values");
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipStreamSynchronize(m->stream));
    HIP_TRY(hipMemcpy(m->params, params, NPARAM * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemsetAsync(m->adam_m, 0, NPARAM * sizeof(float), m->stream));
    HIP_TRY(hipMemsetAsync(m->adam_v, 0, NPARAM * sizeof(float), m->stream));
    m->adam_t = 0;
    if (cast_weights(m)) return fail(OD_ERR_HIP, "od_mlp_cast_kernel launch failed");
    HIP_TRY(hipStreamSynchronize(m->stream));
    return OD_OK;
}

od_status od_mlp_get_params(od_mlp* m, float* params, size_t n) {
    if (!m || !params) return fail(OD_ERR_INVALID, "od_mlp_get_params: null argument");
    if (n != (size_t)NPARAM) return fail(OD_ERR_INVALID, "od_mlp_get_params: expected OD_MLP_PARAMS values");
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipStreamSynchronize(m->stream));
    HIP_TRY(hipMemcpy(params, m->params, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return OD_OK;
}

od_status od_mlp_set_scaling(od_mlp* m, const float shift[12], const float scale[12], const float lo[12],
                             const float hi[12]) {
    if (!m || !shift || !scale || !lo || !hi) return fail(OD_ERR_INVALID, "od_mlp_set_scaling: null argument");
    for (int q = 0; q < NF; ++q) {
        if (!(lo[q] <= hi[q])) return fail(OD_ERR_INVALID, "od_mlp_set_scaling: lo must be <= hi");
        m->sc.shift[q] = shift[q];
        m->sc.scale[q] = scale[q];
        m->sc.lo[q] = lo[q];
        m->sc.hi[q] = hi[q];
    }
    return OD_OK;
}

od_status od_mlp_set_output_scale(od_mlp* m, float s) {
    if (!m) return fail(OD_ERR_INVALID, "od_mlp_set_output_scale: null argument");
    if (!(s > 0.0f) || !std::isfinite(s)) return fail(OD_ERR_INVALID, "od_mlp_set_output_scale: scale must be positive");
    m->sc.out_scale = s;
    return OD_OK;
}

od_status od_mlp_forward(od_mlp* m, const float* x, size_t n, float* y) {
    if (!m || !x || !y) return fail(OD_ERR_INVALID, "od_mlp_forward: null argument");
    if (n == 0) return OD_OK;
    HIP_TRY(hipSetDevice(m->device));
    DevBuf dx, dy;
    HIP_TRY(hipMalloc(&dx.p, n * NF * sizeof(float)));
    HIP_TRY(hipMalloc(&dy.p, n * sizeof(float)));
    HIP_TRY(hipMemcpyAsync(dx.p, x, n * NF * sizeof(float), hipMemcpyHostToDevice, m->stream));
    const unsigned blocks = (unsigned)std::min<size_t>(1024, (n + 63) / 64);
    hipLaunchKernelGGL(od_mlp_forward_kernel, dim3(blocks), dim3(256), 0, m->stream, dev_net(m), m->sc,
                       (const float*)dx.p, (int64_t)n, (float*)dy.p);
    if (check()) return fail(OD_ERR_HIP, "od_mlp_forward_kernel launch failed");
    HIP_TRY(hipMemcpyAsync(y, dy.p, n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    return OD_OK;
}

od_status od_mlp_loss_grad(od_mlp* m, const float* x, const float* t, size_t n, double* loss, float* grad) {
    if (!m || !x || !t) return fail(OD_ERR_INVALID, "od_mlp_loss_grad: null argument");
    if (n == 0) return fail(OD_ERR_INVALID, "od_mlp_loss_grad: empty batch");
    HIP_TRY(hipSetDevice(m->device));
    DevBuf dx, dt;
    HIP_TRY(hipMalloc(&dx.p, n * NF * sizeof(float)));
    HIP_TRY(hipMalloc(&dt.p, n * sizeof(float)));
    HIP_TRY(hipMemcpyAsync(dx.p, x, n * NF * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIP_TRY(hipMemcpyAsync(dt.p, t, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
    float* partial = nullptr;
    size_t cap = 0;
    od_status s = run_grad(m, (const float*)dx.p, (const float*)dt.p, nullptr, (int64_t)n, &partial, &cap);
    std::vector<float> g(NPARAM + 1);
    if (s == OD_OK && hipMemcpyAsync(g.data(), m->grad, (NPARAM + 1) * sizeof(float), hipMemcpyDeviceToHost,
                                     m->stream) != hipSuccess)
        s = fail(OD_ERR_HIP, "od_mlp_loss_grad: copy failed");
    (void)hipStreamSynchronize(m->stream);
    if (partial) (void)hipFree(partial);
    if (s != OD_OK) return s;
    if (loss) *loss = (double)g[NPARAM] / (double)n;
    if (grad) std::memcpy(grad, g.data(), NPARAM * sizeof(float));
    return OD_OK;
}

od_status od_mlp_train(od_mlp* m, const float* x, const float* t, size_t n, uint32_t epochs, uint32_t batch,
                       float lr, uint64_t seed, float* loss_per_epoch) {
    if (!m || !x || !t) return fail(OD_ERR_INVALID, "od_mlp_train: null argument");
    if (n == 0 || batch == 0) return fail(OD_ERR_INVALID, "od_mlp_train: empty data or batch");
    if (n > (size_t)0x7fffffff) return fail(OD_ERR_INVALID, "od_mlp_train: too many samples");
    HIP_TRY(hipSetDevice(m->device));
    DevBuf dx, dt, di;
    HIP_TRY(hipMalloc(&dx.p, n * NF * sizeof(float)));
    HIP_TRY(hipMalloc(&dt.p, n * sizeof(float)));
    HIP_TRY(hipMalloc(&di.p, n * sizeof(int)));
    HIP_TRY(hipMemcpyAsync(dx.p, x, n * NF * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIP_TRY(hipMemcpyAsync(dt.p, t, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
    std::vector<int> perm(n);
    std::mt19937_64 rng(seed);
    float* partial = nullptr;
    size_t cap = 0;
    std::vector<float> g(NPARAM + 1);
    od_status s = OD_OK;
    for (uint32_t ep = 0; ep < epochs && s == OD_OK; ++ep) {
        for (size_t q = 0; q < n; ++q) perm[q] = (int)q;
        for (size_t q = n - 1; q > 0; --q) {  // Fisher-Yates with the seeded generator
            const size_t r = (size_t)(rng() % (q + 1));
            std::swap(perm[q], perm[r]);
        }
        HIP_TRY(hipStreamSynchronize(m->stream));  // the previous epoch's gathers are done with the index
        HIP_TRY(hipMemcpy(di.p, perm.data(), n * sizeof(int), hipMemcpyHostToDevice));
        double sse = 0.0;
        for (size_t b0 = 0; b0 < n && s == OD_OK; b0 += batch) {
            const int64_t nb = (int64_t)std::min<size_t>(batch, n - b0);
            s = run_grad(m, (const float*)dx.p, (const float*)dt.p, (const int*)di.p + b0, nb, &partial, &cap);
            if (s != OD_OK) break;
            m->adam_t += 1;
            const float b1t = (float)std::pow(0.9, (double)m->adam_t), b2t = (float)std::pow(0.999, (double)m->adam_t);
            hipLaunchKernelGGL(od_mlp_adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, m->params,
                               m->grad, m->adam_m, m->adam_v, lr, b1t, b2t);
            if (check() || cast_weights(m)) s = fail(OD_ERR_HIP, "optimizer kernel launch failed");
            if (loss_per_epoch && s == OD_OK) {
                float v = 0.0f;
                if (hipMemcpyAsync(&v, m->grad + NPARAM, sizeof(float), hipMemcpyDeviceToHost, m->stream) !=
                        hipSuccess ||
                    hipStreamSynchronize(m->stream) != hipSuccess)
                    s = fail(OD_ERR_HIP, "od_mlp_train: loss copy failed");
                sse += v;
            }
        }
        if (loss_per_epoch) loss_per_epoch[ep] = (float)(sse / (double)n);
    }
    (void)hipStreamSynchronize(m->stream);
    if (partial) (void)hipFree(partial);
    return s;
}

od_status od_closure_features(od_ctx* c, float* out, size_t n) {
    if (!c || !out) return fail(OD_ERR_INVALID, "od_closure_features: null argument");
    if (n != (size_t)c->E * FS) return fail(OD_ERR_INVALID, "od_closure_features: expected E * 16 values");
    if (!c->state_set) return fail(OD_ERR_STATE, "od_closure_features: state not set");
    HIP_TRY(hipSetDevice(c->device));
    od_status s = ensure_closure_buffers(c);
    if (s != OD_OK) return s;
    if (launch_features(c)) return fail(OD_ERR_HIP, "od_features_kernel launch failed");
    HIP_TRY(hipMemcpyAsync(out, c->feat, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return OD_OK;
}

od_status od_set_eddy_viscosity(od_ctx* c, const void* host, size_t n, od_real dtype) {
    if (!c) return fail(OD_ERR_INVALID, "od_set_eddy_viscosity: null context");
    if (c->mlp) return fail(OD_ERR_STATE, "od_set_eddy_viscosity: a network is attached");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!host) {
        c->mu_t_on = false;
        return OD_OK;
    }
    if (n != (size_t)c->E) return fail(OD_ERR_INVALID, "od_set_eddy_viscosity: expected E values");
    if (!real_ok(dtype)) return fail(OD_ERR_INVALID, "od_set_eddy_viscosity: bad dtype");
    if (!c->p.viscous) return fail(OD_ERR_STATE, "od_set_eddy_viscosity: Euler mode has no viscous fluxes");
    od_status s = ensure_mu_t(c);
    if (s != OD_OK) return s;
    std::vector<unsigned char> tmp(n * c->cbytes);
    if (c->cbytes == 4) {
        if (dtype == OD_F32)
            std::memcpy(tmp.data(), host, n * 4);
        else
            convert((float*)tmp.data(), (const double*)host, n);
    } else {
        if (dtype == OD_F64)
            std::memcpy(tmp.data(), host, n * 8);
        else
            convert((double*)tmp.data(), (const float*)host, n);
    }
    HIP_TRY(hipMemcpy(c->mu_t, tmp.data(), n * c->cbytes, hipMemcpyHostToDevice));
    c->mu_t_on = true;
    return OD_OK;
}

od_status od_get_eddy_viscosity(od_ctx* c, void* host, size_t n, od_real dtype) {
    if (!c || !host) return fail(OD_ERR_INVALID, "od_get_eddy_viscosity: null argument");
    if (n != (size_t)c->E) return fail(OD_ERR_INVALID, "od_get_eddy_viscosity: expected E values");
    if (!real_ok(dtype)) return fail(OD_ERR_INVALID, "od_get_eddy_viscosity: bad dtype");
    HIP_TRY(hipSetDevice(c->device));
    std::vector<unsigned char> tmp(n * c->cbytes, 0);
    if (c->mu_t) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipMemcpy(tmp.data(), c->mu_t, n * c->cbytes, hipMemcpyDeviceToHost));
    }
    if ((size_t)dtype == c->cbytes)
        std::memcpy(host, tmp.data(), n * c->cbytes);
    else if (c->cbytes == 4)
        convert((double*)host, (const float*)tmp.data(), n);
    else
        convert((float*)host, (const double*)tmp.data(), n);
    return OD_OK;
}

od_status od_closure_attach(od_ctx* c, od_mlp* m, const od_closure_opts* opts) {
    if (!c) return fail(OD_ERR_INVALID, "od_closure_attach: null context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!m) {
        c->mlp = nullptr;
        c->mu_t_on = false;
        return OD_OK;
    }
    if (!opts || opts->update_interval < 1) return fail(OD_ERR_INVALID, "od_closure_attach: update_interval >= 1");
    if (!c->p.viscous) return fail(OD_ERR_STATE, "od_closure_attach: Euler mode has no viscous fluxes");
    if (m->device != c->device) return fail(OD_ERR_INVALID, "od_closure_attach: network on another device");
    HIP_TRY(hipStreamSynchronize(m->stream));
    od_status s = ensure_closure_buffers(c);
    if (s == OD_OK) s = ensure_mu_t(c);
    if (s != OD_OK) return s;
    c->mlp = m;
    c->closure = *opts;
    c->steps_done = 0;
    c->ke_prev = -1.0;
    c->eps_prev = 0.0;
    c->closure_active = true;
    for (double& v : c->closure_stats) v = 0.0;
    c->mu_t_on = true;
    return OD_OK;
}

od_status od_closure_stats(od_ctx* c, double out[6]) {
    if (!c || !out) return fail(OD_ERR_INVALID, "od_closure_stats: null argument");
    for (int q = 0; q < 6; ++q) out[q] = c->closure_stats[q];
    return OD_OK;
}

}  // extern "C"
