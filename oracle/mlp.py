"""CPU oracle of the learned closure -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may use
this module; the product path never does.

No reference counterpart (drunkseal/opendynamo has no closure).  Restates,
in numpy float64:
  * the element features of the feature kernel (od_closure.hip), from the
    element's 64 nodal values: velocity gradient by the FR D matrix, Legendre
    modal energies, GL-weighted means;
  * the MLP 12 -> 64 -> 64 -> 64 -> 1 (tanh-GELU hidden layers, softplus
    output) with the BF16 roundings of the GPU path (inputs, weights and
    hidden activations rounded to BF16 before each product; sums exact here,
    FP32 on the GPU), its mean-squared-error loss and the gradient by
    reverse-mode differentiation.
Pinned by: finite differences of loss_grad (bf16=False), closed forms of the
features for linear and single-mode velocity fields, and hand-computed
GELU/softplus values (tests/test_oracle_closure.py).
"""
import math

import numpy as np

NF, NH = 12, 64
OFF = {}
_o = 0
for _name, _size in (("W1", NH * NF), ("b1", NH), ("W2", NH * NH), ("b2", NH), ("W3", NH * NH), ("b3", NH),
                     ("w4", NH), ("b4", 1)):
    OFF[_name] = (_o, _size)
    _o += _size
NPARAM = _o
assert NPARAM == 9217

_A, _B = 0.7978845608028654, 0.044715


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    g = lambda n: p[OFF[n][0]:OFF[n][0] + OFF[n][1]]
    return dict(W1=g("W1").reshape(NH, NF), b1=g("b1"), W2=g("W2").reshape(NH, NH), b2=g("b2"),
                W3=g("W3").reshape(NH, NH), b3=g("b3"), w4=g("w4"), b4=g("b4")[0])


def pack(d):
    out = np.zeros(NPARAM)
    for n, (o, s) in OFF.items():
        out[o:o + s] = np.ravel(d[n])
    return out


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(_A * (x + _B * x ** 3)))


def gelu_d(x):
    t = np.tanh(_A * (x + _B * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * _A * (1.0 + 3.0 * _B * x * x)


def softplus(z):
    return np.where(z > 20.0, z, np.log1p(np.exp(np.minimum(z, 20.0))))


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def _round(x, on):
    # the GPU rounds f32 values; a f64 value is first taken to f32
    return bf16(np.asarray(x, dtype=np.float32)) if on else np.asarray(x, dtype=np.float64)


def forward(params, x, shift=None, scale=None, rounding=True, out_scale=1.0):
    """C = out_scale * softplus(MLP(standardised x)); also returns the cache for loss_grad."""
    P = unpack(params)
    x = np.asarray(x, dtype=np.float64).reshape(-1, NF)
    shift = np.zeros(NF) if shift is None else np.asarray(shift, dtype=np.float64)
    scale = np.ones(NF) if scale is None else np.asarray(scale, dtype=np.float64)
    xs = (x.astype(np.float32) - shift.astype(np.float32)) * scale.astype(np.float32)  # f32 as on the GPU
    h0 = _round(xs, rounding)
    W1, W2, W3, w4 = (_round(P[n], rounding) for n in ("W1", "W2", "W3", "w4"))
    z1 = h0 @ W1.T + P["b1"]
    h1 = _round(gelu(z1), rounding)
    z2 = h1 @ W2.T + P["b2"]
    h2 = _round(gelu(z2), rounding)
    z3 = h2 @ W3.T + P["b3"]
    h3 = _round(gelu(z3), rounding)
    z4 = h3 @ w4 + P["b4"]
    y = float(np.float32(out_scale)) * softplus(z4)
    return y, dict(h0=h0, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, z4=z4, W2=W2, W3=W3, w4=w4)


def loss_grad(params, x, t, shift=None, scale=None, rounding=True, out_scale=1.0):
    """Mean squared error of C against t and its gradient (NPARAM)."""
    y, c = forward(params, x, shift, scale, rounding, out_scale)
    t = np.asarray(t, dtype=np.float64)
    n = y.size
    err = y - t
    loss = float(np.mean(err * err))
    dz4 = 2.0 * err / n * float(np.float32(out_scale)) * sigmoid(c["z4"])
    g = {"b4": np.sum(dz4), "w4": c["h3"].T @ dz4}
    dz3 = np.outer(dz4, c["w4"]) * gelu_d(c["z3"])
    g["W3"], g["b3"] = dz3.T @ c["h2"], dz3.sum(0)
    dz2 = (dz3 @ c["W3"]) * gelu_d(c["z2"])
    g["W2"], g["b2"] = dz2.T @ c["h1"], dz2.sum(0)
    dz1 = (dz2 @ c["W2"]) * gelu_d(c["z1"])
    g["W1"], g["b1"] = dz1.T @ c["h0"], dz1.sum(0)
    if rounding:  # the GPU multiplies BF16-rounded dZ in the weight gradients and back products
        dz3r, dz2r, dz1r = (bf16(np.asarray(d, dtype=np.float32)) for d in (dz3, dz2, dz1))
        g["W3"] = dz3r.T @ c["h2"]
        dz2 = ((dz3r @ c["W3"]) * gelu_d(c["z2"]))
        dz2r = bf16(np.asarray(dz2, dtype=np.float32))
        g["b2"] = dz2.sum(0)
        g["W2"] = dz2r.T @ c["h1"]
        dz1 = ((dz2r @ c["W2"]) * gelu_d(c["z1"]))
        dz1r = bf16(np.asarray(dz1, dtype=np.float32))
        g["b1"] = dz1.sum(0)
        g["W1"] = dz1r.T @ c["h0"]
    return loss, pack(g)


# ---------------------------------------------------------------------------
# Element features
# ---------------------------------------------------------------------------
def _legendre(a, x):
    p0, p1 = np.ones_like(x), x
    if a == 0:
        return p0
    for k in range(2, a + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    return p1


def fr_basics():
    """GL points and weights, D[i][m] = l_m'(x_i), L[a][i] = (2a+1)/2 w_i P_a(x_i)."""
    x, w = np.polynomial.legendre.leggauss(4)
    D = np.zeros((4, 4))
    for i in range(4):
        for m in range(4):
            if i == m:
                D[i, m] = sum(1.0 / (x[i] - x[q]) for q in range(4) if q != i)
            else:
                num = np.prod([x[i] - x[q] for q in range(4) if q not in (i, m)])
                den = np.prod([x[m] - x[q] for q in range(4) if q != m])
                D[i, m] = num / den
    L = np.array([[(2 * a + 1) / 2 * w[i] * _legendre(a, x[i]) for i in range(4)] for a in range(4)])
    return x, w, D, L


def features(state, nx, ny, nz, lx, ly, lz, mu_mol):
    """[E][16] features of every element (layout of od_closure_features)."""
    x, w, D, L = fr_basics()
    E = nx * ny * nz
    u = np.asarray(state, dtype=np.float64).reshape(5, E, 4, 4, 4)  # [v][e][k][j][i]
    rho = u[0]
    vel = u[1:4] / rho
    u2 = (vel ** 2).sum(0)
    ke = 0.5 * rho * u2
    p = 0.4 * (u[4] - ke)
    c = np.sqrt(np.maximum(1.4 * p / rho, 0.0))
    hx, hy, hz = lx / nx, ly / ny, lz / nz
    jac = (2.0 / hx, 2.0 / hy, 2.0 / hz)
    # A[a][d] at every point: d = 0 x (axis i), 1 y (j), 2 z (k)
    A = np.empty((3, 3) + vel.shape[1:])
    for a in range(3):
        A[a, 0] = np.einsum("im,ekjm->ekji", D, vel[a]) * jac[0]
        A[a, 1] = np.einsum("jm,ekmi->ekji", D, vel[a]) * jac[1]
        A[a, 2] = np.einsum("km,emji->ekji", D, vel[a]) * jac[2]
    div = A[0, 0] + A[1, 1] + A[2, 2]
    S = 0.5 * (A + A.transpose(1, 0, 2, 3, 4, 5))
    Sd = S - np.eye(3)[:, :, None, None, None, None] * div / 3.0
    Om = 0.5 * (A - A.transpose(1, 0, 2, 3, 4, 5))
    ss = (Sd ** 2).sum((0, 1))
    oo = (Om ** 2).sum((0, 1))
    sfull = (S ** 2).sum((0, 1))
    tr3 = np.einsum("ab...,bq...,qa...->...", A, A, A)
    smag, omag = np.sqrt(2 * ss), np.sqrt(2 * oo)
    Q = 0.5 * (oo - sfull)
    R = -tr3 / 3.0
    W = np.einsum("k,j,i->kji", w, w, w)
    mean = lambda f: np.einsum("ekji,kji->e", f, W) / 8.0
    # Legendre modal energy of the velocity, by highest degree
    t = np.einsum("ai,bj,ck,vekji->veabc", L, L, L, vel)  # [v][e][a(i)][b(j)][c(k)]
    nrm = 2.0 / (2 * np.arange(4) + 1)
    em = (t ** 2).sum(0) * np.einsum("a,b,c->abc", nrm, nrm, nrm)
    top = np.maximum.outer(np.maximum.outer(np.arange(4), np.arange(4)), np.arange(4))
    e1, e2, e3 = ((em * (top == q)).sum((1, 2, 3)) for q in (1, 2, 3))
    rb, pb, cb = mean(rho), mean(p), mean(c)
    dl = (hx * hy * hz) ** (1.0 / 3.0) / 4.0
    Sb = np.maximum(mean(smag), 1e-6 * cb / dl)
    etot = e1 + e2 + e3 + 1e-30
    mux, muy, muz = (mean(vel[q]) for q in range(3))
    var = lambda f_, m_: mean((f_ - m_[:, None, None, None]) ** 2)  # about the element mean
    f = np.zeros((E, 16))
    f[:, 0] = np.log(rb * dl * dl * Sb / mu_mol)
    f[:, 1] = mean(omag) / Sb
    f[:, 2] = mean(Q) / Sb ** 2
    f[:, 3] = mean(R) / Sb ** 3
    f[:, 4] = mean(div) / Sb
    f[:, 5], f[:, 6], f[:, 7] = e1 / etot, e2 / etot, e3 / etot
    f[:, 8] = np.sqrt(var(rho, rb)) / rb
    f[:, 9] = np.sqrt(var(p, pb)) / pb
    f[:, 10] = dl * Sb / cb
    f[:, 11] = np.sqrt(var(vel[0], mux) + var(vel[1], muy) + var(vel[2], muz)) / (dl * Sb)
    f[:, 12] = rb
    f[:, 13] = dl * dl * Sb
    vol8 = hx * hy * hz / 8.0
    f[:, 14] = np.einsum("ekji,kji->e", ke, W) * vol8
    f[:, 15] = np.einsum("ekji,kji->e", 2 * ss, W) * vol8
    return f


def mu_t_from_features(params, feat, shift, scale, lo, hi, rounding=True, out_scale=1.0):
    """Eddy viscosity per element as the inference kernel forms it."""
    y, _ = forward(params, feat[:, :NF], shift, scale, rounding, out_scale)
    inr = np.all((feat[:, :NF] >= lo) & (feat[:, :NF] <= hi), axis=1)
    return np.where(inr, feat[:, 12] * feat[:, 13] * y, 0.0), inr
