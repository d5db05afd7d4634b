"""Textbook FR P=3 right-hand side assembled with dense tensor algebra.

An independent restatement used to pin the oracle's assembly (derivative
placement, face pairing, trace sides, correction signs, BR1 gradient, slip
wall): flux reconstruction with Radau g-functions on a Cartesian hex mesh,
  R = -sum_d J_d [ D_d F_d + g_L'(xi_d) (F*_L - F_L) + g_R'(xi_d) (F*_R - F_R) ]
with F = F_e - F_v, the BR1 common solution 1/2 (u- + u+) for the gradient
and F* = F_riemann - 1/2 (F_v- + F_v+) at interfaces.  Point fluxes come from
the oracle's point physics, which the physics tests pin on their own; only
the assembly is restated here (different summation order: agreement is to
round-off, not bitwise).
"""
import numpy as np

from oracle import oracle as O


def _apply(kind, *arrays):
    """Evaluate oracle point physics `kind` over stacked inputs (n_in, M)."""
    inp = np.concatenate([np.atleast_2d(a) for a in arrays], axis=0)
    out = np.empty((5, inp.shape[1]))
    for q in range(inp.shape[1]):
        out[:, q] = O.point(8, kind, inp[:, q])
    return out


class Mesh:
    def __init__(self, nx, ny, nz, lx, ly, lz):
        self.n = (nx, ny, nz)
        L = [float(np.float32(x)) for x in (lx, ly, lz)]
        self.jac = [2.0 / (L[d] / self.n[d]) for d in range(3)]
        ops = O.fr_operators()
        self.D = ops["d"].reshape(4, 4)
        self.eL, self.eR = ops["extrap_left"], ops["extrap_right"]
        self.gL, self.gR = ops["g_left_prime"], ops["g_right_prime"]

    # arrays: (comp, nz, ny, nx, k, j, i); direction d: element axis 3-d, point axis 6-d
    @staticmethod
    def eax(d):
        return 3 - d

    @staticmethod
    def pax(d):
        return 6 - d

    def deriv(self, A, d):
        B = np.moveaxis(A, self.pax(d), -1) @ self.D.T
        return np.moveaxis(B, -1, self.pax(d))

    def trace(self, A, d, side):
        c = self.eL if side == 0 else self.eR
        return np.moveaxis(A, self.pax(d), -1) @ c

    def spread(self, F, d, gp):
        """Face array (no point axis d) times g'(xi_d) placed along axis d."""
        shape = [1] * 7
        shape[self.pax(d)] = 4
        return np.expand_dims(F, self.pax(d)) * gp.reshape(shape)


def reflect(T):
    R = T.copy()
    R[1:4] *= -1.0
    return R


def textbook_rhs(state, nx, ny, nz, lx, ly, lz, scheme=2, wall=False, viscous=False, mu=1e-3, bf=0.001):
    m = Mesh(nx, ny, nz, lx, ly, lz)
    U = np.asarray(state, dtype=np.float64).reshape(5, nz, ny, nx, 4, 4, 4)
    flat = lambda A: A.reshape(A.shape[0], -1)
    unflat = lambda A, like: A.reshape((A.shape[0],) + like.shape[1:])

    def neighbour_traces(T_left_face, T_right_face, d):
        """Given each element's own L/R traces, return the neighbour's facing
        trace at the L and R faces (None entries = slip wall)."""
        ax = m.eax(d)
        nbL = np.roll(T_right_face, 1, axis=ax)
        nbR = np.roll(T_left_face, -1, axis=ax)
        return nbL, nbR

    # --- BR1 gradient ---------------------------------------------------
    Q = None
    if viscous:
        Q = np.zeros((15,) + U.shape[1:])
        for d in range(3):
            G = m.deriv(U, d)
            TL, TR = m.trace(U, d, 0), m.trace(U, d, 1)
            nbL, nbR = neighbour_traces(TL, TR, d)
            if wall and d == 1:
                nbL[:, :, 0] = reflect(TL)[:, :, 0]
                nbR[:, :, -1] = reflect(TR)[:, :, -1]
            G = G + m.spread(0.5 * (nbL + TL) - TL, d, m.gL) + m.spread(0.5 * (nbR + TR) - TR, d, m.gR)
            Q[d::3] = m.jac[d] * G

    # --- volume fluxes --------------------------------------------------
    R = np.zeros_like(U)
    for d in range(3):
        Fe = unflat(_apply(1 + d, flat(U)), U)
        F = Fe
        if viscous:
            mus = np.full((1, flat(U).shape[1]), mu)
            F = Fe - unflat(_apply(4 + d, flat(U), flat(Q), mus), U)
        R -= m.jac[d] * m.deriv(F, d)

    # --- interfaces ------------------------------------------------------
    for d in range(3):
        n = np.zeros((3, 1))
        n[d] = 1.0
        TL, TR = m.trace(U, d, 0), m.trace(U, d, 1)     # own left / right face traces
        ax = m.eax(d)
        # interface between element e (left) and e+1 (right): UL = TR[e], UR = TL[e+1]
        UL, UR = TR, np.roll(TL, -1, axis=ax)
        M = flat(UL).shape[1]
        Fr = unflat(_apply(7 + scheme, flat(UL), flat(UR), np.repeat(n, M, axis=1)), UL)
        FeL, FeR = unflat(_apply(1 + d, flat(UL)), UL), unflat(_apply(1 + d, flat(UR)), UR)
        if viscous:
            QL, QR = m.trace(Q, d, 1), np.roll(m.trace(Q, d, 0), -1, axis=ax)
            mus = np.full((1, M), mu)
            FvL = unflat(_apply(4 + d, flat(UL), flat(QL), mus), UL)
            FvR = unflat(_apply(4 + d, flat(UR), flat(QR), mus), UR)
        else:
            FvL = FvR = np.zeros_like(UL)
        Fs = Fr - 0.5 * (FvL + FvR)
        jR = Fs - (FeL - FvL)                      # for the left element's right face
        jL = np.roll(Fs - (FeR - FvR), 1, axis=ax)  # for the right element's left face
        if wall and d == 1:
            # slip walls: F_wall = (0, 0, p, 0, 0) at y = -ly/2 (left face of ey = 0)
            # and y = +ly/2 (right face of ey = ny-1)
            for side, T, sl in ((0, TL, (slice(None), slice(None), 0)), (1, TR, (slice(None), slice(None), -1))):
                Tw = T[sl]
                Fe_w = unflat(_apply(2, flat(Tw)), Tw)
                if viscous:
                    Qw = m.trace(Q, d, side)[sl]
                    Fv_w = unflat(_apply(5, flat(Tw), flat(Qw), np.full((1, flat(Tw).shape[1]), mu)), Tw)
                else:
                    Fv_w = np.zeros_like(Tw)
                W = np.zeros_like(Tw)
                W[2] = Fe_w[2] - Tw[2] * Tw[2] / Tw[0]
                jw = (W - Fv_w) - (Fe_w - Fv_w)
                if side == 0:
                    jL[sl] = jw
                else:
                    jR[sl] = jw
        R -= m.jac[d] * (m.spread(jL, d, m.gL) + m.spread(jR, d, m.gR))
    R[1] += bf
    return R.reshape(5, -1).ravel()
