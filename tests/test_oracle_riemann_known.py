"""Known-answer cases for the oracle's SLAU2 and AUSM+ interface fluxes.

The expected values are derived by hand from the published methods, not from
the oracle's code, and evaluated with mpmath at 40 digits:

* SLAU2 -- mass flux of SLAU (Shima & Kitamura, AIAA J. 49(8), 2011, eqs. for
  m-dot, g, chi, M-hat) and the SLAU2 pressure flux (Kitamura & Shima, JCP 245,
  2013): p~ = (pL+pR)/2 + (P+(ML) - P-(MR)) (pL-pR)/2
                + sqrt((|VL|^2+|VR|^2)/2) (P+(ML) + P-(MR) - 1) rho_bar c_bar,
  P+- the fifth-order AUSM+ polynomials with alpha = 0, c_bar = (cL+cR)/2.
  This is the method the reference cites (SU2 CUpwSLAU_Flow with slau2 = true,
  src/cuda/device_functions.cuh:190-257).
* AUSM+ -- Liou, JCP 129, 1996: M+-(M) = +-(M+-1)^2/4 +- beta (M^2-1)^2,
  P+-(M) = (M+-1)^2 (2-+M)/4 +- alpha M (M^2-1)^2 for |M| < 1, alpha = 3/16,
  beta = 1/8, with the interface speed of sound c_1/2 = max(cL, cR) that the
  reference uses (device_functions.cuh:130-188).

Each case is chosen so that a plausible transcription mistake changes the
answer: the sign of the chi * dp / c_bar term, the g blend of |Vn|, the
full-speed (not normal-speed) square root in the pressure term, the beta/alpha
coefficients, the saturated branches, and max(cL, cR) against the mean.
`test_known_answers_reject_*_mutants` check that claim: a numpy restatement of
each method passes every case, and each mutated restatement fails at least
one of them.
"""
import mpmath as mp
import numpy as np
import pytest

from oracle import oracle as O

mp.mp.dps = 40
G = mp.mpf(14) / 10  # gamma = 1.4


def prim(rho, un, vt, wt, p):
    """Primitive state with the normal velocity un and tangential parts."""
    return dict(rho=mp.mpf(rho), un=mp.mpf(un), vt=mp.mpf(vt), wt=mp.mpf(wt), p=mp.mpf(p))


def total_energy(s):
    return s["p"] / (G - 1) + s["rho"] * (s["un"] ** 2 + s["vt"] ** 2 + s["wt"] ** 2) / 2


def enthalpy(s):
    return (total_energy(s) + s["p"]) / s["rho"]


def sound(s):
    return mp.sqrt(G * s["p"] / s["rho"])


def speed2(s):
    return s["un"] ** 2 + s["vt"] ** 2 + s["wt"] ** 2


def to_flux(mdot, p_face, up, d):
    """Flux in the normal frame [mass, normal mom, t1 mom, t2 mom, energy] ->
    lab-frame components for the face normal e_d (tangents t1, t2 are the
    other two axes in increasing order)."""
    normal = mdot * up["un"] + p_face
    t1, t2 = mdot * up["vt"], mdot * up["wt"]
    mom = [None] * 3
    others = [c for c in range(3) if c != d]
    mom[d], mom[others[0]], mom[others[1]] = normal, t1, t2
    return [mdot, *mom, mdot * enthalpy(up)]


def cons_lab(s, d):
    """Conservative state in the lab frame, normal along axis d."""
    vel = [None] * 3
    others = [c for c in range(3) if c != d]
    vel[d], vel[others[0]], vel[others[1]] = s["un"], s["vt"], s["wt"]
    return np.array([float(s["rho"]), *[float(s["rho"] * v) for v in vel], float(total_energy(s))])


# --------------------------------------------------------------------------
# Closed forms of the published methods
# --------------------------------------------------------------------------
def p_plus0(m):   # AUSM+ pressure polynomial with alpha = 0 (SLAU2's beta+)
    return (m + 1) ** 2 * (2 - m) / 4 if abs(m) < 1 else (mp.mpf(1) if m >= 0 else mp.mpf(0))


def p_minus0(m):
    return (m - 1) ** 2 * (2 + m) / 4 if abs(m) < 1 else (mp.mpf(0) if m >= 0 else mp.mpf(1))


def slau2_exact(L, R):
    cb = (sound(L) + sound(R)) / 2
    ml, mr = L["un"] / cb, R["un"] / cb
    mhat = min(mp.mpf(1), mp.sqrt((speed2(L) + speed2(R)) / 2) / cb)
    chi = (1 - mhat) ** 2
    g = -max(min(ml, 0), -1) * min(max(mr, 0), 1)
    vbar = (L["rho"] * abs(L["un"]) + R["rho"] * abs(R["un"])) / (L["rho"] + R["rho"])
    vp = (1 - g) * vbar + g * abs(L["un"])
    vm = (1 - g) * vbar + g * abs(R["un"])
    mdot = (L["rho"] * (L["un"] + vp) + R["rho"] * (R["un"] - vm) - chi / cb * (R["p"] - L["p"])) / 2
    bl, br = p_plus0(ml), p_minus0(mr)
    pt = (L["p"] + R["p"]) / 2 + (bl - br) * (L["p"] - R["p"]) / 2 \
        + mp.sqrt((speed2(L) + speed2(R)) / 2) * (bl + br - 1) * (L["rho"] + R["rho"]) / 2 * cb
    return mdot, pt, (L if mdot > 0 else R)


def ausm_exact(L, R):
    c = max(sound(L), sound(R))
    ml, mr = L["un"] / c, R["un"] / c
    al, be = mp.mpf(3) / 16, mp.mpf(1) / 8
    if abs(ml) < 1:
        mp_ = (ml + 1) ** 2 / 4 + be * (ml * ml - 1) ** 2
        pp = (ml + 1) ** 2 * (2 - ml) / 4 + al * ml * (ml * ml - 1) ** 2
    else:
        mp_, pp = (ml + abs(ml)) / 2, mp.mpf(1) if ml > 0 else mp.mpf(0)
    if abs(mr) < 1:
        mm = -(mr - 1) ** 2 / 4 - be * (mr * mr - 1) ** 2
        pm = (mr - 1) ** 2 * (2 + mr) / 4 - al * mr * (mr * mr - 1) ** 2
    else:
        mm, pm = (mr - abs(mr)) / 2, mp.mpf(1) if mr < 0 else mp.mpf(0)
    m = mp_ + mm
    up = L if m > 0 else R
    return m * c * up["rho"], pp * L["p"] + pm * R["p"], up


GI = 1 / G
SLAU2_CASES = {
    # both at rest, equal density, pL != pR: M-hat = 0 so chi = 1, g = 0,
    # beta+- = 1/2: m-dot = -(pR - pL) / (2 c_bar), p~ = (pL + pR) / 2
    "rest_pressure_jump": (prim(1, 0, 0, 0, 1), prim(1, 0, 0, 0, "0.5")),
    # subsonic, same direction (g = 0), cL = cR = 1, tangential velocities
    # in the sqrt term: beta+(0.5) = 27/32, beta-(0.25) = 81/256
    "subsonic_compression": (prim(1, "0.5", "0.2", 0, GI), prim(2, "0.25", "-0.1", "0.05", 2 * GI)),
    # expansion ML < 0 < MR: g = 1/8 blends |Vn| towards each side's own
    "subsonic_expansion": (prim(1, "-0.5", "0.2", 0, GI), prim(2, "0.25", "-0.1", 0, 2 * GI)),
    # both supersonic: M-hat clips to 1 (chi = 0), beta+ = 1, beta- = 0 -> p~ = pL
    "supersonic": (prim(1, "1.5", "0.1", 0, GI), prim("0.8", "1.2", 0, "0.1", mp.mpf("0.7") * GI)),
    # unequal sound speeds, opposing flow with a density jump
    "unequal_sound_speed": (prim("1.2", "0.3", 0, "-0.2", "1.1"), prim("0.7", "-0.1", "0.15", 0, "0.6")),
}
AUSM_CASES = {
    # c_1/2 = max(1, 1.1) = 1.1: ML = 0.5, MR = 0.25 (M+(0.5) = 0.6328125,
    # P+(0.5) = 0.896484375)
    "subsonic_max_c": (prim(1, "0.55", "0.1", 0, GI), prim(1, "0.275", 0, "-0.1", mp.mpf("1.21") * GI)),
    # ML = -0.5 subsonic, MR = -1.5 supersonic to the left: right upwinding,
    # P-(-1.5) = 1
    "mixed_left_running": (prim(1, "-0.5", 0, "0.2", GI), prim("1.5", "-1.5", "0.1", 0, mp.mpf("1.5") * GI)),
    # opposing subsonic, m_1/2 < 0
    "opposing": (prim("0.9", "0.2", "0.3", 0, "0.8"), prim("1.3", "-0.4", 0, "0.1", "1.1")),
}


def expected(kind, L, R, d):
    mdot, pf, up = (slau2_exact if kind == "slau2" else ausm_exact)(L, R)
    return np.array([float(x) for x in to_flux(mdot, pf, up, d)])


def oracle_flux(kind, L, R, d, real=8):
    n = np.zeros(3)
    n[d] = 1.0
    return O.point(real, 7 if kind == "slau2" else 8, np.concatenate([cons_lab(L, d), cons_lab(R, d), n]))


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("name", sorted(SLAU2_CASES))
def test_slau2_known_answers(name, d):
    L, R = SLAU2_CASES[name]
    want = expected("slau2", L, R, d)
    got = oracle_flux("slau2", L, R, d)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-14 * np.abs(want).max())


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("name", sorted(AUSM_CASES))
def test_ausm_plus_known_answers(name, d):
    L, R = AUSM_CASES[name]
    want = expected("ausm", L, R, d)
    got = oracle_flux("ausm", L, R, d)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-14 * np.abs(want).max())


def test_slau2_rest_case_in_closed_form():
    """The at-rest case written out without the general formula."""
    L, R = SLAU2_CASES["rest_pressure_jump"]
    cb = (mp.sqrt(G) + mp.sqrt(G / 2)) / 2
    mdot = -(R["p"] - L["p"]) / (2 * cb)
    want = np.array([float(mdot), 0.75, 0.0, 0.0, float(mdot * G / (G - 1))])  # H_L = gamma pL / ((gamma-1) rhoL)
    np.testing.assert_allclose(oracle_flux("slau2", L, R, 0), want, rtol=1e-14, atol=1e-16)


def test_polynomial_values():
    """Hand values of the split polynomials used above."""
    assert p_plus0(mp.mpf("0.5")) == mp.mpf(27) / 32 == mp.mpf("0.84375")
    assert p_minus0(mp.mpf("0.25")) == mp.mpf(81) / 256
    L, R = AUSM_CASES["subsonic_max_c"]
    c = max(sound(L), sound(R))
    assert abs(c - mp.mpf("1.1")) < mp.mpf(10) ** -30
    ml = L["un"] / c
    assert abs((ml + 1) ** 2 / 4 + (ml * ml - 1) ** 2 / 8 - mp.mpf("0.6328125")) < mp.mpf(10) ** -30
    assert abs((ml + 1) ** 2 * (2 - ml) / 4 + 3 * ml * (ml * ml - 1) ** 2 / 16 - mp.mpf("0.896484375")) < mp.mpf(10) ** -30


@pytest.mark.parametrize("kind,cases", [("slau2", SLAU2_CASES), ("ausm", AUSM_CASES)])
def test_f32_oracle_close_to_known_answers(kind, cases):
    for L, R in cases.values():
        want = expected(kind, L, R, 1)
        got = oracle_flux(kind, L, R, 1, real=4)
        np.testing.assert_allclose(got, want, rtol=3e-6, atol=3e-6 * np.abs(want).max())


# --------------------------------------------------------------------------
# Sensitivity of the cases: mutated restatements must fail
# --------------------------------------------------------------------------
def _np_slau2(L, R, d, mut=None):
    """Straight numpy SLAU2 on lab-frame conservative states; `mut` plants one
    plausible transcription mistake."""
    def pv(s):
        rho = s[0]
        vel = s[1:4] / rho
        p = 0.4 * (s[4] - 0.5 * rho * vel @ vel)
        return rho, vel, p, np.sqrt(1.4 * p / rho), (s[4] + p) / rho
    rl, vl, pl, cl, hl = pv(L)
    rr, vr, pr, cr, hr = pv(R)
    cb = 0.5 * (cl + cr)
    unl, unr = vl[d], vr[d]
    ml, mr = unl / cb, unr / cb
    q2 = 0.5 * (vl @ vl + vr @ vr) if mut != "normal_speed_sqrt" else 0.5 * (unl * unl + unr * unr)
    mhat = min(1.0, np.sqrt(q2) / cb)
    chi = (1 - mhat) ** 2
    g = -max(min(ml, 0.0), -1.0) * min(max(mr, 0.0), 1.0)
    if mut == "no_g_blend":
        g = 0.0
    vbar = (rl * abs(unl) + rr * abs(unr)) / (rl + rr)
    vp, vm = (1 - g) * vbar + g * abs(unl), (1 - g) * vbar + g * abs(unr)
    dp = (pr - pl) if mut != "chi_sign" else (pl - pr)
    mdot = 0.5 * (rl * (unl + vp) + rr * (unr - vm) - chi / cb * dp)
    def bp(m): return 0.25 * (2 - m) * (m + 1) ** 2 if abs(m) < 1 else (1.0 if m >= 0 else 0.0)
    def bm(m): return 0.25 * (2 + m) * (m - 1) ** 2 if abs(m) < 1 else (0.0 if m >= 0 else 1.0)
    bl, br = bp(ml), bm(mr)
    if mut == "beta_ausm_alpha":  # alpha = 3/16 left in the pressure polynomials
        if abs(ml) < 1:
            bl += 0.1875 * ml * (ml * ml - 1) ** 2
        if abs(mr) < 1:
            br -= 0.1875 * mr * (mr * mr - 1) ** 2
    pt = 0.5 * (pl + pr) + 0.5 * (bl - br) * (pl - pr)
    if mut != "no_sqrt_term":
        pt += np.sqrt(q2) * (bl + br - 1) * cb * 0.5 * (rl + rr)
    up_v, up_h = (vl, hl) if mdot > 0 else (vr, hr)
    f = np.array([mdot, *(mdot * up_v), mdot * up_h])
    f[1 + d] += pt
    return f


def _np_ausm(L, R, d, mut=None):
    def pv(s):
        rho = s[0]
        vel = s[1:4] / rho
        p = 0.4 * (s[4] - 0.5 * rho * vel @ vel)
        return rho, vel, p, np.sqrt(1.4 * p / rho), (s[4] + p) / rho
    rl, vl, pl, cl, hl = pv(L)
    rr, vr, pr, cr, hr = pv(R)
    c = max(cl, cr) if mut != "mean_c" else 0.5 * (cl + cr)
    al, be = 0.1875, (0.125 if mut != "beta_1_16" else 0.0625)
    if mut == "alpha_0":
        al = 0.0
    ml, mr = vl[d] / c, vr[d] / c
    if abs(ml) < 1:
        mpl = 0.25 * (ml + 1) ** 2 + be * (ml * ml - 1) ** 2
        ppl = 0.25 * (ml + 1) ** 2 * (2 - ml) + al * ml * (ml * ml - 1) ** 2
    else:
        mpl, ppl = 0.5 * (ml + abs(ml)), (1.0 if ml > 0 else 0.0)
    if abs(mr) < 1:
        mmr = -0.25 * (mr - 1) ** 2 - be * (mr * mr - 1) ** 2
        pmr = 0.25 * (mr - 1) ** 2 * (2 + mr) - al * mr * (mr * mr - 1) ** 2
    else:
        mmr, pmr = 0.5 * (mr - abs(mr)), (1.0 if mr < 0 else 0.0)
    m = mpl + mmr
    rho_up, up_v, up_h = (rl, vl, hl) if m > 0 else (rr, vr, hr)
    mdot = m * c * rho_up
    f = np.array([mdot, *(mdot * up_v), mdot * up_h])
    f[1 + d] += ppl * pl + pmr * pr
    return f


def _passes(fn, kind, cases, mut):
    for L, R in cases.values():
        for d in range(3):
            want = expected(kind, L, R, d)
            got = fn(cons_lab(L, d), cons_lab(R, d), d, mut)
            if not np.allclose(got, want, rtol=1e-12, atol=1e-13 * np.abs(want).max()):
                return False
    return True


@pytest.mark.parametrize("mut", ["chi_sign", "no_g_blend", "normal_speed_sqrt", "no_sqrt_term", "beta_ausm_alpha"])
def test_known_answers_reject_slau2_mutants(mut):
    assert _passes(_np_slau2, "slau2", SLAU2_CASES, None), "the faithful restatement must pass"
    assert not _passes(_np_slau2, "slau2", SLAU2_CASES, mut), f"mutant {mut} was not detected"


@pytest.mark.parametrize("mut", ["beta_1_16", "alpha_0", "mean_c"])
def test_known_answers_reject_ausm_mutants(mut):
    assert _passes(_np_ausm, "ausm", AUSM_CASES, None), "the faithful restatement must pass"
    assert not _passes(_np_ausm, "ausm", AUSM_CASES, mut), f"mutant {mut} was not detected"
