"""Pins of the oracle's Riemann solvers and deboost (riemann.pyx:190-1090).

Checks, none of which restates the oracle's own formulas:
  - the reference's HLL known answer (test_riemann.py:88-122);
  - consistency: F(U, U) is the physical Euler flux, for HLL, HLLC and the
    exact solver, in the face frame and in the lab frame with a moving face
    (Galilean invariance of the deboost, Pakmor & Springel 2011 eq. 17);
  - mirror symmetry F(L, R, n) = -F(R, L, -n);
  - supersonic upwinding and an exactly resolved stationary contact (HLLC);
  - wave-speed properties of get_waves (identical states give u -/+ c);
  - the exact solver against Toro's Table 4.2 star states, sampled at
    x/t = 0 with the closed-form star densities;
  - HLLC star states and star fluxes, read off the oracle through two face
    speeds in each star region, against the Rankine-Hugoniot conditions
    across all three waves and the contact conditions (Toro eq. 10.35);
  - the star-pressure estimate of get_waves: the two-rarefaction estimate
    equals the exact star pressure where both waves are rarefactions (Toro
    eq. 4.46 and the root of the pressure function), and the choice of
    estimate follows Toro sec. 9.5 with Q = 2 on data either side of
    p_max/p_min = 2.
tests/test_oracle_mutations.py shows that planted transcription mistakes
fail these pins.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_known_answers.json")))


def euler_flux(dim, gamma, W, n, w=None):
    """Physical flux of primitive state W through a face with normal n moving
    with velocity w (lab frame): F - (w.n) U."""
    d, v, p = W[0], np.asarray(W[1:dim + 1]), W[dim + 1]
    vn = float(v @ n)
    wn = 0.0 if w is None else float(np.asarray(w) @ n)
    E = 0.5 * d * float(v @ v) + p / (gamma - 1.0)
    return np.array([d * (vn - wn)] + list(d * v * (vn - wn) + p * n) + [E * (vn - wn) + p * vn])


def solve(solver, dim, gamma, boost, L, R, n, fv):
    pl = [np.array([x]) for x in L]
    pr = [np.array([x]) for x in R]
    nn = [np.array([x]) for x in n]
    ff = [np.array([x]) for x in fv]
    F, rc, _ = oracle.riemann(solver, dim, gamma, boost, pl, pr, nn, ff)
    assert rc == 0
    return np.array([f[0] for f in F])


def test_hll_reference_known_answer():
    g = GOLD["hll_face_known_answer"]
    for case in g["cases"]:
        st = case["state"]
        W = [st["density"]] + st["velocity"] + [st["pressure"]]
        F = solve("hll", 2, g["gamma"], False, W, W, g["normal"], [0.0, 0.0])
        want = [case["flux"][k] for k in ("mass", "momentum-x", "momentum-y", "energy")]
        np.testing.assert_allclose(F, want, atol=10.0 ** -g["places"])


def _random_state(rng, dim):
    return [10 ** rng.uniform(-1, 1)] + list(rng.uniform(-1, 1, dim)) + [10 ** rng.uniform(-1, 1)]


@pytest.mark.parametrize("solver", ["hll", "hllc", "exact"])
@pytest.mark.parametrize("dim", [2, 3])
def test_consistency_moving_face(solver, dim):
    rng = np.random.default_rng(11)
    gamma = 1.4
    for _ in range(50):
        W = _random_state(rng, dim)
        n = rng.normal(size=dim)
        n /= np.linalg.norm(n)
        w = rng.uniform(-0.5, 0.5, dim)
        want = euler_flux(dim, gamma, W, n, w)
        # boosted face frame + deboost (the integrator's path): the states
        # arrive with face-frame velocities v - w (reconstruction.pyx:846-847)
        Wf = [W[0]] + [W[1 + k] - w[k] for k in range(dim)] + [W[dim + 1]]
        F = solve(solver, dim, gamma, True, Wf, Wf, n, w)
        np.testing.assert_allclose(F, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
        if solver != "exact":
            # lab frame with the face velocity projected (riemann.pyx:327-331)
            F2 = solve(solver, dim, gamma, False, W, W, n, w)
            np.testing.assert_allclose(F2, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("solver", ["hll", "hllc", "exact"])
def test_mirror_symmetry(solver):
    rng = np.random.default_rng(5)
    dim, gamma = 3, 1.4
    for _ in range(200):
        L, R = _random_state(rng, dim), _random_state(rng, dim)
        n = rng.normal(size=dim)
        n /= np.linalg.norm(n)
        F1 = solve(solver, dim, gamma, True, L, R, n, np.zeros(dim))
        F2 = solve(solver, dim, gamma, True, R, L, -n, np.zeros(dim))
        tol = 1e-6 if solver == "exact" else 1e-11
        np.testing.assert_allclose(F1, -F2, rtol=tol, atol=tol * np.abs(F1).max())


def test_hllc_supersonic_upwind_and_contact():
    dim, gamma = 2, 1.4
    n = np.array([1.0, 0.0])
    # both states supersonic to the right: the flux is F(L) exactly
    L = [1.0, 5.0, 0.3, 1.0]
    R = [0.5, 4.0, -0.2, 0.7]
    F = solve("hllc", dim, gamma, True, L, R, n, np.zeros(2))
    np.testing.assert_allclose(F, euler_flux(dim, gamma, L, n), rtol=1e-14)
    # stationary contact: equal pressure and normal velocity 0, density jump
    L = [1.0, 0.0, 0.4, 2.0]
    R = [0.125, 0.0, -0.3, 2.0]
    F = solve("hllc", dim, gamma, True, L, R, n, np.zeros(2))
    assert abs(F[0]) < 1e-15 and abs(F[3]) < 1e-14
    np.testing.assert_allclose(F[1], 2.0, rtol=1e-14)
    assert abs(F[2]) < 1e-15


def test_get_waves_identical_states():
    rng = np.random.default_rng(2)
    gamma = 1.4
    for _ in range(100):
        d, u, p = 10 ** rng.uniform(-1, 1), rng.uniform(-2, 2), 10 ** rng.uniform(-1, 1)
        sl, sc, sr = oracle.get_waves(d, u, p, d, u, p, gamma)
        c = np.sqrt(gamma * p / d)
        np.testing.assert_allclose([sl, sr], [u - c, u + c], rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(sc, u, rtol=1e-12, atol=1e-12)


def test_get_waves_outrun_the_sound_waves():
    rng = np.random.default_rng(3)
    for _ in range(500):
        dl, dr = 10 ** rng.uniform(-1, 1, 2)
        ul, ur = rng.uniform(-1, 1, 2)
        pl, pr = 10 ** rng.uniform(-1, 1, 2)
        sl, sc, sr = oracle.get_waves(dl, ul, pl, dr, ur, pr, 1.4)
        assert sl < sr
        # Toro eq. 10.68-10.69 (q >= 1): never slower than the data's sound waves
        assert sl <= ul - np.sqrt(1.4 * pl / dl) + 1e-12
        assert sr >= ur + np.sqrt(1.4 * pr / dr) - 1e-12


def _sample_exact(gamma, L, R, p_star, u_star):
    """State at x/t = 0 from the published star pressure and velocity
    (Toro eq. 4.53, 4.57, 4.59-4.64)."""
    dl, ul, pl = L
    dr, ur, pr = R
    g1 = (gamma - 1.0) / (gamma + 1.0)
    cl, cr = np.sqrt(gamma * pl / dl), np.sqrt(gamma * pr / dr)
    if u_star >= 0.0:
        if p_star <= pl:
            if ul - cl >= 0.0:
                return L
            cs = cl * (p_star / pl) ** ((gamma - 1) / (2 * gamma))
            if u_star - cs <= 0.0:
                return [dl * (p_star / pl) ** (1.0 / gamma), u_star, p_star]
            c = 2.0 / (gamma + 1) * (cl + 0.5 * (gamma - 1) * ul)
            return [dl * (c / cl) ** (2 / (gamma - 1)), c, pl * (c / cl) ** (2 * gamma / (gamma - 1))]
        sl = ul - cl * np.sqrt((gamma + 1) / (2 * gamma) * p_star / pl + (gamma - 1) / (2 * gamma))
        if sl >= 0.0:
            return L
        return [dl * (p_star / pl + g1) / (g1 * p_star / pl + 1.0), u_star, p_star]
    if p_star >= pr:
        sr = ur + cr * np.sqrt((gamma + 1) / (2 * gamma) * p_star / pr + (gamma - 1) / (2 * gamma))
        if sr <= 0.0:
            return R
        return [dr * (p_star / pr + g1) / (g1 * p_star / pr + 1.0), u_star, p_star]
    if ur + cr <= 0.0:
        return R
    cs = cr * (p_star / pr) ** ((gamma - 1) / (2 * gamma))
    if u_star + cs >= 0.0:
        return [dr * (p_star / pr) ** (1.0 / gamma), u_star, p_star]
    c = 2.0 / (gamma + 1) * (cr - 0.5 * (gamma - 1) * ur)
    return [dr * (c / cr) ** (2 / (gamma - 1)), -c, pr * (c / cr) ** (2 * gamma / (gamma - 1))]


def test_exact_solver_toro_table():
    g = GOLD["toro_exact_star_states"]
    gamma = g["gamma"]
    n = np.array([1.0, 0.0, 0.0])
    for t in g["tests"]:
        L, R = t["left"], t["right"]
        W0 = _sample_exact(gamma, L, R, t["p_star"], t["u_star"])
        want = euler_flux(3, gamma, [W0[0], W0[1], 0.0, 0.0, W0[2]], n)
        got = solve("exact", 3, gamma, True, [L[0], L[1], 0.0, 0.0, L[2]],
                    [R[0], R[1], 0.0, 0.0, R[2]], n, np.zeros(3))
        scale = np.abs(want).max()
        tol = t.get("rel_tol", g["rel_tol"])
        np.testing.assert_allclose(got, want, rtol=tol, atol=tol * scale, err_msg=t["name"])


# ---------------------------------------------------------------------------
# HLLC star states (riemann.pyx:636-672).  In the lab frame (no boost) a face
# moving with normal speed wn inside the left star region carries
# F*L - wn U*L (Toro 2009 eq. 10.71 with the face's own speed), an affine
# function of wn.  Two face speeds inside one region therefore give the star
# state U*K (the slope) and the star flux F*K (the intercept) of the oracle,
# without restating its formulas.  They must satisfy the Rankine-Hugoniot
# conditions across the three waves (Toro eq. 10.35):
#     F*L = FL + SL (U*L - UL),  F*R = FR + SR (U*R - UR),
#     F*R = F*L + S* (U*R - U*L),
# and the contact conditions: u* = S* on both sides, tangential velocity
# carried unchanged.  The third condition holds only for Toro's star energy
# (eq. 10.39); a wrong sign in its p/(rho (S - u)) term breaks it.

def _conserved(dim, gamma, W):
    d, v, p = W[0], np.asarray(W[1:dim + 1]), W[dim + 1]
    return np.array([d] + list(d * v) + [0.5 * d * float(v @ v) + p / (gamma - 1.0)])


def _lab_flux(dim, gamma, L, R, n, wn):
    return solve("hllc", dim, gamma, False, L, R, n, wn * np.asarray(n))


def _star(dim, gamma, L, R, n, w1, w2):
    F1 = _lab_flux(dim, gamma, L, R, n, w1)
    F2 = _lab_flux(dim, gamma, L, R, n, w2)
    U = (F2 - F1) / (w1 - w2)
    return F1 + w1 * U, U


@pytest.mark.parametrize("dim", [2, 3])
def test_hllc_star_states_satisfy_rankine_hugoniot(dim):
    rng = np.random.default_rng(21 + dim)
    gamma = 1.4
    checked = 0
    for _ in range(400):
        L, R = _random_state(rng, dim), _random_state(rng, dim)
        n = rng.normal(size=dim)
        n /= np.linalg.norm(n)
        uL = float(np.asarray(L[1:dim + 1]) @ n)
        uR = float(np.asarray(R[1:dim + 1]) @ n)
        sl, sc, sr = oracle.get_waves(L[0], uL, L[dim + 1], R[0], uR, R[dim + 1], gamma)
        if min(sc - sl, sr - sc) < 1e-3 * (sr - sl):
            continue
        FsL, UsL = _star(dim, gamma, L, R, n, sl + 0.2 * (sc - sl), sl + 0.8 * (sc - sl))
        FsR, UsR = _star(dim, gamma, L, R, n, sc + 0.2 * (sr - sc), sc + 0.8 * (sr - sc))
        UL, UR = _conserved(dim, gamma, L), _conserved(dim, gamma, R)
        FL, FR = euler_flux(dim, gamma, L, n), euler_flux(dim, gamma, R, n)
        scale = max(np.abs(FsL).max(), np.abs(FsR).max(), np.abs(FL).max(), np.abs(FR).max(),
                    abs(sl) * np.abs(UL).max(), abs(sr) * np.abs(UR).max())
        tol = 1e-10 * scale
        np.testing.assert_allclose(FsL, FL + sl * (UsL - UL), rtol=0, atol=tol)
        np.testing.assert_allclose(FsR, FR + sr * (UsR - UR), rtol=0, atol=tol)
        np.testing.assert_allclose(FsR, FsL + sc * (UsR - UsL), rtol=0, atol=tol)
        for Us, W in ((UsL, L), (UsR, R)):
            mom = Us[1:dim + 1]
            np.testing.assert_allclose(float(mom @ n) / Us[0], sc, rtol=1e-9, atol=1e-9)
            v = np.asarray(W[1:dim + 1])
            np.testing.assert_allclose(mom / Us[0] - (float(mom @ n) / Us[0]) * n,
                                       v - float(v @ n) * n, rtol=0, atol=1e-9)
        checked += 1
    assert checked >= 300


def test_hllc_face_frame_equals_lab_frame_star_flux():
    """The boosted solve (face frame, then deboost) of a moving face equals
    the lab-frame flux through that face (Pakmor & Springel 2011 eq. 17),
    here for distinct states in both star regions."""
    rng = np.random.default_rng(31)
    dim, gamma = 3, 1.4
    for _ in range(200):
        L, R = _random_state(rng, dim), _random_state(rng, dim)
        n = rng.normal(size=dim)
        n /= np.linalg.norm(n)
        w = rng.uniform(-0.5, 0.5, dim)
        lab = solve("hllc", dim, gamma, False, L, R, n, w)
        Lf = [L[0]] + [L[1 + k] - w[k] for k in range(dim)] + [L[dim + 1]]
        Rf = [R[0]] + [R[1 + k] - w[k] for k in range(dim)] + [R[dim + 1]]
        boosted = solve("hllc", dim, gamma, True, Lf, Rf, n, w)
        np.testing.assert_allclose(boosted, lab, rtol=1e-11, atol=1e-11 * np.abs(lab).max())


# ---------------------------------------------------------------------------
# get_waves star-pressure estimates (riemann.pyx:375-504, Toro 2009 ch. 9)

def _exact_pstar(gamma, dl, ul, pl, dr, ur, pr):
    """Star pressure of the exact Riemann problem: root of Toro's pressure
    function f_L(p) + f_R(p) + (u_R - u_L) (eq. 4.5-4.7) by bisection."""
    from scipy.optimize import brentq

    def fk(p, d, pk):
        ck = np.sqrt(gamma * pk / d)
        if p <= pk:
            return 2.0 * ck / (gamma - 1.0) * ((p / pk) ** ((gamma - 1.0) / (2.0 * gamma)) - 1.0)
        a, b = 2.0 / ((gamma + 1.0) * d), (gamma - 1.0) / (gamma + 1.0) * pk
        return (p - pk) * np.sqrt(a / (p + b))

    f = lambda p: fk(p, dl, pl) + fk(p, dr, pr) + (ur - ul)
    return brentq(f, 1e-14, 1e3, xtol=1e-300, rtol=1e-15, maxiter=500)


def test_trrs_estimate_is_exact_for_two_rarefactions():
    """Where both waves are rarefactions the two-rarefaction estimate is the
    exact star pressure (Toro 2009 sec. 9.4.1); checked against the closed
    form of eq. 4.46 and the root of the pressure function."""
    rng = np.random.default_rng(41)
    gamma = 1.4
    z = (gamma - 1.0) / (2.0 * gamma)
    seen = 0
    for _ in range(2000):
        dl, dr = 10 ** rng.uniform(-1, 1, 2)
        pl, pr = 10 ** rng.uniform(-1, 1, 2)
        cl, cr = np.sqrt(gamma * pl / dl), np.sqrt(gamma * pr / dr)
        ul = rng.uniform(-1.0, 0.5)
        ur = ul + rng.uniform(0.0, 1.5) * (cl + cr)     # diverging, no vacuum
        if ur - ul >= 2.0 * (cl + cr) / (gamma - 1.0):
            continue
        sl, sc, sr, ps, br = oracle.get_waves_pstar(dl, ul, pl, dr, ur, pr, gamma)
        p_exact = _exact_pstar(gamma, dl, ul, pl, dr, ur, pr)
        if br != 1 or p_exact >= min(pl, pr):
            continue
        closed = ((cl + cr - 0.5 * (gamma - 1.0) * (ur - ul)) /
                  (cl / pl ** z + cr / pr ** z)) ** (1.0 / z)
        np.testing.assert_allclose(ps, closed, rtol=1e-12)
        np.testing.assert_allclose(ps, p_exact, rtol=1e-11)
        # both waves are rarefactions: the speeds are the data's sound waves
        assert sl == ul - cl and sr == ur + cr
        seen += 1
    assert seen >= 200


def _pvrs(gamma, dl, ul, pl, dr, ur, pr):
    """Toro 2009 eq. 9.20 (primitive-variable estimate), floored at zero."""
    cl, cr = np.sqrt(gamma * pl / dl), np.sqrt(gamma * pr / dr)
    return max(0.0, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (dl + dr) * (cl + cr))


def _tsrs(gamma, dl, ul, pl, dr, ur, pr, p0):
    """Toro 2009 eq. 9.42 (two-shock estimate) about p0."""
    def g(d, pk):
        a, b = 2.0 / ((gamma + 1.0) * d), (gamma - 1.0) / (gamma + 1.0) * pk
        return np.sqrt(a / (p0 + b))
    gl, gr = g(dl, pl), g(dr, pr)
    return (gl * pl + gr * pr - (ur - ul)) / (gl + gr)


@pytest.mark.parametrize("ratio", [1.5, 1.9, 1.99, 2.0, 2.01, 2.1, 3.0, 5.0])
def test_wave_estimate_switch_follows_q2(ratio):
    """Toro 2009 sec. 9.5 with Q_user = 2: the primitive-variable estimate
    when p_max/p_min < 2 and it lies between p_min and p_max, the
    two-rarefaction estimate when it is at most p_min, the two-shock estimate
    about it otherwise."""
    rng = np.random.default_rng(int(ratio * 1000))
    gamma = 1.4
    for _ in range(100):
        dl, dr = 10 ** rng.uniform(-1, 1, 2)
        pmin = 10 ** rng.uniform(-1, 0.5)
        pl, pr = (pmin * ratio, pmin) if rng.random() < 0.5 else (pmin, pmin * ratio)
        ul = rng.uniform(-1, 1)
        # velocities close enough that the primitive-variable estimate stays
        # strictly between p_min and p_max
        ur = ul + rng.uniform(-0.05, 0.05) * min(pl, pr) / (dl + dr)
        ppv = _pvrs(gamma, dl, ul, pl, dr, ur, pr)
        assert min(pl, pr) < ppv < max(pl, pr)
        sl, sc, sr, ps, br = oracle.get_waves_pstar(dl, ul, pl, dr, ur, pr, gamma)
        if ratio < 2.0:
            assert br == 0, (ratio, pl, pr)
            np.testing.assert_allclose(ps, ppv, rtol=1e-13)
        else:
            assert br == 2, (ratio, pl, pr)
            np.testing.assert_allclose(ps, _tsrs(gamma, dl, ul, pl, dr, ur, pr, ppv), rtol=1e-13)
        # fastest signal speeds from the estimate (Toro 2009 eq. 10.59-10.60)
        cl, cr = np.sqrt(gamma * pl / dl), np.sqrt(gamma * pr / dr)
        ql = 1.0 if ps <= pl else np.sqrt(1.0 + (gamma + 1.0) / (2.0 * gamma) * (ps / pl - 1.0))
        qr = 1.0 if ps <= pr else np.sqrt(1.0 + (gamma + 1.0) / (2.0 * gamma) * (ps / pr - 1.0))
        np.testing.assert_allclose([sl, sr], [ul - cl * ql, ur + cr * qr], rtol=1e-13, atol=1e-13)


def test_wave_estimate_two_rarefaction_branch_when_below_pmin():
    """Strongly diverging data put the primitive-variable estimate below
    p_min: the two-rarefaction branch, also for p_max/p_min < 2."""
    gamma = 1.4
    for ratio in (1.2, 1.8, 4.0):
        dl, dr, pl, pr = 1.0, 0.5, 1.0, 1.0 / ratio
        ul, ur = -0.8, 0.9
        assert _pvrs(gamma, dl, ul, pl, dr, ur, pr) <= min(pl, pr)
        _, _, _, ps, br = oracle.get_waves_pstar(dl, ul, pl, dr, ur, pr, gamma)
        assert br == 1
        np.testing.assert_allclose(ps, _exact_pstar(gamma, dl, ul, pl, dr, ur, pr), rtol=1e-11)
