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
    x/t = 0 with the closed-form star densities.
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
