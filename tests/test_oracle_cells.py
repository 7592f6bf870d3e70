"""Pins of the oracle's cell and face loops other than the Riemann solvers.

  gradients  (reconstruction.pyx:563-737): exact for linear fields (the
             divergence-theorem identity behind Springel 2010 eq. 21);
             constant fields give zero; the AREPO limiter keeps every face
             extrapolation inside the neighbour min/max and is maximal.
  ghost gradients (boundary.pyx:265-307): enumerated flip rule.
  add_spatial (741-921): the reference's constant-state fixture; linear
             fields reach the face centroid exactly; GIZMO bounds.
  add_temporal (923-1072): the primitive Euler equations in matrix form.
  update_from_fluxes (mesh.pyx:479-551): conservation; the first failing
             face in face order.
  EOS, CFL, generator and face velocities: closed forms.
"""
import json
import os

import numpy as np
import pytest

from cpu_mesh import ghost_copy, lattice
from oracle import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_known_answers.json")))


def _linear_prims(a, rng, base=None):
    dim = a["dim"]
    b = [rng.normal(size=dim) for _ in range(dim + 2)]
    base = base or [2.0, 0.3, -0.2, 0.1, 3.0][:dim + 2]
    if dim == 3:
        base = [2.0, 0.3, -0.2, 0.1, 3.0]
    else:
        base = [2.0, 0.3, -0.2, 3.0]
    return [base[n] + sum(0.2 * b[n][k] * a["x"][k] for k in range(dim))
            for n in range(dim + 2)], b


@pytest.mark.parametrize("dim,n", [(2, 40), (3, 12)])
def test_gradient_exact_for_linear_fields(dim, n):
    a = lattice(n, dim, seed=3, boundary=1)   # periodic images keep the field linear
    rng = np.random.default_rng(0)
    a["prim"], b = _linear_prims(a, rng)
    grad, rc, _ = oracle.compute_gradients(a, limiter=1)   # TESS limits density only
    assert rc == 0
    nr = a["n_real"]
    for f in range(1, dim + 2):
        for k in range(dim):
            np.testing.assert_allclose(grad[f * dim + k][:nr], 0.2 * b[f][k], rtol=0, atol=5e-12)


def test_gradient_zero_for_constant_field():
    a = lattice(10, 3, seed=1, boundary=0)
    a["prim"] = [np.full(a["n"], v) for v in (1.0, 0.5, -0.5, 0.25, 2.0)]
    grad, rc, _ = oracle.compute_gradients(a, limiter=0)
    assert rc == 0
    for g in grad:
        assert np.abs(g[:a["n_real"]]).max() < 1e-12


@pytest.mark.parametrize("dim,n", [(2, 30), (3, 10)])
def test_arepo_limiter_bounds_and_maximality(dim, n):
    a = lattice(n, dim, seed=4, boundary=1)
    rng = np.random.default_rng(1)
    nf = dim + 2
    a["prim"] = [1.0 + 0.3 * rng.random(a["n"]) for _ in range(nf)]
    a["prim"] = ghost_copy(a, a["prim"], flip_velocity=False)
    lim, rc, _ = oracle.compute_gradients(a, limiter=0)
    free, rc2, _ = oracle.compute_gradients(a, limiter=1)   # fields >= 1 unlimited
    assert rc == 0 and rc2 == 0
    nr = a["n_real"]
    off, fl = a["nbr_off"], a["nbr_face"]
    pi, pj = a["pair_i"], a["pair_j"]
    for i in range(0, nr, 7):
        faces = fl[off[i]:off[i + 1]]
        nb = np.where(pi[faces] == i, pj[faces], pi[faces])
        cx = np.array([a["x"][k][i] + a["dcom"][k][i] for k in range(dim)])
        rel = np.stack([a["com"][k][faces] for k in range(dim)], 1) - cx
        for f in range(1, nf):
            phi = a["prim"][f]
            lo, hi = min(phi[i], phi[nb].min()), max(phi[i], phi[nb].max())
            g = np.array([lim[f * dim + k][i] for k in range(dim)])
            g0 = np.array([free[f * dim + k][i] for k in range(dim)])
            ext = phi[i] + rel @ g
            assert np.all(ext <= hi + 1e-12) and np.all(ext >= lo - 1e-12)
            # the limited gradient is alpha * unlimited with 0 <= alpha <= 1 ...
            nz = np.abs(g0) > 1e-300
            if nz.any():
                alpha = (g[nz] / g0[nz])[0]
                assert -1e-12 <= alpha <= 1 + 1e-12
                np.testing.assert_allclose(g, alpha * g0, rtol=1e-12, atol=1e-14)
                # ... and maximal: alpha < 1 only if some face touches a bound
                if alpha < 1 - 1e-12:
                    gap = np.minimum(hi - ext, ext - lo)
                    assert gap.min() < 1e-10


@pytest.mark.parametrize("dim,n", [(2, 30), (3, 10)])
def test_tess_limiter_scales_density_only(dim, n):
    """reconstruction.pyx:705-736: one alpha, min over every field and face
    of max((phi_j - phi_i)/dphi, 0), applied to the density gradient only.
    Density is given the pressure's values, so its unlimited gradient is the
    (unlimited) pressure gradient bit for bit."""
    a = lattice(n, dim, seed=6, boundary=1)
    rng = np.random.default_rng(2)
    nf = dim + 2
    a["prim"] = [1.0 + 0.3 * rng.random(a["n"]) for _ in range(nf)]
    a["prim"][0] = a["prim"][nf - 1].copy()
    a["prim"] = ghost_copy(a, a["prim"], flip_velocity=False)
    tess, rc, _ = oracle.compute_gradients(a, limiter=1)
    assert rc == 0
    nr = a["n_real"]
    off, fl = a["nbr_off"], a["nbr_face"]
    pi, pj = a["pair_i"], a["pair_j"]
    touched = 0
    for i in range(0, nr, 5):
        g0 = np.array([tess[(nf - 1) * dim + k][i] for k in range(dim)])
        gd = np.array([tess[k][i] for k in range(dim)])
        if np.abs(g0).max() < 1e-300:
            continue
        alpha = gd[np.argmax(np.abs(g0))] / g0[np.argmax(np.abs(g0))]
        assert -1e-15 <= alpha <= 1.0 + 1e-15
        np.testing.assert_allclose(gd, alpha * g0, rtol=1e-13, atol=1e-15)
        faces = fl[off[i]:off[i + 1]]
        nb = np.where(pi[faces] == i, pj[faces], pi[faces])
        cx = np.array([a["x"][k][i] + a["dcom"][k][i] for k in range(dim)])
        rel = np.stack([a["com"][k][faces] for k in range(dim)], 1) - cx
        tight = False
        for f in range(nf):
            gf = np.array([tess[f * dim + k][i] for k in range(dim)]) if f else g0
            dphi = rel @ gf
            jump = a["prim"][f][nb] - a["prim"][f][i]
            nz = dphi != 0.0
            psi = np.maximum(jump[nz] / dphi[nz], 0.0)
            # alpha never exceeds any face's ratio, for any field ...
            assert np.all(alpha <= psi + 1e-12)
            # ... and is attained by one of them (or is 1)
            tight |= bool(np.any(np.abs(psi - alpha) < 1e-12))
        assert tight or abs(alpha - 1.0) < 1e-15
        touched += alpha < 1.0
    assert touched > 0


def test_ghost_gradient_flip_rule():
    """boundary.pyx:291-307: only d v_q/d x_q flips, and only for exterior
    ghosts outside the box on axis q; periodic ghosts are pure copies."""
    dim, n = 2, 5
    x = [np.array([0.2, 0.7, -0.2, 0.7, 0.5]), np.array([0.3, 0.6, 0.3, 1.4, 0.5])]
    tag = np.array([0, 0, 1, 1, 1], np.int8)
    typ = np.array([4, 4, 2, 2, 2], np.int8)
    mp = np.array([0, 0, 0, 1, 1], np.int32)
    a = {"dim": dim, "n": n, "tag": tag, "type": typ, "map": mp, "x": x,
         "lo": [0.0, 0.0], "hi": [1.0, 1.0], "boundary": 0}
    G = dim * (dim + 2)
    base = [np.array([10.0 * g + 1, 10.0 * g + 2, 0, 0, 0]) for g in range(G)]
    grad = oracle.ghost_gradients(a, [b.copy() for b in base])
    want = [b.copy() for b in base]
    for g in range(G):
        want[g][2:] = base[g][mp[2:]]
    # ghost 2 is left of x=0: flips dv_x/dx = component (1+0)*2+0 = 2
    want[2][2] *= -1
    # ghost 3 is above y=1: flips dv_y/dy = component (1+1)*2+1 = 5
    want[5][3] *= -1
    for g in range(G):
        np.testing.assert_array_equal(grad[g], want[g])
    a["boundary"] = 1
    grad = oracle.ghost_gradients(a, [b.copy() for b in base])
    for g in range(G):
        w = base[g].copy()
        w[2:] = base[g][mp[2:]]
        np.testing.assert_array_equal(grad[g], w)


def test_ghost_refresh_eos_hand_placed():
    """domain_manager.pyx:757-774 + boundary.pyx:336-348 + the EOS: each
    ghost takes its image's conserved values, momentum (and velocity) flip on
    each axis where an exterior ghost is outside the box -- twice for a
    corner ghost -- and the primitives follow from the conserved fields
    (checked against the closed-form EOS on exact binary values)."""
    dim, n = 2, 6
    x = [np.array([0.25, 0.75, -0.25, 0.75, -0.25, 0.5]),
         np.array([0.25, 0.5, 0.25, 1.5, -0.5, 0.5])]
    tag = np.array([0, 0, 1, 1, 1, 1], np.int8)
    typ = np.array([4, 4, 2, 2, 2, 2], np.int8)
    mp = np.array([0, 0, 0, 1, 0, 1], np.int32)
    vol = np.array([0.25, 0.5, 0.25, 0.5, 0.25, 0.5])
    cons = [np.array([0.5, 1.0, 9, 9, 9, 9.0]),          # mass
            np.array([0.25, -0.5, 9, 9, 9, 9.0]),        # momentum-x
            np.array([-0.125, 1.0, 9, 9, 9, 9.0]),       # momentum-y
            np.array([2.0, 4.0, 9, 9, 9, 9.0])]          # energy
    a = {"dim": dim, "n": n, "tag": tag, "type": typ, "map": mp, "x": x, "volume": vol,
         "lo": [0.0, 0.0], "hi": [1.0, 1.0], "boundary": 0, "gamma": 1.4}
    c, p = oracle.ghost_refresh_eos(a, cons)
    sign = {2: (-1.0, 1.0), 3: (1.0, -1.0), 4: (-1.0, -1.0), 5: (1.0, 1.0)}
    for i in range(2, 6):
        s = mp[i]
        assert c[0][i] == cons[0][s] and c[3][i] == cons[3][s]
        assert c[1][i] == sign[i][0] * cons[1][s] and c[2][i] == sign[i][1] * cons[2][s]
    for i in range(n):
        d = c[0][i] / vol[i]
        v = [c[1][i] / c[0][i], c[2][i] / c[0][i]]
        pres = (c[3][i] / vol[i] - 0.5 * d * (v[0] * v[0] + v[1] * v[1])) * 0.4
        assert p[0][i] == d and p[1][i] == v[0] and p[2][i] == v[1]
        assert abs(p[3][i] - pres) <= 1e-15 * abs(pres)
    # cell 0: rho 2, v (0.5, -0.25), E/V 8 -> p = (8 - 0.3125) * 0.4
    assert p[0][0] == 2.0 and p[1][0] == 0.5 and p[2][0] == -0.25
    assert abs(p[3][0] - 3.075) < 1e-15
    a["boundary"] = 1
    c, p = oracle.ghost_refresh_eos(a, cons)
    for q in range(4):
        np.testing.assert_array_equal(c[q][2:], cons[q][mp[2:]])


def test_constant_state_reconstruction_fixture():
    v = GOLD["constant_state_reconstruction"]["value"]
    a = lattice(8, 2, seed=0)
    a["prim"] = [np.full(a["n"], v) for _ in range(4)]
    grad = [np.zeros(a["n"]) for _ in range(8)]
    for gizmo in (True, False):
        pl, pr = oracle.add_spatial(a, grad, boost=False, gizmo=gizmo)
        for q in range(4):
            assert np.all(pl[q] == v) and np.all(pr[q] == v)
        oracle.add_temporal(a, grad, 0.01, pl, pr, boost=False)
        for q in range(4):
            assert np.all(pl[q] == v) and np.all(pr[q] == v)
    pl, pr = oracle.states_constant(a, boost=False)
    for q in range(4):
        assert np.all(pl[q] == v) and np.all(pr[q] == v)


@pytest.mark.parametrize("dim,n", [(2, 16), (3, 6)])
def test_spatial_linear_field_reaches_face_centroid(dim, n):
    a = lattice(n, dim, seed=2, boundary=0)
    rng = np.random.default_rng(4)
    nf = dim + 2
    b = [rng.normal(size=dim) for _ in range(nf)]
    c0 = [2.0] + [0.1] * dim + [3.0]
    cen = [a["x"][k] + a["dcom"][k] for k in range(dim)]
    a["prim"] = [c0[q] + sum(0.1 * b[q][k] * cen[k] for k in range(dim)) for q in range(nf)]
    grad = [np.full(a["n"], 0.1 * b[q][k]) for q in range(nf) for k in range(dim)]
    pl, pr = oracle.add_spatial(a, grad, boost=False, gizmo=False)
    for q in range(nf):
        want = c0[q] + sum(0.1 * b[q][k] * a["com"][k] for k in range(dim))
        np.testing.assert_allclose(pl[q], want, rtol=0, atol=1e-13)
        np.testing.assert_allclose(pr[q], want, rtol=0, atol=1e-13)


def test_gizmo_limiter_bounds():
    """Hopkins (2015) B4: the limited state lies in [phi_minus, phi_plus]."""
    a = lattice(20, 2, seed=5, boundary=1)
    rng = np.random.default_rng(6)
    a["prim"] = [0.5 + rng.random(a["n"]) for _ in range(4)]
    a["prim"] = ghost_copy(a, a["prim"], flip_velocity=False)
    grad = [rng.normal(size=a["n"]) * 3.0 for _ in range(8)]
    pl, pr = oracle.add_spatial(a, grad, boost=False, gizmo=True)
    sl, sr = oracle.add_spatial(a, [np.zeros(a["n"]) for _ in range(8)], boost=False, gizmo=True)
    for q in range(4):
        d1 = 0.5 * np.abs(sl[q] - sr[q])
        lo, hi = np.minimum(sl[q], sr[q]) - d1, np.maximum(sl[q], sr[q]) + d1
        assert np.all(pl[q] >= lo - 1e-14) and np.all(pl[q] <= hi + 1e-14)
        assert np.all(pr[q] >= lo - 1e-14) and np.all(pr[q] <= hi + 1e-14)


def _two_cell_face(prim_i, prim_j, grad_i, grad_j, dcom_i, dcom_j, com):
    """One 2D face between a cell at the origin and one at (1, 0)."""
    a = {"dim": 2, "nf": 1, "n": 2,
         "prim": [np.array([prim_i[q], prim_j[q]]) for q in range(4)],
         "x": [np.array([0.0, 1.0]), np.array([0.0, 0.0])],
         "dcom": [np.array([dcom_i[k], dcom_j[k]]) for k in range(2)],
         "com": [np.array([com[k]]) for k in range(2)],
         "fv": [np.zeros(1), np.zeros(1)],
         "pair_i": np.array([0], dtype=np.int32), "pair_j": np.array([1], dtype=np.int32)}
    grad = [np.array([grad_i[g], grad_j[g]]) for g in range(8)]
    pl, pr = oracle.add_spatial(a, grad, boost=False, gizmo=True)
    return [float(x[0]) for x in pl], [float(x[0]) for x in pr]


def test_gizmo_limiter_hand_worked_case():
    """Hopkins (2015) appendix B4 worked by hand on one face whose cell
    centroids sit at unequal distances from the face centroid: generators at
    x = 0 and 1 (|x_j - x_i| = 1), centroids at x = -0.25 and 1, face centroid
    at x = 0.5, so |sep_i| = 0.75 and |sep_j| = 0.5.  psi1 = 1/2, psi2 = 1/4;
    phibar_l = phi_l + |sep_i|/|x_j - x_i| (phi_r - phi_l) and its mirror.
    Every value below is exact in binary except the one fallback quotient."""
    prim_i = [1.0, -1.0, 0.5, 3.0]
    prim_j = [2.0, -0.25, 0.5, 1.0]
    # per field (d/dx, d/dy): density 2, vx 2 / -2, vy 1, pressure -4 / -6
    grad_i = [2.0, 0.0, 2.0, 0.0, 1.0, 0.0, -4.0, 0.0]
    grad_j = [2.0, 0.0, -2.0, 0.0, 1.0, 0.0, -6.0, 0.0]
    L, R = _two_cell_face(prim_i, prim_j, grad_i, grad_j, (-0.25, 0.0), (0.0, 0.0), (0.5, 0.0))
    # density (phi_i < phi_j): delta1 = 1/2, delta2 = 1/4, phibar_l = 1.75,
    # phibar_r = 1.5; extrapolations 2.5 and 1.0 are cut to phibar -/+ delta2
    assert L[0] == 2.0 and R[0] == 1.25
    # vx (phi_i < phi_j, both negative): delta1 = 3/8, delta2 = 3/16;
    # phibar_l = -7/16 and the left extrapolation 0.5 is cut to
    # phibar_l + delta2 = -1/4; phi_max + delta1 changes sign, so
    # phi_plus = phi_max / (1 + delta1/|phi_max|) and caps the right
    # extrapolation 0.75
    assert L[1] == -0.25 and R[1] == -0.25 / 2.5
    # vy equal on both sides: the unextrapolated states
    assert L[2] == 0.5 and R[2] == 0.5
    # pressure (phi_i > phi_j): delta1 = 1, delta2 = 1/2, phibar_l = 1.5,
    # phibar_r = 2; extrapolations 0 and 4 are cut to phibar_l - delta2 and
    # phibar_r + delta2
    assert L[3] == 1.0 and R[3] == 2.5
    # the same face with the roles of the cells swapped (centroid offsets
    # swapped with them) mirrors the result
    L2, R2 = _two_cell_face(prim_j, prim_i, [-g for g in grad_j], [-g for g in grad_i],
                            (0.0, 0.0), (0.25, 0.0), (0.5, 0.0))
    assert L2 == R and R2 == L


@pytest.mark.parametrize("dim", [2, 3])
def test_temporal_is_primitive_euler_predictor(dim):
    """L = W - dt * sum_k A_k(W) dW/dx_k with the face-frame velocity in the
    advective terms; A_k is the primitive-variable Jacobian of the Euler
    equations (Toro eq. 3.65)."""
    a = lattice(6 if dim == 3 else 12, dim, seed=8, boundary=1)
    rng = np.random.default_rng(9)
    nf = dim + 2
    n = a["n"]
    a["gamma"] = 1.4
    a["prim"] = [1.0 + rng.random(n)] + [rng.uniform(-1, 1, n) for _ in range(dim)] + [1.0 + rng.random(n)]
    a["fv"] = [rng.uniform(-0.3, 0.3, a["nf"]) for _ in range(dim)]
    grad = [0.2 * rng.normal(size=n) for _ in range(nf * dim)]
    dt = 1e-3
    pl, pr = oracle.states_constant(a, boost=True)
    pl0 = [x.copy() for x in pl]
    pr0 = [x.copy() for x in pr]
    oracle.add_temporal(a, grad, dt, pl, pr, boost=True)
    for side, pos, S0, S in (("L", a["pair_i"], pl0, pl), ("R", a["pair_j"], pr0, pr)):
        for m in range(0, a["nf"], 5):
            c = pos[m]
            rho, p = a["prim"][0][c], a["prim"][dim + 1][c]
            u = np.array([a["prim"][1 + k][c] - a["fv"][k][m] for k in range(dim)])
            dW = np.array([[grad[q * dim + k][c] for k in range(dim)] for q in range(nf)])
            rhs = np.zeros(nf)
            for k in range(dim):
                A = np.zeros((nf, nf))
                A[0, 0] = u[k]
                A[0, 1 + k] = rho
                for q in range(dim):
                    A[1 + q, 1 + q] = u[k]
                A[1 + k, nf - 1] = 1.0 / rho
                A[nf - 1, 1 + k] = a["gamma"] * p
                A[nf - 1, nf - 1] = u[k]
                rhs += A @ dW[:, k]
            want = np.array([S0[q][m] for q in range(nf)]) - dt * rhs
            got = np.array([S[q][m] for q in range(nf)])
            if want[0] < 0 or want[-1] < 0:
                continue
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14, err_msg=side)


def test_update_conservation_and_first_failure():
    a = lattice(10, 3, seed=3, boundary=1)
    rng = np.random.default_rng(7)
    nf = a["nf"]
    a["cons"] = [10.0 + rng.random(a["n"]) for _ in range(5)]
    flux = [rng.normal(size=nf) for _ in range(5)]
    dt = 1e-3
    cons, rc, _ = oracle.update_from_fluxes(a, flux, dt)
    assert rc == 0
    nr = a["n_real"]
    ghost_face = a["pair_j"] >= nr
    for q in range(5):
        got = cons[q][:nr].sum() - a["cons"][q][:nr].sum()
        want = -(dt * a["area"][ghost_face] * flux[q][ghost_face]).sum()
        assert abs(got - want) < 1e-10 * np.abs(a["cons"][q][:nr]).sum() * 1e-2
    # a face that drains one cell's mass: the first such face is reported
    flux[0][:] = 0.0
    k = nf // 2
    i = a["pair_i"][k]
    flux[0][k] = 2.0 * a["cons"][0][i] / (dt * a["area"][k])
    cons, rc, bad = oracle.update_from_fluxes(a, flux, dt)
    assert rc == 2 and bad == k
    flux[0][k] = 0.0
    flux[4][k] = 2.0 * a["cons"][4][i] / (dt * a["area"][k])
    cons, rc, bad = oracle.update_from_fluxes(a, flux, dt)
    assert rc == 3 and bad == k


def test_eos_round_trip_and_known_values():
    gamma = 1.4
    prim = [np.array([1.0, 2.0]), np.array([1.0, -0.5]), np.array([0.0, 0.25]), np.array([1.0, 3.0])]
    vol = np.array([2.0, 0.5])
    cons = oracle.cons_from_prim(2, gamma, prim, vol)
    np.testing.assert_allclose([c[0] for c in cons], [2.0, 2.0, 0.0, (0.5 + 1.0 / 0.4) * 2.0])
    back = oracle.prim_from_cons(2, gamma, cons, vol)
    for q in range(4):
        np.testing.assert_allclose(back[q], prim[q], rtol=1e-14, atol=1e-15)


def test_cfl_minimum():
    gamma, cfl = 1.4, 0.3
    d = np.array([1.0, 1.0, 1.0, 5.0])
    p = np.array([1.0, 2.0, 1.0, 1.0])
    v0 = np.array([0.0, 0.5, 0.0, 0.0])
    vol = np.array([0.01, 0.01, 0.0001, 0.0001])
    tag = np.array([0, 0, 0, 1], np.int8)        # the ghost (tiny, slow) is ignored
    prim = [d, v0, np.zeros(4), p]
    dt = oracle.time_step(2, gamma, cfl, True, tag, prim, vol)
    np.testing.assert_allclose(dt, cfl * np.sqrt(0.0001 / np.pi) / np.sqrt(gamma), rtol=1e-15)
    dt = oracle.time_step(2, gamma, cfl, False, tag, prim, vol)
    np.testing.assert_allclose(dt, cfl * np.sqrt(0.0001 / np.pi) / np.sqrt(gamma), rtol=1e-15)
    vol3 = np.array([1.0, 1.0, 8.0, 1e-6])
    dt3 = oracle.time_step(3, gamma, 1.0, True, np.zeros(4, np.int8),
                           [d, v0, np.zeros(4), np.zeros(4), p], vol3)
    np.testing.assert_allclose(dt3, (3e-6 / (4 * np.pi)) ** (1 / 3) / np.sqrt(gamma / 5.0),
                               rtol=1e-14)


def test_generator_velocity_regularisation_branches():
    """Springel (2010) eq. 63: no correction below 0.9 eta R, a linear ramp
    up to 1.1 eta R, then |w - v| = c along dcom."""
    gamma, eta = 1.4, 0.25
    vol = np.full(3, np.pi)                       # R = 1 in 2D
    R = 1.0
    d = np.array([0.5, 1.0, 2.0]) * eta * R      # below, inside and above the ramp
    dcom = [d * 0.6, d * 0.8]
    prim = [np.ones(3), np.array([0.1, 0.2, 0.3]), np.zeros(3), np.ones(3)]
    w = oracle.generator_velocities(2, gamma, eta, True, prim, dcom, vol)
    c = np.sqrt(gamma)
    dw = np.stack([w[0] - prim[1], w[1] - prim[2]], 1)
    assert np.all(dw[0] == 0)
    np.testing.assert_allclose(np.linalg.norm(dw[1]), c * (1.0 - 0.9) / 0.2, rtol=1e-14)
    np.testing.assert_allclose(np.linalg.norm(dw[2]), c, rtol=1e-14)
    np.testing.assert_allclose(dw[2] / c, [0.6, 0.8], rtol=1e-14)
    w0 = oracle.generator_velocities(2, gamma, eta, False, prim, dcom, vol)
    np.testing.assert_array_equal(w0[0], prim[1])


def test_face_velocity_is_bisector_normal_velocity():
    """Springel (2010) eq. 32-33: the face velocity's normal component is the
    normal velocity of the bisector plane of two moving generators."""
    a = lattice(10, 3, seed=2, boundary=0)
    rng = np.random.default_rng(3)
    w = [rng.normal(size=a["n"]) for _ in range(3)]
    fv = oracle.face_velocities(3, a["x"], w, a["com"], a["pair_i"], a["pair_j"])
    i, j = a["pair_i"], a["pair_j"]
    X = np.stack(a["x"], 1)
    W = np.stack(w, 1)
    dx = X[j] - X[i]
    nrm = dx / np.linalg.norm(dx, axis=1)[:, None]
    mid = 0.5 * (X[i] + X[j])
    f = np.stack(a["com"], 1)
    # d/dt of (y - mid).(x_j - x_i) = 0 for the moving plane
    want = (0.5 * (W[i] + W[j]) * dx).sum(1) + ((W[i] - W[j]) * (f - mid)).sum(1)
    want /= np.linalg.norm(dx, axis=1)
    got = (np.stack(fv, 1) * nrm).sum(1)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13)
