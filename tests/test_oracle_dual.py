"""Pins of the forward-mode (dual-number) build of the oracle
(oracle/oracle_dual.cpp), the taped reference the GPU adjoint is checked
against: its values are the plain oracle's step by step, bit for bit, and its
directional derivatives match central differences of the plain oracle's
rollout away from branch switches."""
import numpy as np
import pytest

from cpu_mesh import ghost_copy, lattice


def _problem(n, dim, boundary, seed):
    a = lattice(n, dim, seed=seed, boundary=boundary)
    a["gamma"] = 5.0 / 3.0 if boundary == 1 else 1.4
    rng = np.random.default_rng(seed + 7)
    x = a["x"]
    na = a["n"]
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * x[0]) * np.cos(2 * np.pi * x[1]) + 0.02 * rng.normal(size=na)
    vel = [0.2 * np.cos(2 * np.pi * x[(k + 1) % dim]) + 0.02 * rng.normal(size=na) for k in range(dim)]
    p = 1.0 + 0.1 * np.cos(2 * np.pi * x[0]) + 0.02 * rng.normal(size=na)
    return a, [rho] + vel + [p]


def _plain_rollout(a, prim0, dts):
    from oracle import oracle
    dim, g = a["dim"], a["gamma"]
    cons = oracle.cons_from_prim(dim, g, prim0, a["volume"])

    def refresh(cons):
        cons = ghost_copy(a, cons)
        return cons, oracle.prim_from_cons(dim, g, cons, a["volume"])
    cons, prim = refresh(cons)
    for dt in dts:
        b = dict(a, prim=prim, cons=cons)
        r = oracle.core_step(b, dt, solver=1, boost=True, gizmo=True, limiter=0)
        assert r["rc"] == 0
        cons, prim = refresh(r["cons"])
    return prim


CASES = [(10, 2, 0, 1), (10, 2, 1, 2), (4, 3, 0, 3)]


@pytest.mark.parametrize("n,dim,boundary,seed", CASES)
def test_dual_values_are_the_oracle(n, dim, boundary, seed):
    from oracle import oracle
    a, prim0 = _problem(n, dim, boundary, seed)
    dts = [2e-3, 1.5e-3, 2.5e-3]
    ref = _plain_rollout(a, prim0, dts)
    got, _ = oracle.rollout_dual(a, prim0, [np.zeros_like(q) for q in prim0], dts)
    for q in range(dim + 2):
        assert np.array_equal(got[q], ref[q]), q


@pytest.mark.parametrize("n,dim,boundary,seed", CASES)
def test_dual_tangents_match_central_differences(n, dim, boundary, seed):
    from oracle import oracle
    a, prim0 = _problem(n, dim, boundary, seed)
    dts = [2e-3, 1.5e-3, 2.5e-3]
    real = a["tag"] == 0
    rng = np.random.default_rng(seed + 11)
    w = rng.normal(size=a["n"]) * real
    checked = 0
    for trial in range(6):
        v = [rng.normal(size=a["n"]) * real * (0.1 if q else 1.0) for q in range(dim + 2)]
        _, dout = oracle.rollout_dual(a, prim0, v, dts)
        dL = float(np.dot(w, dout[0]))

        def L(h):
            pr = _plain_rollout(a, [p + h * vq for p, vq in zip(prim0, v)], dts)
            return float(np.dot(w, pr[0]))
        l0 = L(0.0)
        fds, gaps = [], []
        for k in (1, 2, 3):
            h = k * 1e-7
            lp, lm = L(h), L(-h)
            fds.append((lp - lm) / (2 * h))
            gaps.append((lp - l0) / h - (l0 - lm) / h)
        scale = max(abs(fds[0]), 1e-8)
        if any(abs(gaps[k - 1] - k * gaps[0]) > 0.05 * abs(gaps[k - 1]) + 1e-6 * scale
               for k in (2, 3)):
            continue                       # a branch switch inside the stencil
        assert abs(fds[0] - dL) <= 1e-5 * scale, (trial, fds[0], dL)
        checked += 1
    assert checked >= 3


def test_dual_tangent_is_linear():
    from oracle import oracle
    a, prim0 = _problem(10, 2, 0, 5)
    dts = [2e-3, 2e-3]
    rng = np.random.default_rng(3)
    real = a["tag"] == 0
    v1 = [rng.normal(size=a["n"]) * real for _ in prim0]
    v2 = [rng.normal(size=a["n"]) * real for _ in prim0]
    _, d1 = oracle.rollout_dual(a, prim0, v1, dts)
    _, d2 = oracle.rollout_dual(a, prim0, v2, dts)
    _, d3 = oracle.rollout_dual(a, prim0, [2.0 * x - 0.5 * y for x, y in zip(v1, v2)], dts)
    for q in range(4):
        ref = 2.0 * d1[q] - 0.5 * d2[q]
        assert np.allclose(d3[q], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("n,dim,boundary,seed", CASES)
def test_openmp_build_gives_the_serial_bits(n, dim, boundary, seed):
    """libphd_oracle_omp.so (the all-cores CPU line) runs the independent
    loops in parallel; every output equals the serial oracle's bit for bit."""
    from oracle import oracle
    a, prim0 = _problem(n, dim, boundary, seed)
    a = dict(a, prim=prim0, cons=oracle.cons_from_prim(dim, a["gamma"], prim0, a["volume"]))
    r1 = oracle.core_step(a, 2e-3)
    r2 = oracle.core_step(a, 2e-3, omp=True)
    assert r1["rc"] == r2["rc"] == 0
    for key in ("grad", "left", "right", "flux", "cons"):
        for q, (x, y) in enumerate(zip(r1[key], r2[key])):
            assert np.array_equal(x, y), (key, q)
    assert oracle.omp_threads() >= 1
