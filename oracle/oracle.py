"""ctypes binding of the CPU oracle (oracle/phd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Arrays are host numpy arrays
in the layout produced by phd.problems.host_arrays (reference field order).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PHD_ORACLE_LIB names another build of phd_oracle.c for the serial oracle:
# tests/test_oracle_mutations.py points it at copies with a planted mistake
# to show that the pins catch it.
LIB_PATH = os.environ.get("PHD_ORACLE_LIB") or os.path.join(HERE, "libphd_oracle.so")
OMP_PATH = os.path.join(HERE, "libphd_oracle_omp.so")

_lib = None
_omp = None
dp = ctypes.POINTER(ctypes.c_double)


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return LIB_PATH


def lib(omp=False):
    """The serial oracle, or (omp=True) the same source with its independent
    loops run by OpenMP threads (same bits)."""
    global _lib, _omp
    if omp:
        if _omp is None:
            if not os.path.exists(OMP_PATH):
                build()
            _omp = ctypes.CDLL(OMP_PATH)
            _omp.orc_compute_time_step.restype = ctypes.c_double
        return _omp
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_compute_time_step.restype = ctypes.c_double
    return _lib


def omp_threads():
    return int(lib(omp=True).orc_threads())


def _pp(arrs):
    """double** over a list of float64 arrays (kept alive by the caller)."""
    a = (ctypes.c_void_p * max(1, len(arrs)))()
    for q, x in enumerate(arrs):
        assert x.dtype == np.float64 and x.flags["C_CONTIGUOUS"]
        a[q] = x.ctypes.data
    return a


def _p(x):
    return ctypes.c_void_p(x.ctypes.data) if x is not None and x.size else ctypes.c_void_p(0)


def _f64(lst):
    return [np.ascontiguousarray(x, dtype=np.float64) for x in lst]


def cons_from_prim(dim, gamma, prim, vol):
    prim = _f64(prim)
    cons = [np.zeros_like(prim[0]) for _ in range(dim + 2)]
    lib().orc_cons_from_prim(dim, prim[0].size, ctypes.c_double(gamma), _pp(prim), _p(vol),
                             _pp(cons))
    return cons


def prim_from_cons(dim, gamma, cons, vol):
    cons = _f64(cons)
    prim = [np.zeros_like(cons[0]) for _ in range(dim + 2)]
    lib().orc_prim_from_cons(dim, cons[0].size, ctypes.c_double(gamma), _pp(cons), _p(vol),
                             _pp(prim))
    return prim


def time_step(dim, gamma, cfl, boost, tag, prim, vol):
    prim = _f64(prim)
    return lib().orc_compute_time_step(dim, prim[0].size, ctypes.c_double(gamma),
                                       ctypes.c_double(cfl), int(boost), _p(tag), _pp(prim),
                                       _p(vol))


def generator_velocities(dim, gamma, eta, regularize, prim, dcom, vol):
    prim, dcom = _f64(prim), _f64(dcom)
    w = [np.zeros_like(prim[0]) for _ in range(dim)]
    lib().orc_generator_velocities(dim, prim[0].size, ctypes.c_double(gamma),
                                   ctypes.c_double(eta), int(regularize), _pp(prim), _pp(dcom),
                                   _p(vol), _pp(w))
    return w


def face_velocities(dim, x, w, com, pair_i, pair_j):
    x, w, com = _f64(x), _f64(w), _f64(com)
    fv = [np.zeros_like(com[0]) for _ in range(dim)]
    lib().orc_face_velocities(dim, com[0].size, _pp(x), _pp(w), _pp(com), _p(pair_i),
                              _p(pair_j), _pp(fv))
    return fv


def compute_gradients(a, limiter=0):
    """Returns (grad list of dim*(dim+2) arrays, rc, bad_cell)."""
    dim, n = a["dim"], a["n"]
    grad = [np.zeros(n) for _ in range(dim * (dim + 2))]
    bad = ctypes.c_int32(-1)
    rc = lib().orc_compute_gradients(dim, n, _p(a["tag"]), _pp(a["x"]), _pp(a["dcom"]),
                                     _p(a["volume"]), _pp(a["prim"]), _p(a["area"]),
                                     _pp(a["com"]), _p(a["pair_i"]), _p(a["pair_j"]),
                                     _p(a["nbr_off"]), _p(a["nbr_face"]), int(limiter),
                                     _pp(grad), ctypes.byref(bad))
    return grad, rc, bad.value


def ghost_gradients(a, grad):
    dim = a["dim"]
    lo = np.array(a["lo"] + [0.0] * (3 - dim))
    hi = np.array(a["hi"] + [0.0] * (3 - dim))
    lib().orc_ghost_gradients(dim, a["n"], _p(a["tag"]), _p(a["type"]), _p(a["map"]),
                              _pp(a["x"]), _p(lo), _p(hi), int(a["boundary"]), _pp(grad))
    return grad


def add_spatial(a, grad, boost=True, gizmo=True):
    dim, nf = a["dim"], a["nf"]
    pl = [np.zeros(nf) for _ in range(dim + 2)]
    pr = [np.zeros(nf) for _ in range(dim + 2)]
    lib().orc_add_spatial(dim, nf, _pp(a["prim"]), _pp(a["x"]), _pp(a["dcom"]), _pp(grad),
                          _pp(a["com"]), _pp(a["fv"]), _p(a["pair_i"]), _p(a["pair_j"]),
                          int(boost), int(gizmo), _pp(pl), _pp(pr))
    return pl, pr


def add_temporal(a, grad, dt, pl, pr, boost=True):
    dim, nf = a["dim"], a["nf"]
    lib().orc_add_temporal(dim, nf, ctypes.c_double(a["gamma"]), ctypes.c_double(dt),
                           _pp(a["prim"]), _pp(grad), _pp(a["fv"]), _p(a["pair_i"]),
                           _p(a["pair_j"]), int(boost), _pp(pl), _pp(pr))
    return pl, pr


def states_constant(a, boost=True):
    dim, nf = a["dim"], a["nf"]
    pl = [np.zeros(nf) for _ in range(dim + 2)]
    pr = [np.zeros(nf) for _ in range(dim + 2)]
    lib().orc_states_constant(dim, nf, _pp(a["prim"]), _pp(a["fv"]), _p(a["pair_i"]),
                              _p(a["pair_j"]), int(boost), _pp(pl), _pp(pr))
    return pl, pr


def riemann(solver, dim, gamma, boost, pl, pr, normal, fv):
    """solver: 'hll' | 'hllc' | 'exact'.  Returns (flux, rc, bad_face)."""
    pl, pr, normal, fv = _f64(pl), _f64(pr), _f64(normal), _f64(fv)
    nf = pl[0].size
    flux = [np.zeros(nf) for _ in range(dim + 2)]
    L = lib()
    if solver == "exact":
        bad = ctypes.c_int32(-1)
        rc = L.orc_exact(dim, nf, ctypes.c_double(gamma), _pp(pl), _pp(pr), _pp(normal),
                         _pp(fv), _pp(flux), ctypes.byref(bad))
        return flux, rc, bad.value
    fn = L.orc_hll if solver == "hll" else L.orc_hllc
    fn(dim, nf, ctypes.c_double(gamma), int(boost), _pp(pl), _pp(pr), _pp(normal), _pp(fv),
       _pp(flux))
    return flux, 0, -1


def get_waves(dl, ul, pl, dr, ur, pr, gamma):
    sl, sc, sr = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().orc_get_waves(*[ctypes.c_double(v) for v in (dl, ul, pl, dr, ur, pr, gamma)],
                        ctypes.byref(sl), ctypes.byref(sc), ctypes.byref(sr))
    return sl.value, sc.value, sr.value


def get_waves_pstar(dl, ul, pl, dr, ur, pr, gamma):
    """get_waves plus the star-pressure estimate its wave speeds are built
    from and the estimate used: 0 PVRS, 1 TRRS, 2 TSRS.  Returns
    (sl, sc, sr, p_star, branch)."""
    sl, sc, sr, ps = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    br = ctypes.c_int()
    lib().orc_get_waves_pstar(*[ctypes.c_double(v) for v in (dl, ul, pl, dr, ur, pr, gamma)],
                              ctypes.byref(sl), ctypes.byref(sc), ctypes.byref(sr),
                              ctypes.byref(ps), ctypes.byref(br))
    return sl.value, sc.value, sr.value, ps.value, br.value


def update_from_fluxes(a, flux, dt, cons=None):
    """Returns (cons after update, rc, bad_face)."""
    dim = a["dim"]
    cons = _f64([c.copy() for c in (a["cons"] if cons is None else cons)])
    bad = ctypes.c_int32(-1)
    rc = lib().orc_update_from_fluxes(dim, a["nf"], ctypes.c_double(dt), _p(a["tag"]),
                                      _p(a["area"]), _p(a["pair_i"]), _p(a["pair_j"]),
                                      _pp(_f64(flux)), _pp(cons), ctypes.byref(bad))
    return cons, rc, bad.value


def core_step(a, dt, solver=1, boost=True, gizmo=True, limiter=0, omp=False):
    """Full core step on copies; returns dict(grad, left, right, flux, cons, rc, bad)."""
    dim, n, nf = a["dim"], a["n"], a["nf"]
    grad = [np.zeros(n) for _ in range(dim * (dim + 2))]
    pl = [np.zeros(nf) for _ in range(dim + 2)]
    pr = [np.zeros(nf) for _ in range(dim + 2)]
    flux = [np.zeros(nf) for _ in range(dim + 2)]
    cons = _f64([c.copy() for c in a["cons"]])
    lo = np.array(a["lo"] + [0.0] * (3 - dim))
    hi = np.array(a["hi"] + [0.0] * (3 - dim))
    bad = ctypes.c_int32(-1)
    rc = lib(omp).orc_core_step(dim, n, nf, ctypes.c_double(a["gamma"]), ctypes.c_double(dt),
                             int(solver), int(boost), int(gizmo), int(limiter),
                             int(a["boundary"]), _p(a["tag"]), _p(a["type"]), _p(a["map"]),
                             _pp(a["x"]), _pp(a["dcom"]), _p(a["volume"]), _pp(a["prim"]),
                             _pp(cons), _p(a["area"]), _pp(a["com"]), _pp(a["normal"]),
                             _pp(a["fv"]), _p(a["pair_i"]), _p(a["pair_j"]), _p(a["nbr_off"]),
                             _p(a["nbr_face"]), _p(lo), _p(hi), _pp(grad), _pp(pl), _pp(pr),
                             _pp(flux), ctypes.byref(bad))
    return {"grad": grad, "left": pl, "right": pr, "flux": flux, "cons": cons, "rc": rc,
            "bad": bad.value}


def ghost_refresh_eos(a, cons, omp=False):
    """Static-mesh epilogue (orc_ghost_refresh_eos) on copies of the conserved
    fields `cons`; returns (cons, prim)."""
    dim, n = a["dim"], a["n"]
    cons = _f64([c.copy() for c in cons])
    prim = [np.zeros(n) for _ in range(dim + 2)]
    lo = np.array(a["lo"] + [0.0] * (3 - dim))
    hi = np.array(a["hi"] + [0.0] * (3 - dim))
    lib(omp).orc_ghost_refresh_eos(dim, n, ctypes.c_double(a["gamma"]), _p(a["tag"]),
                                   _p(a["type"]), _p(a["map"]), _pp(_f64(a["x"])), _p(lo), _p(hi),
                                   int(a["boundary"]), _p(a["volume"]), _pp(cons), _pp(prim))
    return cons, prim


def static_rollout(a, dts, solver=1, boost=True, gizmo=True, limiter=0, omp=False):
    """StaticMeshMUSCLHancock steps (integrate.py:349-378) with the given
    time steps: core step, then ghost refresh + EOS.  `a` holds the starting
    state (prim and cons consistent, ghosts refreshed) and is not changed.
    Returns (cons, prim, rc, bad, steps done)."""
    b = dict(a)
    b["prim"] = _f64([x.copy() for x in a["prim"]])
    b["cons"] = _f64([x.copy() for x in a["cons"]])
    for s, dt in enumerate(dts):
        r = core_step(b, float(dt), solver=solver, boost=boost, gizmo=gizmo, limiter=limiter,
                      omp=omp)
        if r["rc"] != 0:
            return r["cons"], b["prim"], r["rc"], r["bad"], s
        b["cons"], b["prim"] = ghost_refresh_eos(b, r["cons"], omp=omp)
    return b["cons"], b["prim"], 0, -1, len(dts)


DUAL_PATH = os.path.join(HERE, "libphd_oracle_dual.so")
_dual = None


def dual_lib():
    global _dual
    if _dual is None:
        if not os.path.exists(DUAL_PATH):
            build()
        _dual = ctypes.CDLL(DUAL_PATH)
    return _dual


def rollout_dual(a, prim0, dprim0, dts, boost=True, gizmo=True):
    """Static-mesh rollout (oracle_dual.cpp: ord_rollout) with tangents:
    initial primitives prim0 (all cells) and their tangents dprim0; time
    steps dts.  Returns (final primitives, their tangents), or raises on a
    fault."""
    dim, n, nf = a["dim"], a["n"], a["nf"]
    prim0, dprim0 = _f64(prim0), _f64(dprim0)
    out = [np.zeros(n) for _ in range(dim + 2)]
    dout = [np.zeros(n) for _ in range(dim + 2)]
    dt = np.ascontiguousarray(dts, dtype=np.float64)
    lo = np.array(a["lo"] + [0.0] * (3 - dim))
    hi = np.array(a["hi"] + [0.0] * (3 - dim))
    x, dcom = _f64(a["x"]), _f64(a["dcom"])
    com, normal = _f64(a["com"]), _f64(a["normal"])
    rc = dual_lib().ord_rollout(dim, n, nf, ctypes.c_double(a["gamma"]), len(dt), _p(dt),
                                int(boost), int(gizmo), int(a["boundary"]), _p(a["tag"]),
                                _p(a["type"]), _p(a["map"]), _pp(x), _pp(dcom),
                                _p(a["volume"]), _p(a["area"]), _pp(com), _pp(normal),
                                _p(a["pair_i"]), _p(a["pair_j"]), _p(a["nbr_off"]),
                                _p(a["nbr_face"]), _p(lo), _p(hi), _pp(prim0), _pp(dprim0),
                                _pp(out), _pp(dout))
    if rc != 0:
        raise RuntimeError("oracle rollout fault %d" % rc)
    return out, dout


def extract_geometry(dim, x, topo, lo, hi, boundary, omp=False):
    """Mesh geometry from Delaunay topology (orc_extract_geometry).

    x: the real generators' coordinates (dim arrays of n); topo: a dict with
    ghost_map, ghost_code, pair_i, pair_j, ring_off, ring_k, ring_code,
    nbr_off, nbr_face (phd.mesh.build_voronoi's keys).  Returns a dict with
    x (reals then ghost images), volume, dcom (n + n_ghost), area, com,
    normal (n_faces)."""
    n = x[0].size
    ng, nf = topo["ghost_map"].size, topo["pair_i"].size
    X = [np.concatenate([np.asarray(x[k], dtype=np.float64), np.zeros(ng)]) for k in range(dim)]
    out = {"x": X, "volume": np.zeros(n + ng), "dcom": [np.zeros(n + ng) for _ in range(dim)],
           "area": np.zeros(nf), "com": [np.zeros(nf) for _ in range(dim)],
           "normal": [np.zeros(nf) for _ in range(dim)]}
    ws = np.zeros(max(1, nf * 2 * (dim + 1)))
    lo_ = np.array(list(lo) + [0.0] * (3 - dim), dtype=np.float64)
    hi_ = np.array(list(hi) + [0.0] * (3 - dim), dtype=np.float64)
    i32 = {k: np.ascontiguousarray(topo[k], dtype=np.int32)
           for k in ("ghost_map", "pair_i", "pair_j", "ring_off", "ring_k", "nbr_off", "nbr_face")}
    i8 = {k: np.ascontiguousarray(topo[k], dtype=np.int8) for k in ("ghost_code", "ring_code")}
    lib(omp).orc_extract_geometry(
        dim, int(boundary), _p(lo_), _p(hi_), n, ng, nf, _pp(X), _p(i32["ghost_map"]),
        _p(i8["ghost_code"]), _p(i32["pair_i"]), _p(i32["pair_j"]), _p(i32["ring_off"]),
        _p(i32["ring_k"]), _p(i8["ring_code"]), _p(i32["nbr_off"]), _p(i32["nbr_face"]),
        _p(out["volume"]), _pp(out["dcom"]), _p(out["area"]), _pp(out["com"]),
        _pp(out["normal"]), _p(ws))
    return out
