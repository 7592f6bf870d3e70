"""Host-only problem arrays for the CPU tests: a Voronoi mesh from the host
builder plus ghost cells, laid out exactly like phd.problems.host_arrays
(reals first, then ghosts; reference field order), without touching a GPU."""
import numpy as np

from phd.mesh import build_voronoi
from phd.problems import jittered_lattice


def mesh_arrays(x, dim, boundary=0, lo=None, hi=None, gamma=1.4):
    lo = [0.0] * dim if lo is None else list(lo)
    hi = [1.0] * dim if hi is None else list(hi)
    g = build_voronoi(dim, x, lo, hi, boundary)
    nr, ng = g["n_real"], g["n_ghost"]
    n = nr + ng
    gm = g["ghost_map"]
    X = [np.concatenate([np.asarray(x[k], dtype=np.float64), g["ghost_x"][k]]) for k in range(dim)]
    nf = g["area"].size
    return {
        "dim": dim, "n": n, "nf": nf, "n_real": nr,
        "tag": np.concatenate([np.zeros(nr, np.int8), np.ones(ng, np.int8)]),
        "type": np.concatenate([np.full(nr, 4, np.int8), np.full(ng, 2, np.int8)]),
        "map": np.concatenate([np.zeros(nr, np.int32), gm.astype(np.int32)]),
        "x": X, "dcom": g["dcom"], "volume": g["volume"],
        "area": g["area"], "com": g["com"], "normal": g["normal"],
        "fv": [np.zeros(nf) for _ in range(dim)],
        "pair_i": g["pair_i"], "pair_j": g["pair_j"],
        "nbr_off": g["nbr_off"], "nbr_face": g["nbr_face"],
        "lo": lo, "hi": hi, "boundary": boundary, "gamma": gamma,
    }


def ghost_copy(a, field_list, flip_velocity=True):
    """Ghost fields = image fields, velocities negated on each axis where a
    reflective ghost lies outside (boundary.pyx:336-348)."""
    nr = a["n_real"]
    out = []
    for q, f in enumerate(field_list):
        f = f.copy()
        f[nr:] = f[a["map"][nr:]]
        out.append(f)
    if flip_velocity and a["boundary"] == 0:
        dim = a["dim"]
        for k in range(dim):
            xk = a["x"][k]
            sel = (a["type"] == 2) & ((xk < a["lo"][k]) | (xk > a["hi"][k]))
            out[1 + k][sel] *= -1
    return out


def lattice(n, dim, seed=0, boundary=0, jitter=0.1):
    return mesh_arrays(jittered_lattice(n, dim, seed, jitter=jitter), dim, boundary)
