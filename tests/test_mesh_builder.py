"""Host Voronoi builder (libphd_mesh.so): the reference's mesh fixtures and
the extraction conventions the flux pipeline relies on.

Face-level geometry of the reference comes from CGAL, which is not part of
this build; its conventions (not its bits) are checked here, and scipy's
Qhull Voronoi is an independent cross-check of areas and neighbour sets.
"""
import json
import os

import numpy as np
import pytest

from phd.mesh import build_voronoi, face_lists_from_pairs
from phd.problems import jittered_lattice

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_known_answers.json")))


@pytest.mark.parametrize("dim", [2, 3])
def test_lattice_volumes(dim):
    n = GOLD["lattice_volumes"]["n_%dd" % dim]
    idx = np.indices((n,) * dim).reshape(dim, -1)
    x = [(idx[a] + 0.5) / n for a in range(dim)]
    g = build_voronoi(dim, x, [0.0] * dim, [1.0] * dim, 0)
    nr = g["n_real"]
    np.testing.assert_allclose(g["volume"][:nr], (1.0 / n) ** dim, rtol=1e-12)
    for k in range(dim):
        assert np.abs(g["dcom"][k][:nr]).max() < 1e-14
    np.testing.assert_allclose(g["volume"][:nr].sum(), 1.0, rtol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_random_box_total_volume(dim):
    n = GOLD["random_box_total_volume"]["n_%dd" % dim]
    rng = np.random.default_rng(0)
    x = [rng.uniform(size=n) for _ in range(dim)]
    g = build_voronoi(dim, x, [0.0] * dim, [1.0] * dim, 0)
    np.testing.assert_allclose(g["volume"][:n].sum(), 1.0, rtol=1e-12)


def test_reflective_ghost_images():
    g0 = GOLD["reflective_ghost_positions"]
    x = [np.array([g0["particle"][0], 0.75, 0.75, 0.25]), np.array([g0["particle"][1], 0.25, 0.75, 0.75])]
    g = build_voronoi(2, x, [0.0, 0.0], [1.0, 1.0], 0)
    imgs = {tuple(np.round([g["ghost_x"][0][k], g["ghost_x"][1][k]], 12))
            for k in range(g["n_ghost"]) if g["ghost_map"][k] == 0}
    for im in g0["images"]:
        assert tuple(im) in imgs


@pytest.mark.parametrize("dim,boundary", [(2, 0), (2, 1), (3, 0), (3, 1)])
def test_conventions(dim, boundary):
    n = 24 if dim == 2 else 8
    x = jittered_lattice(n, dim, seed=1)
    g = build_voronoi(dim, x, [0.0] * dim, [1.0] * dim, boundary)
    nr, ng = g["n_real"], g["n_ghost"]
    pi, pj = g["pair_i"], g["pair_j"]
    # faces are emitted from their lower-id real cell, in ascending pair_i
    assert np.all(pi < pj) and np.all(pi < nr)
    assert np.all(np.diff(pi) >= 0)
    # normal = (x_j - x_i)/|x_j - x_i|
    X = [np.concatenate([x[k], g["ghost_x"][k]]) for k in range(dim)]
    d = np.stack([X[k][pj] - X[k][pi] for k in range(dim)], 1)
    nrm = np.stack(g["normal"], 1)
    np.testing.assert_allclose(nrm, d / np.linalg.norm(d, axis=1)[:, None], rtol=0, atol=1e-14)
    # face centroid lies on the bisector plane
    com = np.stack(g["com"], 1)
    mid = 0.5 * np.stack([X[k][pj] + X[k][pi] for k in range(dim)], 1)
    assert np.abs(((com - mid) * nrm).sum(1)).max() < 1e-12
    # closed cells: sum over faces of area * outward normal = 0
    closure = np.zeros((nr, dim))
    np.add.at(closure, pi, g["area"][:, None] * nrm)
    real_j = pj < nr
    np.add.at(closure, pj[real_j], -(g["area"][real_j, None] * nrm[real_j]))
    assert np.abs(closure).max() < 1e-12
    # per-cell lists: ascending face ids, exactly the faces touching the cell
    off, lst = face_lists_from_pairs(pi, pj, nr + ng)
    np.testing.assert_array_equal(off, g["nbr_off"])
    np.testing.assert_array_equal(lst, g["nbr_face"])
    for c in range(0, nr, max(1, nr // 50)):
        seg = g["nbr_face"][off[c]:off[c + 1]]
        assert np.all(np.diff(seg) > 0)
    # ghosts copy their image's volume and dcom, unreflected (mesh.pyx:141-142)
    gm = g["ghost_map"]
    np.testing.assert_array_equal(g["volume"][nr:], g["volume"][gm])
    for k in range(dim):
        np.testing.assert_array_equal(g["dcom"][k][nr:], g["dcom"][k][gm])
    np.testing.assert_allclose(g["volume"][:nr].sum(), 1.0, rtol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_against_qhull(dim):
    """Neighbour sets and face areas of interior cells against scipy's
    Qhull Voronoi of the same generators plus their mirror images."""
    scipy_spatial = pytest.importorskip("scipy.spatial")
    from scipy.spatial import ConvexHull, Voronoi
    rng = np.random.default_rng(3)
    n = 300 if dim == 2 else 400
    x = [rng.uniform(0.05, 0.95, n) for _ in range(dim)]
    g = build_voronoi(dim, x, [0.0] * dim, [1.0] * dim, 0)
    P = np.stack(x, 1)
    pts = [P]
    for a in range(dim):
        for w in (0.0, 1.0):
            Q = P.copy()
            Q[:, a] = 2 * w - Q[:, a]
            pts.append(Q)
    vor = Voronoi(np.concatenate(pts))
    ours = {}
    for f in range(g["area"].size):
        i, j = g["pair_i"][f], g["pair_j"][f]
        if j < n:
            ours[(i, j)] = g["area"][f]
    checked = 0
    for (a, b), rv in zip(vor.ridge_points, vor.ridge_vertices):
        if a >= n or b >= n or -1 in rv:
            continue
        i, j = min(a, b), max(a, b)
        V = vor.vertices[rv]
        if dim == 2:
            area = np.linalg.norm(V[1] - V[0])
        else:
            c = V.mean(0)
            nrm = P[j] - P[i]
            nrm /= np.linalg.norm(nrm)
            u = np.cross(nrm, [1.0, 0, 0] if abs(nrm[0]) < 0.9 else [0, 1.0, 0])
            u /= np.linalg.norm(u)
            v = np.cross(nrm, u)
            area = ConvexHull(np.stack([(V - c) @ u, (V - c) @ v], 1)).volume
        if area < 1e-9:
            continue
        assert (i, j) in ours, (i, j, area)
        np.testing.assert_allclose(ours[(i, j)], area, rtol=1e-9)
        checked += 1
    assert checked > n
