"""Pins of the oracle's geometry extraction (orc_extract_geometry), the
restatement of Tess2d/Tess3d::extract_geometry (tess.cpp:238-398,
tess3.cpp:334-562) from the host builder's Delaunay topology.

None of these retypes the oracle's formulas:
  - a Cartesian lattice with a dyadic spacing, where every quantity is exact
    (cell volume h^D, centroid offset 0, face area h^(D-1), face centroid the
    midpoint, normal an axis);
  - the reference's known answers (lattice volumes, total volume of a random
    box; tests/golden/reference_known_answers.json);
  - the divergence theorem on jittered meshes, which ties volume, centroid,
    face area, face centroid and normal together independently of how a
    cell is cut into triangles or tetrahedra: sum_f a_f n_f = 0,
    V = (1/D) sum_f a_f d_f and V dcom = D/(D+1) (1/D) sum_f a_f d_f
    (c_f - x), with d_f = (c_f - x).n_f (each face's cone from the
    generator has its centroid at D/(D+1) of the way to the face centroid);
  - the host builder's own geometry, an independent computation (the
    vertices of the clipped polytope, not circumcentres);
  - the reference's ghost image formulas (boundary.pyx:103-104, 378-380).
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle
from phd.mesh import build_voronoi
from phd.problems import jittered_lattice

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_known_answers.json")))


def _oracle_mesh(dim, x, boundary, lo=None, hi=None):
    lo = [0.0] * dim if lo is None else lo
    hi = [1.0] * dim if hi is None else hi
    t = build_voronoi(dim, x, lo, hi, boundary, geometry=False)
    return t, oracle.extract_geometry(dim, x, t, lo, hi, boundary)


def _lattice(n, dim):
    idx = np.indices((n,) * dim).reshape(dim, -1)
    return [(idx[a] + 0.5) / n for a in range(dim)]


@pytest.mark.parametrize("dim,boundary", [(2, 0), (2, 1), (3, 0), (3, 1)])
def test_dyadic_lattice_is_exact(dim, boundary):
    n = 8
    h = 1.0 / n
    x = _lattice(n, dim)
    t, o = _oracle_mesh(dim, x, boundary)
    nr = n ** dim
    assert np.all(o["volume"] == h ** dim)
    assert np.all(o["area"] == h ** (dim - 1))
    for a in range(dim):
        assert np.all(o["dcom"][a] == 0.0)
    X = o["x"]
    pi, pj = t["pair_i"], t["pair_j"]
    for a in range(dim):
        np.testing.assert_array_equal(o["com"][a], 0.5 * (X[a][pi] + X[a][pj]))
    nrm = np.stack(o["normal"], 1)
    assert np.all(np.abs(nrm).sum(1) == 1.0) and np.all(np.abs(nrm).max(1) == 1.0)
    # 2D (2n(n-1) + 4n) and 3D (3n^2(n-1) + 6n^2) face counts, reflective;
    # periodic has the same count (wall faces become image faces)
    assert pi.size == (2 * n * (n - 1) + 4 * n if dim == 2 else 3 * n * n * (n - 1) + 6 * n * n)
    assert o["volume"][:nr].sum() == 1.0


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_lattice_volumes(dim):
    n = GOLD["lattice_volumes"]["n_%dd" % dim]
    t, o = _oracle_mesh(dim, _lattice(n, dim), 0)
    nr = n ** dim
    np.testing.assert_allclose(o["volume"][:nr], (1.0 / n) ** dim, rtol=1e-12)
    for a in range(dim):
        assert np.abs(o["dcom"][a][:nr]).max() < 1e-14


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_random_box_total_volume(dim):
    n = GOLD["random_box_total_volume"]["n_%dd" % dim]
    rng = np.random.default_rng(0)
    x = [rng.uniform(size=n) for _ in range(dim)]
    t, o = _oracle_mesh(dim, x, 0)
    np.testing.assert_allclose(o["volume"][:n].sum(), 1.0, rtol=1e-12)


def _per_cell_identities(dim, t, o, nr):
    """closure, volume and centroid from the divergence theorem, per real
    cell, from face data only."""
    X = np.stack(o["x"], 1)
    pi, pj = t["pair_i"], t["pair_j"]
    area, com, nrm = o["area"], np.stack(o["com"], 1), np.stack(o["normal"], 1)
    closure = np.zeros((nr, dim))
    vol = np.zeros(nr)
    mom = np.zeros((nr, dim))
    for side, cells, sgn in ((0, pi, 1.0), (1, pj, -1.0)):
        keep = cells < nr
        c = cells[keep]
        n_out = sgn * nrm[keep]
        rel = com[keep] - X[c]
        d = (rel * n_out).sum(1)
        np.add.at(closure, c, area[keep, None] * n_out)
        np.add.at(vol, c, area[keep] * d / dim)
        np.add.at(mom, c, (area[keep] * d / dim)[:, None] * (dim / (dim + 1.0)) * rel)
    return closure, vol, mom / vol[:, None]


@pytest.mark.parametrize("dim,boundary,n", [(2, 0, 24), (2, 1, 24), (3, 0, 7), (3, 1, 7)])
def test_divergence_theorem_identities(dim, boundary, n):
    x = jittered_lattice(n, dim, seed=11)
    t, o = _oracle_mesh(dim, x, boundary)
    nr = n ** dim
    h = 1.0 / n
    closure, vol, dc = _per_cell_identities(dim, t, o, nr)
    assert np.abs(closure).max() < 1e-13 * h ** (dim - 1)
    np.testing.assert_allclose(o["volume"][:nr], vol, rtol=1e-12)
    for a in range(dim):
        assert np.abs(o["dcom"][a][:nr] - dc[:, a]).max() < 1e-13 * h
    np.testing.assert_allclose(o["volume"][:nr].sum(), 1.0, rtol=1e-13)
    # ghost cells carry their image's volume and centroid offset
    gm = t["ghost_map"]
    np.testing.assert_array_equal(o["volume"][nr:], o["volume"][gm])
    for a in range(dim):
        np.testing.assert_array_equal(o["dcom"][a][nr:], o["dcom"][a][gm])


@pytest.mark.parametrize("dim,boundary,n", [(2, 0, 30), (2, 1, 30), (3, 0, 8), (3, 1, 8)])
def test_matches_clipping_geometry(dim, boundary, n):
    """The host builder's geometry comes from the vertices of the clipped
    polytope (plane intersections along its edges); the oracle's from
    circumcentres of the Delaunay simplices.  Same cells, different
    arithmetic."""
    x = jittered_lattice(n, dim, seed=5)
    g = build_voronoi(dim, x, [0.0] * dim, [1.0] * dim, boundary)
    t, o = _oracle_mesh(dim, x, boundary)
    h = 1.0 / n
    np.testing.assert_allclose(o["volume"], g["volume"], rtol=1e-12)
    np.testing.assert_allclose(o["area"], g["area"], rtol=1e-11, atol=1e-14 * h ** (dim - 1))
    for a in range(dim):
        assert np.abs(o["com"][a] - g["com"][a]).max() < 1e-12 * h
        assert np.abs(o["dcom"][a] - g["dcom"][a]).max() < 1e-13 * h
        assert np.abs(o["normal"][a] - g["normal"][a]).max() < 1e-15


# an offset box with unequal sides, so that a mixed-up axis shows
BOX_LO = [-0.5, 0.25, 1.0]
BOX_HI = [1.5, 0.75, 1.625]


def _box_points(n, dim, seed):
    x = jittered_lattice(n, dim, seed=seed)
    return [BOX_LO[a] + (BOX_HI[a] - BOX_LO[a]) * x[a] for a in range(dim)]


@pytest.mark.parametrize("dim,boundary", [(2, 0), (2, 1), (3, 0), (3, 1)])
def test_ghost_images(dim, boundary):
    n = 16 if dim == 2 else 6
    lo, hi = BOX_LO[:dim], BOX_HI[:dim]
    x = _box_points(n, dim, 2)
    t, o = _oracle_mesh(dim, x, boundary, lo, hi)
    nr = n ** dim
    gm, gc = t["ghost_map"], t["ghost_code"].astype(int)
    assert gm.size and np.all(gc > 0)
    for a in range(dim):
        src, img = x[a][gm], o["x"][a][nr:]
        np.testing.assert_array_equal(img, t["ghost_x"][a])
        if boundary == 1:
            # boundary.pyx:378-380: x + (index - 1) translate, index the base-3 digit
            shift = ((gc - 1) // 3 ** a) % 3 - 1
            np.testing.assert_array_equal(img, src + shift * (hi[a] - lo[a]))
        else:
            # boundary.pyx:103-104, 125-126: x - 2 (x - edge) on the wall's axis
            axis, side = (gc - 1) >> 1, (gc - 1) & 1
            on = axis == a
            edge = np.where(side[on] == 1, hi[a], lo[a])
            np.testing.assert_array_equal(img[~on], src[~on])
            np.testing.assert_array_equal(img[on], src[on] - 2 * (src[on] - edge))


@pytest.mark.parametrize("dim,boundary", [(2, 0), (2, 1), (3, 0), (3, 1)])
def test_anisotropic_box_identities(dim, boundary):
    n = 20 if dim == 2 else 6
    lo, hi = BOX_LO[:dim], BOX_HI[:dim]
    t, o = _oracle_mesh(dim, _box_points(n, dim, 9), boundary, lo, hi)
    nr = n ** dim
    closure, vol, dc = _per_cell_identities(dim, t, o, nr)
    np.testing.assert_allclose(o["volume"][:nr], vol, rtol=1e-12)
    np.testing.assert_allclose(o["volume"][:nr].sum(), np.prod(np.subtract(hi, lo)), rtol=1e-13)
    scale = min(np.subtract(hi, lo)) / n
    for a in range(dim):
        assert np.abs(o["dcom"][a][:nr] - dc[:, a]).max() < 1e-12 * scale


def test_reference_reflective_ghost_positions():
    g0 = GOLD["reflective_ghost_positions"]
    x = [np.array([g0["particle"][0], 0.75, 0.75, 0.25]),
         np.array([g0["particle"][1], 0.25, 0.75, 0.75])]
    t, o = _oracle_mesh(2, x, 0)
    imgs = {tuple(np.round([o["x"][0][4 + k], o["x"][1][4 + k]], 12))
            for k in range(t["ghost_map"].size) if t["ghost_map"][k] == 0}
    for im in g0["images"]:
        assert tuple(im) in imgs


@pytest.mark.skipif(bool(os.environ.get("PHD_ORACLE_LIB")),
                    reason="compares the two builds of the unmodified source")
def test_serial_and_parallel_builds_agree():
    x = jittered_lattice(7, 3, seed=4)
    t = build_voronoi(3, x, [0.0] * 3, [1.0] * 3, 1, geometry=False)
    a = oracle.extract_geometry(3, x, t, [0.0] * 3, [1.0] * 3, 1)
    b = oracle.extract_geometry(3, x, t, [0.0] * 3, [1.0] * 3, 1, omp=True)
    for k in ("volume", "area"):
        np.testing.assert_array_equal(a[k], b[k])
    for k in ("dcom", "com", "normal", "x"):
        for u, v in zip(a[k], b[k]):
            np.testing.assert_array_equal(u, v)
