"""phd_extract_geometry (the device half of Tess2d/Tess3d::extract_geometry,
tess.cpp:238-398, tess3.cpp:334-562) against the oracle's
orc_extract_geometry on the same host topology, bit for bit: ghost
positions, cell volume and centroid offset, face area, centroid and normal.

Cases: 2D and 3D, reflective and periodic, jittered lattices, uniform
random points, a dyadic Cartesian lattice (exact answers), the 128^3
benchmark mesh, the Mesh.build_geometry path of the integrators, and
invalid topology entries (reported, not read).
"""
import ctypes

import numpy as np
import pytest
import torch

from parity import assert_bitwise

pytestmark = pytest.mark.gpu


def _device_geometry(dim, x, topo, lo, hi, boundary):
    from phd.mesh import extract_geometry, upload_topology
    dev = torch.device("cuda:0")
    n, ng, nf = x[0].size, topo["ghost_map"].size, topo["pair_i"].size
    t = upload_topology(topo, dev)
    X = [torch.zeros(n + ng, dtype=torch.float64, device=dev) for _ in range(dim)]
    for a in range(dim):
        X[a][:n] = torch.as_tensor(x[a], device=dev)
    vol = torch.zeros(n + ng, dtype=torch.float64, device=dev)
    dcom = [torch.zeros(n + ng, dtype=torch.float64, device=dev) for _ in range(dim)]
    area = torch.zeros(nf, dtype=torch.float64, device=dev)
    com = [torch.zeros(nf, dtype=torch.float64, device=dev) for _ in range(dim)]
    normal = [torch.zeros(nf, dtype=torch.float64, device=dev) for _ in range(dim)]
    extract_geometry(t, dim, lo, hi, boundary, X, vol, dcom, area=area, com=com, normal=normal)
    torch.cuda.synchronize()

    def h(v):
        return [u.cpu().numpy() for u in v] if isinstance(v, list) else v.cpu().numpy()
    return {"x": h(X), "volume": h(vol), "dcom": h(dcom), "area": h(area), "com": h(com),
            "normal": h(normal)}


def _compare(dim, x, boundary, name, lo=None, hi=None, omp=False):
    from oracle import oracle
    from phd.mesh import build_voronoi
    lo = [0.0] * dim if lo is None else lo
    hi = [1.0] * dim if hi is None else hi
    topo = build_voronoi(dim, x, lo, hi, boundary, geometry=False)
    print("case %s: %d cells, %d ghosts, %d faces, %d ring entries"
          % (name, x[0].size, topo["ghost_map"].size, topo["pair_i"].size,
             topo["ring_k"].size), flush=True)
    got = _device_geometry(dim, x, topo, lo, hi, boundary)
    want = oracle.extract_geometry(dim, x, topo, lo, hi, boundary, omp=omp)
    for k in ("volume", "area"):
        assert_bitwise(got[k], want[k], "%s %s" % (name, k))
    for k in ("x", "dcom", "com", "normal"):
        for a in range(dim):
            assert_bitwise(got[k][a], want[k][a], "%s %s[%d]" % (name, k, a))
    return topo, got


@pytest.mark.parametrize("dim,boundary", [(2, 0), (2, 1), (3, 0), (3, 1)])
def test_jittered_matches_oracle(gpu, dim, boundary):
    from phd.problems import jittered_lattice
    n = 40 if dim == 2 else 12
    _compare(dim, jittered_lattice(n, dim, seed=3), boundary,
             "jittered %dD %s" % (dim, ("reflective", "periodic")[boundary]))


@pytest.mark.parametrize("dim,boundary", [(2, 0), (2, 1), (3, 0), (3, 1)])
def test_random_points_offset_box_match_oracle(gpu, dim, boundary):
    rng = np.random.default_rng(17 + dim + boundary)
    lo, hi = [-0.5, 0.25, 1.0][:dim], [1.5, 0.75, 1.625][:dim]
    n = 1500 if dim == 2 else 900
    x = [lo[a] + (hi[a] - lo[a]) * rng.uniform(0.001, 0.999, size=n) for a in range(dim)]
    _compare(dim, x, boundary, "random %dD %s" % (dim, ("reflective", "periodic")[boundary]),
             lo, hi)


@pytest.mark.parametrize("dim", [2, 3])
def test_dyadic_lattice_exact(gpu, dim):
    n = 8
    h = 1.0 / n
    idx = np.indices((n,) * dim).reshape(dim, -1)
    x = [(idx[a] + 0.5) * h for a in range(dim)]
    topo, got = _compare(dim, x, 1, "lattice %dD" % dim)
    assert np.all(got["volume"] == h ** dim) and np.all(got["area"] == h ** (dim - 1))
    for a in range(dim):
        assert np.all(got["dcom"][a] == 0.0)


@pytest.mark.parametrize("cap", ["0", "97"])
def test_unstaged_faces_match_oracle(gpu, monkeypatch, cap):
    """The 3D face kernel stages each block's ring entries in LDS and falls
    back to reading (and recomputing) them from global memory for entries
    past its LDS capacity; PHD_GEOMETRY_STAGE_CAP lowers that capacity so
    every face (0) or most faces (97 entries per block) take the fallback."""
    from phd.problems import jittered_lattice
    monkeypatch.setenv("PHD_GEOMETRY_STAGE_CAP", cap)
    _compare(3, jittered_lattice(10, 3, seed=8), 1, "jittered 3D periodic, stage cap %s" % cap)


def test_benchmark_mesh_matches_oracle(gpu):
    """The 3D 128^3 jittered lattice of the Sedov benchmark (2.1 M cells,
    16.6 M faces); the oracle runs its OpenMP build (same bits)."""
    from phd.problems import jittered_lattice
    _compare(3, jittered_lattice(128, 3, seed=0), 0, "sedov mesh 128^3", omp=True)


def test_build_geometry_device_and_host_paths(gpu):
    """Mesh.build_geometry (mesh.pyx:255-323) on the device path gives the
    oracle's geometry from the same topology, and agrees with the host
    builder's own geometry (an independent computation) to rounding."""
    from oracle import oracle
    from phd import problems
    from phd.mesh import build_voronoi
    p = problems.kh_2d(n=48, seed=1)
    pc = p.particles
    assert p.mesh.device_geometry
    nr = p.n_real
    x = [pc[f][:nr].cpu().numpy() for f in pc.carray_named_groups["position"]]
    dm = p.domain_manager
    boundary = dm.boundary_condition.native_code
    topo = build_voronoi(2, x, dm.xmin, dm.xmax, boundary, geometry=False)
    want = oracle.extract_geometry(2, x, topo, dm.xmin, dm.xmax, boundary)
    fc = p.mesh.faces
    assert_bitwise(pc["volume"].cpu().numpy(), want["volume"], "volume")
    assert_bitwise(fc["area"].cpu().numpy(), want["area"], "area")
    for a, f in enumerate(pc.carray_named_groups["position"]):
        assert_bitwise(pc[f].cpu().numpy(), want["x"][a], "position[%d]" % a)
    for a, f in enumerate(fc.carray_named_groups["com"]):
        assert_bitwise(fc[f].cpu().numpy(), want["com"][a], "com[%d]" % a)
    host = build_voronoi(2, x, dm.xmin, dm.xmax, boundary)
    np.testing.assert_allclose(pc["volume"].cpu().numpy(), host["volume"], rtol=1e-12)
    np.testing.assert_allclose(fc["area"].cpu().numpy(), host["area"], rtol=1e-11,
                               atol=1e-15)


@pytest.mark.parametrize("dim,field,value,expect", [(2, "ring_k", 10 ** 6, "face"),
                                                    (2, "pair_j", -3, "face"),
                                                    (2, "nbr_face", 10 ** 7, "cell"),
                                                    (2, "ghost_map", 10 ** 6, "cell"),
                                                    (3, "ring_k", 10 ** 6, "face"),
                                                    (3, "ring_code", 100, "face"),
                                                    (3, "pair_i", 10 ** 6, "face")])
def test_invalid_topology_is_reported(gpu, dim, field, value, expect):
    """Every index is range-checked on the device before it is used: an
    out-of-range entry gives PHD_E_ARG and names the face or cell."""
    from phd.mesh import build_voronoi
    from phd.problems import jittered_lattice
    x = jittered_lattice(10 if dim == 2 else 5, dim, seed=1)
    box = [0.0] * dim, [1.0] * dim
    topo = build_voronoi(dim, x, box[0], box[1], 0, geometry=False)
    topo[field] = topo[field].copy()
    pos = topo[field].size // 2
    topo[field][pos] = value
    print("case invalid %dD %s[%d] = %d" % (dim, field, pos, value), flush=True)
    with pytest.raises(RuntimeError, match="invalid topology entry \\(%s" % expect):
        _device_geometry(dim, x, topo, box[0], box[1], 0)
    from phd import _native as nat
    nat.raise_device_fault(nat.handle())   # the step error word is untouched
