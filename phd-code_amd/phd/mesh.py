"""Voronoi mesh: host topology + device geometry, GPU velocities and
conserved update.

Mesh keeps the reference's constructor, field registration and face schema
(mesh/mesh.pyx:16-166).  build_geometry replaces the CGAL tessellation and
ghost creation (mesh.pyx:168-323): the host builder libphd_mesh.so finds the
Delaunay topology (faces, ghost images, each face's ring of bounding
generators, per-cell face lists), which is uploaded, and
phd_extract_geometry computes ghost positions, face area, centroid and
normal and cell volume and centroid on the device (tess.cpp:238-398,
tess3.cpp:334-562).  device_geometry=False keeps the host builder's own
geometry instead.  The three hot methods (assign_generator_velocities,
assign_face_velocities, update_from_fluxes; mesh.pyx:363-551) are HIP
kernels.
"""
import ctypes

import numpy as np
import torch

from . import _native as nat
from .containers import CarrayContainer
from .utils import ParticleTAGS
from .views import cells_view, faces_view

face_vars_2d = {"area": "double", "pair-i": "long", "pair-j": "long",
                "com-x": "double", "com-y": "double",
                "velocity-x": "double", "velocity-y": "double",
                "normal-x": "double", "normal-y": "double"}
named_group_2d = {"com": ["com-x", "com-y"], "velocity": ["velocity-x", "velocity-y"],
                  "normal": ["normal-x", "normal-y"]}
face_vars_3d = dict(face_vars_2d, **{"com-z": "double", "normal-z": "double",
                                     "velocity-z": "double"})
named_group_3d = {"com": ["com-x", "com-y", "com-z"],
                  "velocity": ["velocity-x", "velocity-y", "velocity-z"],
                  "normal": ["normal-x", "normal-y", "normal-z"]}
fields_to_register_2d = {"volume": "double", "dcom-x": "double", "dcom-y": "double",
                         "w-x": "double", "w-y": "double"}
fields_to_register_3d = dict(fields_to_register_2d, **{"dcom-z": "double", "w-z": "double"})


class FaceLists(object):
    """Per-cell face lists (the reference's mesh.neighbors, mesh.pxd:34) as
    CSR: faces of cell i are faces[offsets[i]:offsets[i+1]], ascending."""

    def __init__(self, offsets, faces):
        self.offsets = offsets
        self.faces = faces

    def __len__(self):
        return self.offsets.numel() - 1

    def __getitem__(self, i):
        o = self.offsets
        return self.faces[int(o[i]):int(o[i + 1])].tolist()

    def size(self):
        return len(self)


def face_lists_from_pairs(pair_i, pair_j, n_cells):
    """Build ascending per-cell face lists from the pair arrays, the way
    extract_geometry pushes face ids (tess.cpp:379-382)."""
    pi = np.asarray(pair_i, dtype=np.int64)
    pj = np.asarray(pair_j, dtype=np.int64)
    nf = pi.size
    cells = np.concatenate([pi, pj])
    fids = np.concatenate([np.arange(nf), np.arange(nf)])
    order = np.lexsort((fids, cells))
    counts = np.bincount(cells, minlength=n_cells)
    off = np.zeros(n_cells + 1, dtype=np.int32)
    np.cumsum(counts, out=off[1:])
    return off, fids[order].astype(np.int32)


class Mesh(object):
    def __init__(self, regularize=True, relax_iterations=0, eta=0.25, num_neighbors=128,
                 max_iterations=20, device_geometry=True):
        self.particle_fields_registered = False
        self.max_iterations = max_iterations
        self.relax_iterations = relax_iterations
        self.eta = float(eta)
        self.regularize = bool(regularize)
        self.num_neighbors = num_neighbors
        self.faces = None
        self.neighbors = None
        self.n_threads = 0
        self._plan = None
        self._plan_key = None
        self.version = 0
        self.device_geometry = bool(device_geometry)
        self.topology = None

    def register_fields(self, particles):
        dim = len(particles.carray_named_groups["position"])
        axes = "xyz"[:dim]
        particles.carray_named_groups["w"] = ["w-" + a for a in axes]
        particles.carray_named_groups["dcom"] = ["dcom-" + a for a in axes]
        if dim == 2:
            self.face_fields, self.face_field_groups = face_vars_2d, named_group_2d
            reg = fields_to_register_2d
        elif dim == 3:
            self.face_fields, self.face_field_groups = face_vars_3d, named_group_3d
            reg = fields_to_register_3d
        else:
            raise RuntimeError("Wrong dimension specified")
        n = particles.get_carray_size()
        for field, dtype in reg.items():
            if field not in particles.carrays:
                particles.register_carray(n, field, dtype)
        self.update_ghost_fields = list(particles.carray_named_groups["dcom"]) + ["volume"]
        self.particle_fields_registered = True

    def initialize(self, device=None):
        if not self.particle_fields_registered:
            raise RuntimeError("ERROR: Fields not registered in particles by Mesh!")
        self.faces = CarrayContainer(carrays_to_register=self.face_fields, device=device)
        self.faces.carray_named_groups = {k: list(v) for k, v in self.face_field_groups.items()}
        dev = self.faces.device
        self.neighbors = FaceLists(torch.zeros(1, dtype=torch.int32, device=dev),
                                   torch.zeros(0, dtype=torch.int32, device=dev))

    # -- geometry (host) -----------------------------------------------------
    def tessellate(self, particles, domain_manager):
        self.build_geometry(particles, domain_manager)

    def reset_mesh(self):
        pass

    def build_geometry(self, particles, domain_manager):
        """Remove old ghosts, build the Voronoi mesh of the real generators
        with ghost images (reflective mirrors / periodic copies), append the
        ghosts as copies of their images (domain_manager.pyx:516-539), and
        set faces, volumes, dcom and the per-cell face lists."""
        particles.remove_tagged_particles(ParticleTAGS.Ghost)
        dim = len(particles.carray_named_groups["position"])
        n = particles.get_carray_size()
        xs = [np.ascontiguousarray(particles[f].cpu().numpy(), dtype=np.float64)
              for f in particles.carray_named_groups["position"]]
        boundary = domain_manager.boundary_condition.native_code
        lo, hi = domain_manager.xmin, domain_manager.xmax
        geo = build_voronoi(dim, xs, lo, hi, boundary, self.n_threads,
                            geometry=not self.device_geometry)
        ng = geo["ghost_map"].size
        dev = particles.device
        if ng:
            gmap = torch.as_tensor(geo["ghost_map"], device=dev, dtype=torch.long)
            ghosts = particles.extract_items(gmap)
            if not self.device_geometry:
                for k, f in enumerate(particles.carray_named_groups["position"]):
                    ghosts[f][:] = torch.as_tensor(geo["ghost_x"][k], device=dev)
            ghosts["tag"][:] = ParticleTAGS.Ghost
            ghosts["type"][:] = ParticleTAGS.Exterior
            if "map" in ghosts.carrays:
                ghosts["map"][:] = gmap.to(torch.int32)
            particles.append_container(ghosts)
        if not self.device_geometry:
            particles["volume"][:] = torch.as_tensor(geo["volume"], device=dev)
            for k, f in enumerate(particles.carray_named_groups["dcom"]):
                particles[f][:] = torch.as_tensor(geo["dcom"][k], device=dev)
            self.set_faces(geo)
            return
        self.topology = upload_topology(geo, dev)
        self.faces.resize(geo["pair_i"].size)
        extract_geometry(self.topology, dim, lo, hi, boundary,
                         [particles[f] for f in particles.carray_named_groups["position"]],
                         particles["volume"],
                         [particles[f] for f in particles.carray_named_groups["dcom"]],
                         self.faces)
        self.set_faces(geo, device_geometry=True)

    def set_faces(self, geo, device_geometry=False):
        """Upload face arrays and face lists (host numpy dict); with
        device_geometry the area, com and normal arrays are already on the
        device and only the pairs and face lists are taken from geo."""
        nf = geo["pair_i"].size
        fc = self.faces
        fc.resize(nf)
        dev = fc.device
        fc["pair-i"][:] = torch.as_tensor(geo["pair_i"], device=dev)
        fc["pair-j"][:] = torch.as_tensor(geo["pair_j"], device=dev)
        if not device_geometry:
            fc["area"][:] = torch.as_tensor(geo["area"], device=dev)
            for k, f in enumerate(fc.carray_named_groups["com"]):
                fc[f][:] = torch.as_tensor(geo["com"][k], device=dev)
            for k, f in enumerate(fc.carray_named_groups["normal"]):
                fc[f][:] = torch.as_tensor(geo["normal"][k], device=dev)
        for f in fc.carray_named_groups["velocity"]:
            fc[f][:] = 0.0
        self.neighbors = FaceLists(torch.as_tensor(geo["nbr_off"], device=dev),
                                   torch.as_tensor(geo["nbr_face"], device=dev))
        self.invalidate_plan()

    # -- device plan ---------------------------------------------------------
    def invalidate_plan(self):
        """Call after changing face arrays, positions, dcom or volumes."""
        self.version += 1
        if self._plan is not None:
            self._plan.close()
        self._plan = None
        self._plan_key = None

    def plan(self, particles):
        """The device plan of this mesh (built on first use after a rebuild)."""
        key = (self.version, particles.get_carray_size(), self.faces.get_carray_size(),
               particles.pointer("tag"), self.faces.pointer("area"))
        if self._plan is None or self._plan_key != key:
            if self._plan is not None:
                self._plan.close()
            self._plan = nat.Plan(cells_view(particles), faces_view(self))
            self._plan_key = key
        return self._plan.ptr

    # -- hot path (GPU) ------------------------------------------------------
    def assign_generator_velocities(self, particles, equation_state):
        """mesh.pyx:363-426: w = v + Springel (2010) eq. 63 regularisation."""
        lib, h = nat.flux(), nat.handle()
        c = cells_view(particles)
        nat.check(lib.phd_assign_generator_velocities(
            h, ctypes.byref(c), equation_state.get_gamma(), self.eta, int(self.regularize)),
            "assign_generator_velocities")

    def assign_face_velocities(self, particles):
        """mesh.pyx:428-477: face velocity, Springel (2010) eq. 32-33."""
        lib, h = nat.flux(), nat.handle()
        c, f = cells_view(particles), faces_view(self)
        nat.check(lib.phd_assign_face_velocities(h, ctypes.byref(c), ctypes.byref(f)),
                  "assign_face_velocities")

    def update_from_fluxes(self, particles, riemann, dt, check_errors=True):
        """mesh.pyx:479-551: deterministic per-cell walk over ascending face
        ids; raises the reference's RuntimeError on non-positive mass or
        energy."""
        lib, h = nat.flux(), nat.handle()
        c, f = cells_view(particles), faces_view(self)
        fl = nat.ptr_array(riemann.fluxes.group("conservative"))
        nat.check(lib.phd_update_from_fluxes(h, self.plan(particles), ctypes.byref(c),
                                             ctypes.byref(f), fl, float(dt)),
                  "update_from_fluxes")
        if check_errors:
            nat.raise_device_fault(h)


TOPOLOGY_KEYS = ("ghost_map", "ghost_code", "pair_i", "pair_j", "ring_off", "ring_k",
                 "ring_code", "nbr_off", "nbr_face")


def upload_topology(geo, device):
    """Device copies of the topology arrays of a build_voronoi result."""
    t = {k: torch.as_tensor(geo[k], device=device) for k in TOPOLOGY_KEYS}
    t["n_real"] = int(geo["n_real"])
    return t


def extract_geometry(topo, dim, lo, hi, boundary, x, volume, dcom, faces=None, area=None,
                     com=None, normal=None):
    """phd_extract_geometry on device tensors: topo from upload_topology;
    x, dcom: dim tensors of n_real + n_ghost (ghost rows of x are written);
    volume likewise; the face outputs either a face container (faces) or
    the area, com and normal tensors."""
    if faces is not None:
        area = faces["area"]
        com = [faces[f] for f in faces.carray_named_groups["com"]]
        normal = [faces[f] for f in faces.carray_named_groups["normal"]]
    lib, h = nat.flux(), nat.handle()
    nf = int(topo["pair_i"].numel())
    t = nat.Topology()
    t.dim, t.boundary = dim, int(boundary)
    t.n_real, t.n_ghost, t.n_faces = topo["n_real"], int(topo["ghost_map"].numel()), nf
    for a in range(dim):
        t.lo[a], t.hi[a] = float(lo[a]), float(hi[a])
    for k in TOPOLOGY_KEYS:
        v = topo[k]
        setattr(t, k, v.data_ptr() if v.numel() else None)

    def ptr(v):
        assert v.dtype == torch.float64 and v.is_contiguous()
        return v.data_ptr()

    g = nat.Geometry()
    for a in range(dim):
        g.x[a], g.dcom[a] = ptr(x[a]), ptr(dcom[a])
        if nf:
            g.com[a], g.normal[a] = ptr(com[a]), ptr(normal[a])
    g.volume = ptr(volume)
    g.area = ptr(area) if nf else None
    nbytes = int(lib.phd_geometry_workspace_bytes(dim, topo["n_real"], nf))
    ws = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device=volume.device)
    bad = ctypes.c_int64(-1)
    rc = lib.phd_extract_geometry(h, ctypes.byref(t), ctypes.byref(g), ws.data_ptr(), nbytes,
                                  ctypes.byref(bad))
    if rc == nat.PHD_E_ARG and bad.value != -1:
        raise RuntimeError("phd_extract_geometry: invalid topology entry (%s %d)" % (
            ("face", bad.value) if bad.value >= 0 else ("cell", -2 - bad.value)))
    nat.check(rc, "extract_geometry")


def build_voronoi(dim, xs, lo, hi, boundary, n_threads=0, geometry=True):
    """Run the host builder; returns a dict of numpy arrays: the topology
    (ghost_x, ghost_map, ghost_code, pair_i, pair_j, nbr_off, nbr_face and
    the face rings ring_off, ring_k, ring_code) and, with geometry=True, the
    host-computed geometry (volume, dcom, area, com, normal).  geometry=False
    runs the topology-only build, whose geometry phd_extract_geometry
    computes on the device."""
    lib = nat.mesh_lib()
    n = xs[0].size
    xs = [np.ascontiguousarray(x, dtype=np.float64) for x in xs]
    xp = (ctypes.c_void_p * 3)(*[x.ctypes.data for x in xs], *([0] * (3 - dim)))
    lo_ = (ctypes.c_double * 3)(*[float(v) for v in lo], *([0.0] * (3 - dim)))
    hi_ = (ctypes.c_double * 3)(*[float(v) for v in hi], *([0.0] * (3 - dim)))
    st = ctypes.c_int32(0)
    build = lib.phdm_build if geometry else lib.phdm_build_topology
    m = build(dim, n, ctypes.cast(xp, ctypes.c_void_p), ctypes.cast(lo_, ctypes.c_void_p),
              ctypes.cast(hi_, ctypes.c_void_p), int(boundary), int(n_threads), ctypes.byref(st))
    if not m:
        msg = {-1: "bad arguments", -2: "particle left domain (generator outside the box)",
               -3: "Mesh failed to converged!", -4: "too many faces"}.get(st.value, "error")
        raise RuntimeError("mesh build failed: %s" % msg)
    try:
        ng, nf, nl = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        lib.phdm_counts(m, ctypes.byref(ng), ctypes.byref(nf), ctypes.byref(nl))
        ng, nf, nl = ng.value, nf.value, nl.value
        nr_ = ctypes.c_int64()
        lib.phdm_ring_count(m, ctypes.byref(nr_))
        na = n + ng
        out = {
            "ghost_x": [np.zeros(ng) for _ in range(dim)],
            "ghost_map": np.zeros(ng, dtype=np.int32),
            "ghost_code": np.zeros(ng, dtype=np.int8),
            "pair_i": np.zeros(nf, dtype=np.int32),
            "pair_j": np.zeros(nf, dtype=np.int32),
            "nbr_off": np.zeros(na + 1, dtype=np.int32),
            "nbr_face": np.zeros(nl, dtype=np.int32),
            "ring_off": np.zeros(nf + 1, dtype=np.int32),
            "ring_k": np.zeros(nr_.value, dtype=np.int32),
            "ring_code": np.zeros(nr_.value, dtype=np.int8),
        }
        if geometry:
            out.update({"volume": np.zeros(na), "dcom": [np.zeros(na) for _ in range(dim)],
                        "area": np.zeros(nf), "com": [np.zeros(nf) for _ in range(dim)],
                        "normal": [np.zeros(nf) for _ in range(dim)]})
        keep = []

        def parr(key):
            if key not in out:
                return None
            a = (ctypes.c_void_p * 3)(*[x.ctypes.data for x in out[key]], *([0] * (3 - dim)))
            keep.append(a)
            return ctypes.cast(a, ctypes.c_void_p)

        def ptr(key):
            return out[key].ctypes.data if key in out else None

        rc = lib.phdm_fill(m, parr("ghost_x"), ptr("ghost_map"), ptr("volume"), parr("dcom"),
                           ptr("area"), parr("com"), parr("normal"), ptr("pair_i"),
                           ptr("pair_j"), ptr("nbr_off"), ptr("nbr_face"))
        if rc == 0:
            rc = lib.phdm_fill_topology(m, ptr("ghost_code"), ptr("ring_off"), ptr("ring_k"),
                                        ptr("ring_code"))
        if rc != 0:
            raise RuntimeError("mesh fill failed")
        out["n_real"] = n
        out["n_ghost"] = ng
        return out
    finally:
        lib.phdm_free(m)
