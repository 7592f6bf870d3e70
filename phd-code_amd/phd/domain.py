"""Domain manager and boundary conditions (serial, one GPU).

DomainManager follows domain/domain_manager.pyx:47-170 (constructor, field
registration, partition) and its serial ghost refresh
(update_ghost_fields 712-774, update_ghost_gradients 776-837,
move_generators 566-590).  Reflective / Periodic follow domain/boundary.pyx
(migrate_particles 236-263 / 502-538, update_gradients 265-307 / 457-478,
update_fields 309-348 / 480-500).  Ghost creation itself happens in the host
mesh builder (Mesh.build_geometry).
"""
import ctypes

import torch

from . import _native as nat
from .utils import ParticleTAGS
from .views import cells_view


class BoundaryConditionBase(object):
    native_code = None

    def migrate_particles(self, particles, domain_manager):
        raise NotImplementedError("BoundaryBase::migrate_particles called!")

    def update_gradients(self, particles, gradients, domain_manager):
        raise NotImplementedError("BoundaryBase::update_gradients called!")

    def update_fields(self, particles, domain_manager):
        raise NotImplementedError("BoundaryBase::update_fields called!")


def _outside_masks(particles, domain_manager):
    dim = len(particles.carray_named_groups["position"])
    ext = particles["type"] == ParticleTAGS.Exterior
    lo = [(particles[f] < domain_manager.xmin[k]) & ext
          for k, f in enumerate(particles.carray_named_groups["position"])]
    hi = [(particles[f] > domain_manager.xmax[k]) & ext
          for k, f in enumerate(particles.carray_named_groups["position"])]
    return dim, lo, hi


class Reflective(BoundaryConditionBase):
    native_code = nat.BOUNDARY_REFLECTIVE

    def migrate_particles(self, particles, domain_manager):
        real = particles["tag"] == ParticleTAGS.Real
        inside = torch.ones_like(real)
        for k, f in enumerate(particles.carray_named_groups["position"]):
            x = particles[f]
            inside &= (x >= domain_manager.xmin[k]) & (x <= domain_manager.xmax[k])
        if bool((real & ~inside).any()):
            raise RuntimeError("particle left domain in reflective boundary condition!!")

    def update_gradients(self, particles, gradients, domain_manager):
        """Diagonal-only flip of the ghost velocity gradient
        (boundary.pyx:296-307), on the GPU."""
        domain_manager._ghost_gradients(particles, gradients, nat.BOUNDARY_REFLECTIVE)

    def update_fields(self, particles, domain_manager):
        """Negate velocity and momentum of exterior ghosts on each axis where
        they lie outside the box (boundary.pyx:336-348)."""
        dim, lo, hi = _outside_masks(particles, domain_manager)
        for k in range(dim):
            v = particles[particles.carray_named_groups["velocity"][k]]
            mv = particles[particles.carray_named_groups["momentum"][k]]
            for mask in (lo[k], hi[k]):
                v[mask] *= -1
                mv[mask] *= -1


class Periodic(BoundaryConditionBase):
    native_code = nat.BOUNDARY_PERIODIC

    def migrate_particles(self, particles, domain_manager):
        """Wrap real particles that left the box (boundary.pyx:523-538)."""
        real = particles["tag"] == ParticleTAGS.Real
        pos = [particles[f] for f in particles.carray_named_groups["position"]]
        outside = torch.zeros_like(real)
        for k, x in enumerate(pos):
            outside |= (x < domain_manager.xmin[k]) | (x > domain_manager.xmax[k])
        sel = real & outside
        for k, x in enumerate(pos):
            L = domain_manager.xmax[k] - domain_manager.xmin[k]
            x[sel & (x <= domain_manager.xmin[k])] += L
            x[sel & (x >= domain_manager.xmax[k])] -= L

    def update_gradients(self, particles, gradients, domain_manager):
        domain_manager._ghost_gradients(particles, gradients, nat.BOUNDARY_PERIODIC)

    def update_fields(self, particles, domain_manager):
        pass


class DomainManager(object):
    def __init__(self, xmin, xmax, initial_radius=0.1, search_radius_factor=2.0):
        xmin, xmax = list(xmin), list(xmax)
        if len(xmin) not in [2, 3]:
            raise RuntimeError("Wrong dimension specified")
        if len(xmin) != len(xmax):
            raise RuntimeError("Wrong dimensions for xmin and xmax")
        for a, b in zip(xmin, xmax):
            if b < a:
                raise RuntimeError("ERROR: Boundary inconsistent")
        self.dim = len(xmin)
        self.xmin = [float(v) for v in xmin]
        self.xmax = [float(v) for v in xmax]
        self.translate = [b - a for a, b in zip(self.xmin, self.xmax)]
        self.initial_radius = initial_radius
        self.search_radius_factor = search_radius_factor
        self.boundary_condition = None
        self.load_balance = None
        self.particle_fields_registered = False

    def register_fields(self, particles):
        if self.dim != len(particles.carray_named_groups["position"]):
            raise RuntimeError("ERROR: Particle dim does not match")
        n = particles.get_carray_size()
        for field, dtype in (("map", "long"), ("radius", "double"), ("old_radius", "double")):
            if field not in particles.carrays:
                particles.register_carray(n, field, dtype)
        particles["radius"][:] = self.initial_radius
        particles["old_radius"][:] = self.initial_radius
        self.particle_fields_registered = True

    def initialize(self):
        if not self.particle_fields_registered:
            raise RuntimeError("ERROR: Fields not registered in particles by Mesh!")
        if not self.boundary_condition:
            raise RuntimeError("Not all setters defined in DomainMangaer")

    def set_boundary_condition(self, boundary_condition):
        self.boundary_condition = boundary_condition

    def set_load_balance(self, load_balance):
        self.load_balance = load_balance

    def check_for_partition(self, particles, integrator):
        return True

    def partition(self, particles):
        self.boundary_condition.migrate_particles(particles, self)

    def move_generators(self, particles, dt):
        """x += dt*w for real cells (domain_manager.pyx:566-590)."""
        real = particles["tag"] == ParticleTAGS.Real
        for f, w in zip(particles.carray_named_groups["position"],
                        particles.carray_named_groups["w"]):
            x = particles[f]
            x[real] += dt * particles[w][real]

    def update_ghost_fields(self, particles, fields, apply_boundary_condition=False):
        """Copy fields from each ghost's image (domain_manager.pyx:759-774)."""
        ghost = torch.nonzero(particles["tag"] == ParticleTAGS.Ghost).flatten()
        if ghost.numel():
            src = particles["map"][ghost].long()
            for field in fields:
                a = particles[field]
                a[ghost] = a[src]
        if apply_boundary_condition:
            self.boundary_condition.update_fields(particles, self)

    def update_ghost_gradients(self, particles, gradients, mesh=None):
        """Ghost gradients from their images plus the boundary rule
        (domain_manager.pyx:819-837)."""
        if mesh is not None:
            self._mesh = mesh
        self.boundary_condition.update_gradients(particles, gradients, self)

    def _ghost_gradients(self, particles, gradients, boundary):
        lib, h = nat.flux(), nat.handle()
        if getattr(self, "_mesh", None) is None:
            raise RuntimeError("update_ghost_gradients needs the mesh whose plan holds the "
                               "cell centroids (pass mesh=)")
        plan = self._mesh.plan(particles)
        c = cells_view(particles)
        lo = (ctypes.c_double * 3)(*(self.xmin + [0.0] * (3 - self.dim)))
        hi = (ctypes.c_double * 3)(*(self.xmax + [0.0] * (3 - self.dim)))
        nat.check(lib.phd_update_ghost_gradients(h, plan, ctypes.byref(c), boundary,
                                                 ctypes.cast(lo, ctypes.c_void_p),
                                                 ctypes.cast(hi, ctypes.c_void_p),
                                                 gradients.data.data_ptr(), gradients.ld),
                  "update_ghost_gradients")
