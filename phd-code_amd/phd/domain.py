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

from . import _native as nat
from .views import cells_view


class BoundaryConditionBase(object):
    native_code = None

    def migrate_particles(self, particles, domain_manager):
        raise NotImplementedError("BoundaryBase::migrate_particles called!")

    def update_gradients(self, particles, gradients, domain_manager):
        raise NotImplementedError("BoundaryBase::update_gradients called!")

    def update_fields(self, particles, domain_manager):
        raise NotImplementedError("BoundaryBase::update_fields called!")


def _box(domain_manager):
    lo = (ctypes.c_double * 3)(*(domain_manager.xmin + [0.0] * (3 - domain_manager.dim)))
    hi = (ctypes.c_double * 3)(*(domain_manager.xmax + [0.0] * (3 - domain_manager.dim)))
    return lo, hi


class Reflective(BoundaryConditionBase):
    native_code = nat.BOUNDARY_REFLECTIVE

    def migrate_particles(self, particles, domain_manager):
        """boundary.pyx:236-263: raise if a real particle left the box (the
        check runs on the GPU and synchronises)."""
        lib, h = nat.flux(), nat.handle()
        c = cells_view(particles)
        lo, hi = _box(domain_manager)
        first = ctypes.c_int64(-1)
        nat.check(lib.phd_first_outside(h, ctypes.byref(c), lo, hi, ctypes.byref(first)),
                  "migrate_particles")
        if first.value >= 0:
            raise RuntimeError("particle left domain in reflective boundary condition!!")

    def update_gradients(self, particles, gradients, domain_manager):
        """Diagonal-only flip of the ghost velocity gradient
        (boundary.pyx:296-307), on the GPU."""
        domain_manager._ghost_gradients(particles, gradients, nat.BOUNDARY_REFLECTIVE)

    def update_fields(self, particles, domain_manager):
        """Negate velocity and momentum of exterior ghosts on each axis where
        they lie outside the box (boundary.pyx:336-348), on the GPU."""
        lib, h = nat.flux(), nat.handle()
        c = cells_view(particles)
        lo, hi = _box(domain_manager)
        nat.check(lib.phd_reflect_fields(h, ctypes.byref(c), lo, hi), "update_fields")


class Periodic(BoundaryConditionBase):
    native_code = nat.BOUNDARY_PERIODIC

    def migrate_particles(self, particles, domain_manager):
        """Wrap real particles that left the box (boundary.pyx:523-538), on
        the GPU; both tests of an axis see the position before its wrap."""
        lib, h = nat.flux(), nat.handle()
        c = cells_view(particles)
        lo, hi = _box(domain_manager)
        nat.check(lib.phd_migrate_periodic(h, ctypes.byref(c), lo, hi), "migrate_particles")

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
        """x += dt*w for real cells (domain_manager.pyx:566-590), on the GPU."""
        lib, h = nat.flux(), nat.handle()
        c = cells_view(particles)
        nat.check(lib.phd_move_generators(h, ctypes.byref(c), float(dt)), "move_generators")

    def update_ghost_fields(self, particles, fields, apply_boundary_condition=False):
        """Copy fields from each ghost's image (domain_manager.pyx:757-774),
        on the GPU."""
        fields = list(fields)
        if fields and particles.get_carray_size():
            lib, h = nat.flux(), nat.handle()
            c = cells_view(particles)
            ptrs = (ctypes.c_void_p * len(fields))(*[particles.pointer(f) for f in fields])
            size = (ctypes.c_int32 * len(fields))(
                *[particles.carrays[f].element_size() for f in fields])
            nat.check(lib.phd_update_ghost_fields(h, ctypes.byref(c), ptrs, size, len(fields)),
                      "update_ghost_fields")
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
