"""Face-state reconstruction plugins (reconstruction/reconstruction.pyx).

PieceWiseLinear keeps the reference's constructor, add_fields / initialize
and compute_gradients / add_spatial / add_temporal / compute_states
signatures (reconstruction.pyx:355-1079); the loops run as HIP kernels.

compute_states defers its work: the integrator always follows it with
riemann.compute_fluxes, which then runs one fused kernel that builds the
states in registers and never writes them to memory.  Reading left_states or
right_states in between materialises them with the stand-alone states
kernel (bit-identical), so the public attributes keep their meaning.
"""
import ctypes

from . import _native as nat
from .containers import CarrayContainer, GradientContainer
from .views import cells_view, faces_view


class ReconstructionBase(object):
    order = None

    def __init__(self):
        self.fields_registered = False
        self.has_passive_scalars = False
        self._left = None
        self._right = None
        self._pending = None

    def initialize(self):
        raise NotImplementedError("Reconstruction::initialize called!")

    def add_fields(self, particles):
        raise NotImplementedError("Reconstruction::initialize called!")

    def compute_gradients(self, particles, mesh, domain_manager):
        raise NotImplementedError("Reconstruction::_compute_gradients called!")

    def compute_states(self, particles, mesh, gamma, domain_manager, dt, boost):
        raise NotImplementedError("Reconstruction::compute called!")

    # -- deferred face states -------------------------------------------------
    def _defer(self, particles, mesh, gamma, dt, boost):
        nf = mesh.faces.get_carray_size()
        self._left.resize(nf)
        self._right.resize(nf)
        self._pending = (particles, mesh, float(gamma), float(dt), bool(boost))

    def take_pending(self):
        """Hand the deferred state build to a fused flux kernel."""
        p, self._pending = self._pending, None
        return p

    @property
    def left_states(self):
        self._materialize()
        return self._left

    @property
    def right_states(self):
        self._materialize()
        return self._right

    def _materialize(self):
        if self._pending is None:
            return
        particles, mesh, gamma, dt, boost = self._pending
        self._pending = None
        self._run_states(particles, mesh, gamma, dt, boost,
                         nat.STAGE_SPATIAL | nat.STAGE_TEMPORAL)

    def _run_states(self, particles, mesh, gamma, dt, boost, stages):
        lib, h = nat.flux(), nat.handle()
        c, f = cells_view(particles), faces_view(mesh)
        grad, ld = self._grad_ptr()
        L = nat.ptr_array(self._left.group("primitive"))
        R = nat.ptr_array(self._right.group("primitive"))
        nat.check(lib.phd_compute_states(h, mesh.plan(particles), ctypes.byref(c),
                                         ctypes.byref(f), grad, ld, gamma, dt,
                                         int(boost), int(self._gizmo()), self.order, stages, L, R),
                  "compute_states")

    def _grad_ptr(self):
        return None, 0

    def _gizmo(self):
        return False

    def _state_containers(self):
        self._left = CarrayContainer(carrays_to_register=self.reconstruct_fields)
        self._right = CarrayContainer(carrays_to_register=self.reconstruct_fields)
        self._left.carray_named_groups = dict(self.reconstruct_field_groups)
        self._right.carray_named_groups = dict(self.reconstruct_field_groups)


class PieceWiseConstant(ReconstructionBase):
    """reconstruction.pyx:131-352: face states are the cell values (face
    frame velocities when boosted)."""
    order = nat.ORDER_CONSTANT

    def initialize(self):
        if not self.fields_registered:
            raise RuntimeError("Reconstruction did not set fields to reconstruct!")
        self._state_containers()

    def add_fields(self, particles):
        g = particles.carray_named_groups
        if "primitive" not in g or "velocity" not in g:
            raise RuntimeError("ERROR: Missing fields in particles!")
        self.reconstruct_fields = {f: "double" for f in g["primitive"]}
        self.reconstruct_field_groups = {"primitive": list(g["primitive"]),
                                         "velocity": list(g["velocity"])}
        self.fields_registered = True

    def compute_gradients(self, particles, mesh, domain_manager):
        pass  # no gradients for constant reconstruction

    def compute_states(self, particles, mesh, gamma, domain_manager, dt, boost):
        self._defer(particles, mesh, gamma, dt, boost)


class PieceWiseLinear(ReconstructionBase):
    """AREPO-style linear reconstruction (reconstruction.pyx:355-1079)."""
    order = nat.ORDER_LINEAR

    def __init__(self, limiter="arepo", gizmo_limiter=True, **kwargs):
        super(PieceWiseLinear, self).__init__()
        if limiter == "arepo":
            self.slope_limiter = nat.LIMITER_AREPO
        elif limiter == "tess":
            self.slope_limiter = nat.LIMITER_TESS
        else:
            raise RuntimeError("ERROR: Unrecognized limiter")
        self.limiter = limiter
        self.gizmo_limiter = bool(gizmo_limiter)
        self.grad = None

    def add_fields(self, particles):
        g = particles.carray_named_groups
        if "primitive" not in g or "velocity" not in g:
            raise RuntimeError("ERROR: Missing fields in particles!")
        self.dim = len(g["position"])
        self.reconstruct_fields = {f: "double" for f in g["primitive"]}
        self.reconstruct_field_groups = {"primitive": list(g["primitive"]),
                                         "velocity": list(g["velocity"])}
        self.grad_field_names = list(g["primitive"])
        self.fields_registered = True

    def initialize(self):
        if not self.fields_registered:
            raise RuntimeError("Reconstruction did not set fields to reconstruct!")
        self._state_containers()
        self.grad = GradientContainer(self.grad_field_names, self.dim,
                                      nat.record_ld(self.dim), device=self._left.device)
        self.reconstruct_grads = {g: "double" for g in self.grad.names}
        self.reconstruct_grad_groups = self.grad.carray_named_groups

    def _grad_ptr(self):
        return self.grad.data.data_ptr(), self.grad.ld

    def _gizmo(self):
        return self.gizmo_limiter

    def compute_gradients(self, particles, mesh, domain_manager, check_errors=True):
        """Eq. 21 gradients + limiter for real cells (563-737), then the ghost
        transfer (739)."""
        self.grad.resize(particles.get_carray_size())
        lib, h = nat.flux(), nat.handle()
        c = cells_view(particles)
        nat.check(lib.phd_compute_gradients(h, mesh.plan(particles), ctypes.byref(c),
                                            self.slope_limiter, self.grad.data.data_ptr(),
                                            self.grad.ld), "compute_gradients")
        if check_errors:
            nat.raise_device_fault(h)
        domain_manager.update_ghost_gradients(particles, self.grad, mesh)

    def add_spatial(self, particles, mesh, gamma, domain_manager, dt, boost):
        self._materialize()
        nf = mesh.faces.get_carray_size()
        self._left.resize(nf)
        self._right.resize(nf)
        self._run_states(particles, mesh, float(gamma), float(dt), boost, nat.STAGE_SPATIAL)

    def add_temporal(self, particles, mesh, gamma, domain_manager, dt, boost):
        self._materialize()
        self._run_states(particles, mesh, float(gamma), float(dt), boost, nat.STAGE_TEMPORAL)

    def compute_states(self, particles, mesh, gamma, domain_manager, dt, boost):
        self._defer(particles, mesh, gamma, dt, boost)
