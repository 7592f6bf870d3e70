"""Riemann solver plugins (riemann/riemann.pyx): HLL, HLLC, Exact.

Same constructors, add_fields / initialize, compute_fluxes and
compute_time_step as the reference (riemann.pyx:15-222, 225-241, 507, 684-698);
the face loops and the CFL reduction run as HIP kernels.
"""
import ctypes

from . import _native as nat
from .containers import CarrayContainer
from .views import cells_view, faces_view


class RiemannBase(object):
    solver = None

    def __init__(self, cfl=0.5, **kwargs):
        self.cfl = float(cfl)
        self.fields_registered = False
        self.fluxes = None

    def add_fields(self, particles):
        g = particles.carray_named_groups
        if "conservative" not in g or "momentum" not in g:
            raise RuntimeError("ERROR: Missing fields in particles!")
        self.dim = len(g["position"])
        self.flux_fields = {f: "double" for f in g["conservative"]}
        self.flux_field_groups = {"conservative": list(g["conservative"]),
                                  "momentum": list(g["momentum"])}
        self.fields_registered = True

    def initialize(self):
        if not self.fields_registered:
            raise RuntimeError("Riemann did not set fields for flux!")
        self.fluxes = CarrayContainer(carrays_to_register=self.flux_fields)
        self.fluxes.carray_named_groups = dict(self.flux_field_groups)

    def compute_fluxes(self, particles, mesh, reconstruction, eos, check_errors=True):
        """riemann.pyx:74-98.  With deferred states from compute_states the
        states and the Riemann solve run as one fused kernel."""
        self.fluxes.resize(mesh.faces.get_carray_size())
        lib, h = nat.flux(), nat.handle()
        gamma = eos.get_gamma()
        fl = nat.ptr_array(self.fluxes.group("conservative"))
        f = faces_view(mesh)
        pending = reconstruction.take_pending()
        if pending is not None:
            pparticles, pmesh, pgamma, dt, boost = pending
            c = cells_view(pparticles)
            grad, ld = reconstruction._grad_ptr()
            rc = lib.phd_face_flux(h, pmesh.plan(pparticles), ctypes.byref(c), ctypes.byref(f),
                                   grad, ld, pgamma, dt,
                                   int(boost), int(reconstruction._gizmo()),
                                   reconstruction.order, self.solver, fl)
            nat.check(rc, "compute_fluxes")
        else:
            L = nat.ptr_array(reconstruction.left_states.group("primitive"))
            R = nat.ptr_array(reconstruction.right_states.group("primitive"))
            nat.check(lib.phd_compute_fluxes(h, ctypes.byref(f), self.dim, L, R, gamma,
                                             int(self.boost), self.solver, fl), "compute_fluxes")
        if check_errors and self.solver == nat.SOLVER_EXACT:
            nat.raise_device_fault(h)

    def compute_time_step(self, particles, eos):
        """riemann.pyx:120-188 (synchronising: returns a host float)."""
        lib, h = nat.flux(), nat.handle()
        c = cells_view(particles)
        out = ctypes.c_double()
        nat.check(lib.phd_compute_time_step(h, ctypes.byref(c), eos.get_gamma(), self.cfl,
                                            int(self.boost), ctypes.byref(out)),
                  "compute_time_step")
        return out.value


class HLL(RiemannBase):
    """Toro chapter 10 HLL (riemann.pyx:225-504)."""
    solver = nat.SOLVER_HLL

    def __init__(self, cfl=0.5, boost=True, **kwargs):
        self.boost = bool(boost)
        super(HLL, self).__init__(cfl, **kwargs)


class HLLC(HLL):
    """Toro chapter 10 HLLC (riemann.pyx:507-682)."""
    solver = nat.SOLVER_HLLC


class Exact(RiemannBase):
    """Toro chapter 4 exact solver (riemann.pyx:684-1090).  As in the
    reference the constructor forces boost on and cfl to 0.5 (695-698)."""
    solver = nat.SOLVER_EXACT

    def __init__(self, cfl=0.5, **kwargs):
        super(Exact, self).__init__(0.5, **kwargs)
        self.boost = True
