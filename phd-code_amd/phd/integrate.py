"""Integrators (integrate/integrate.py): the callers of the hot path.

StaticMeshMUSCLHancock / MovingMeshMUSCLHancock keep the reference's setters,
initialize / before_loop / compute_time_step / evolve_timestep and call order
(integrate.py:17-428); each hot-path call lands in a HIP kernel.
"""
import logging

import numpy as np

from .containers import CarrayContainer
from .domain import BoundaryConditionBase, DomainManager
from .equation_state import EquationStateBase
from .mesh import Mesh
from .reconstruction import ReconstructionBase
from .riemann import RiemannBase
from .utils import check_class

phdLogger = logging.getLogger("phd")


class IntegrateBase(object):
    def __init__(self, dt=0., time=0., iteration=0, old_dt=-np.inf, max_dt_change=2.0,
                 initial_timestep_factor=1.0, restart=False, **kwargs):
        self.dt = dt
        self.time = time
        self.old_dt = old_dt
        self.iteration = iteration
        self.restart = restart
        self.max_dt_change = max_dt_change
        self.initial_timestep_factor = initial_timestep_factor
        self.mesh = None
        self.riemann = None
        self.particles = None
        self.reconstruction = None
        self.domain_manager = None
        self.boundary_condition = None
        self.equation_state = None
        self.source_terms = {}

    @check_class(BoundaryConditionBase)
    def set_boundary_condition(self, boundary_condition):
        self.boundary_condition = boundary_condition

    @check_class(DomainManager)
    def set_domain_manager(self, domain_manager):
        self.domain_manager = domain_manager

    @check_class(EquationStateBase)
    def set_equation_state(self, equation_state):
        self.equation_state = equation_state

    def set_load_balance(self, load_balance):
        self.load_balance = load_balance

    @check_class(CarrayContainer)
    def set_particles(self, particles):
        self.particles = particles

    @check_class(Mesh)
    def set_mesh(self, mesh):
        self.mesh = mesh

    @check_class(ReconstructionBase)
    def set_reconstruction(self, reconstruction):
        self.reconstruction = reconstruction

    @check_class(RiemannBase)
    def set_riemann(self, riemann):
        self.riemann = riemann

    def add_source_term(self, source_term):
        self.source_terms[source_term.__class__.__name__] = source_term

    def initialize(self):
        raise NotImplementedError("IntegrateBase::initialize called!")

    def compute_source(self, term):
        for source in self.source_terms.values():
            if term == "motion":
                source.apply_motion(self)
            elif term == "primitive":
                source.apply_primitive(self)
            elif term == "conservative":
                source.apply_conservative(self)
            elif term == "flux":
                source.apply_flux(self)
            elif term == "compute":
                source.compute_source(self)
            else:
                raise RuntimeError("ERROR: Unknown source term method")

    def before_loop(self, simulation=None):
        """integrate.py:188-231 (mesh relaxation is not part of this build)."""
        self.domain_manager.partition(self.particles)
        phdLogger.info("IntegrateBase: Building initial mesh")
        self.mesh.build_geometry(self.particles, self.domain_manager)
        self.domain_manager.boundary_condition.update_fields(self.particles, self.domain_manager)
        self.equation_state.conservative_from_primitive(self.particles)
        self.compute_source("compute")

    def compute_time_step(self):
        """integrate.py:233-284 (single GPU: no dt all-reduce)."""
        dt = self.riemann.compute_time_step(self.particles, self.equation_state)
        self.dt = dt
        if self.source_terms:
            dt_source = np.inf
            for source in self.source_terms.values():
                dt_source = min(source.compute_time_step(self), dt_source)
            dt = min(dt, dt_source)
        self.dt = dt
        return dt

    def evolve_timestep(self):
        raise NotImplementedError("IntegrateBase::evolve_timestep called!")

    def after_loop(self, simulation=None):
        pass


class StaticMeshMUSCLHancock(IntegrateBase):
    """integrate.py:295-378."""

    def initialize(self):
        if (not self.mesh or not self.riemann or self.particles is None or
                not self.domain_manager or not self.equation_state or
                not self.reconstruction or not self.boundary_condition):
            raise RuntimeError("ERROR: Not all setters defined in %s!" % self.__class__.__name__)
        self.domain_manager.set_boundary_condition(self.boundary_condition)
        self.domain_manager.register_fields(self.particles)
        self.domain_manager.initialize()
        self.mesh.register_fields(self.particles)
        self.mesh.initialize()
        self.reconstruction.add_fields(self.particles)
        self.reconstruction.initialize()
        self.riemann.add_fields(self.particles)
        self.riemann.initialize()

    def before_loop(self, simulation=None):
        super(StaticMeshMUSCLHancock, self).before_loop(simulation)
        dim = len(self.particles.carray_named_groups["position"])
        for a in "xyz"[:dim]:
            self.particles["w-" + a][:] = 0.
            self.mesh.faces["velocity-" + a][:] = 0.

    def evolve_timestep(self):
        phdLogger.info("StaticMeshMUSCLHancock: Starting integration")
        self.reconstruction.compute_gradients(self.particles, self.mesh, self.domain_manager)
        self.reconstruction.compute_states(self.particles, self.mesh,
                                           self.equation_state.get_gamma(), self.domain_manager,
                                           0.5 * self.dt, self.riemann.boost)
        self.compute_source("primitive")
        self.riemann.compute_fluxes(self.particles, self.mesh, self.reconstruction,
                                    self.equation_state)
        self.mesh.update_from_fluxes(self.particles, self.riemann, self.dt)
        self.compute_source("flux")
        self.compute_source("compute")
        self.compute_source("conservative")
        self.domain_manager.update_ghost_fields(
            self.particles, self.particles.carray_named_groups["conservative"], True)
        self.equation_state.primitive_from_conservative(self.particles)
        self.iteration += 1
        self.time += self.dt


class MovingMeshMUSCLHancock(StaticMeshMUSCLHancock):
    """integrate.py:380-428."""

    def before_loop(self, simulation=None):
        IntegrateBase.before_loop(self, simulation)

    def evolve_timestep(self):
        phdLogger.info("MovingMeshMUSCLHancock: Starting integration")
        self.mesh.assign_generator_velocities(self.particles, self.equation_state)
        self.compute_source("motion")
        self.mesh.assign_face_velocities(self.particles)
        self.reconstruction.compute_gradients(self.particles, self.mesh, self.domain_manager)
        self.reconstruction.compute_states(self.particles, self.mesh,
                                           self.equation_state.get_gamma(), self.domain_manager,
                                           0.5 * self.dt, self.riemann.boost)
        self.compute_source("primitive")
        self.riemann.compute_fluxes(self.particles, self.mesh, self.reconstruction,
                                    self.equation_state)
        self.mesh.update_from_fluxes(self.particles, self.riemann, self.dt)
        self.compute_source("flux")
        self.domain_manager.move_generators(self.particles, self.dt)
        self.domain_manager.partition(self.particles)
        self.mesh.build_geometry(self.particles, self.domain_manager)
        self.domain_manager.boundary_condition.update_fields(self.particles, self.domain_manager)
        self.compute_source("compute")
        self.compute_source("conservative")
        self.equation_state.primitive_from_conservative(self.particles)
        self.iteration += 1
        self.time += self.dt
