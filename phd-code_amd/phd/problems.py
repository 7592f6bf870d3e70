"""Synthetic problem setups on seeded jittered lattices, and a frozen-mesh
driver of the fused core step.

Generators sit at ((i,j[,k]) + 0.5)/n + U(-0.1, 0.1)/n per axis
(numpy.random.default_rng(seed)), optionally renumbered along a Morton curve
of the lattice index so that neighbouring cells are close in memory.

Configurations (sizes are arguments):
  sod_2d   : reflective unit box, gamma 1.4, x < 0.5: rho 1, p 1; else rho
             0.125, p 0.1; plain mesh motion (Mesh(regularize=False)), HLLC
             with cfl 0.3.
  kh_2d    : periodic unit box, gamma 5/3, rho 2 for |y-0.5| < 0.25 else 1,
             vx +0.5 inside / -0.5 outside, p 2.5, vy 0.01 sin(4 pi x).
  sedov_3d : reflective unit box, gamma 1.4, rho 1, p 1e-5, v 0, E = 1 put
             in the 32 generators nearest the centre with uniform energy
             density (p = (gamma-1) E / sum V).
"""
import ctypes
import math

import numpy as np
import torch

from . import _native as nat
from .domain import DomainManager, Periodic, Reflective
from .equation_state import IdealGas
from .integrate import MovingMeshMUSCLHancock
from .mesh import Mesh
from .reconstruction import PieceWiseLinear
from .riemann import HLLC, HLL
from .utils import HydroParticleCreator, ParticleTAGS
from .views import cells_view, faces_view


def _morton_key(idx, dim, bits):
    key = np.zeros(idx.shape[1], dtype=np.uint64)
    for b in range(bits):
        for a in range(dim):
            key |= ((idx[a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(dim * b + a)
    return key


def jittered_lattice(n, dim, seed=0, jitter=0.1, order="morton"):
    """Generators of an n^dim jittered lattice in the unit box (list of
    dim float64 arrays)."""
    rng = np.random.default_rng(seed)
    idx = np.indices((n,) * dim).reshape(dim, -1)
    x = [(idx[a] + 0.5) / n + rng.uniform(-jitter, jitter, idx.shape[1]) / n
         for a in range(dim)]
    if order == "morton":
        bits = max(1, int(math.ceil(math.log2(n))))
        perm = np.argsort(_morton_key(idx.astype(np.uint64), dim, bits), kind="stable")
        x = [xa[perm] for xa in x]
    return [np.ascontiguousarray(xa) for xa in x]


class Problem(object):
    """A set-up integrator (mesh built, conserved fields computed)."""

    def __init__(self, name, dim, x, gamma, boundary, init, riemann, mesh, recon=None,
                 device=None, n_threads=0):
        self.name = name
        self.dim = dim
        n = x[0].size
        particles = HydroParticleCreator(num=n, dim=dim, device=device)
        for k, f in enumerate(particles.carray_named_groups["position"]):
            particles[f][:] = torch.as_tensor(x[k], device=particles.device)
        particles["ids"][:] = torch.arange(n, dtype=torch.int32, device=particles.device)
        self.particles = particles
        self.eos = IdealGas(gamma)
        self.domain_manager = DomainManager([0.0] * dim, [1.0] * dim, 0.1)
        self.boundary = boundary
        self.mesh = mesh
        self.mesh.n_threads = n_threads
        self.recon = recon if recon is not None else PieceWiseLinear()
        self.riemann = riemann
        integ = MovingMeshMUSCLHancock()
        integ.set_mesh(self.mesh)
        integ.set_equation_state(self.eos)
        integ.set_riemann(self.riemann)
        integ.set_particles(particles)
        integ.set_domain_manager(self.domain_manager)
        integ.set_boundary_condition(boundary)
        integ.set_reconstruction(self.recon)
        integ.initialize()
        self.integrator = integ
        # mesh first (volumes are needed by the Sedov set-up), then state
        self.domain_manager.partition(particles)
        self.mesh.build_geometry(particles, self.domain_manager)
        init(self)
        self.domain_manager.update_ghost_fields(particles, particles.carray_named_groups["primitive"])
        boundary.update_fields(particles, self.domain_manager)
        self.eos.conservative_from_primitive(particles)

    @property
    def n_real(self):
        return int((self.particles["tag"] == ParticleTAGS.Real).sum().item())

    @property
    def n_faces(self):
        return self.mesh.faces.get_carray_size()


def sod_2d(n=100, seed=0, device=None, order="morton"):
    x = jittered_lattice(n, 2, seed, order=order)

    def init(p):
        pc = p.particles
        left = pc["position-x"] < 0.5
        pc["density"][:] = torch.where(left, 1.0, 0.125)
        pc["pressure"][:] = torch.where(left, 1.0, 0.1)
        pc["velocity-x"][:] = 0.0
        pc["velocity-y"][:] = 0.0
    return Problem("sod_2d", 2, x, 1.4, Reflective(), init, HLLC(cfl=0.3),
                   Mesh(regularize=False), device=device)


def kh_2d(n=512, seed=0, device=None, order="morton"):
    x = jittered_lattice(n, 2, seed, order=order)

    def init(p):
        pc = p.particles
        y = pc["position-y"]
        band = (y - 0.5).abs() < 0.25
        pc["density"][:] = torch.where(band, 2.0, 1.0)
        pc["velocity-x"][:] = torch.where(band, 0.5, -0.5)
        pc["velocity-y"][:] = 0.01 * torch.sin(4.0 * math.pi * pc["position-x"])
        pc["pressure"][:] = 2.5
    return Problem("kh_2d", 2, x, 5.0 / 3.0, Periodic(), init, HLLC(), Mesh(), device=device)


def sedov_3d(n=128, seed=0, device=None, order="morton", n_threads=0, energy=1.0, n_hot=32):
    x = jittered_lattice(n, 3, seed, order=order)
    gamma = 1.4

    def init(p):
        pc = p.particles
        pc["density"][:] = 1.0
        for f in pc.carray_named_groups["velocity"]:
            pc[f][:] = 0.0
        pc["pressure"][:] = 1.0e-5
        real = pc["tag"] == ParticleTAGS.Real
        r2 = sum((pc[f] - 0.5) ** 2 for f in pc.carray_named_groups["position"])
        r2 = torch.where(real, r2, torch.full_like(r2, float("inf")))
        hot = torch.argsort(r2, stable=True)[:n_hot]
        vol = pc["volume"][hot].sum()
        pc["pressure"][hot] = (gamma - 1.0) * energy / vol
    return Problem("sedov_3d", 3, x, gamma, Reflective(), init, HLLC(), Mesh(), device=device,
                   n_threads=n_threads)


PROBLEMS = {"sod_2d": sod_2d, "kh_2d": kh_2d, "sedov_3d": sedov_3d}


class FrozenStep(object):
    """The core face-flux step of MovingMeshMUSCLHancock.evolve_timestep
    (integrate.py:396-408) on a frozen mesh, called straight through the C
    ABI: gradients + limiter, ghost gradients, fused states + Riemann +
    deboost, conserved update.  The prologue (generator and face velocities)
    and the epilogue (ghost refresh + EOS) are separate calls."""

    def __init__(self, problem):
        self.p = problem
        self.lib = nat.flux()
        pc, mesh = problem.particles, problem.mesh
        self.cells = cells_view(pc)
        self.faces = faces_view(mesh)
        recon = problem.recon
        recon.grad.resize(pc.get_carray_size())
        self.grad = recon.grad
        problem.riemann.fluxes.resize(mesh.faces.get_carray_size())
        self.flux_ptrs = nat.ptr_array(problem.riemann.fluxes.group("conservative"))
        dm = problem.domain_manager
        dim = problem.dim
        self.lo = (ctypes.c_double * 3)(*(dm.xmin + [0.0] * (3 - dim)))
        self.hi = (ctypes.c_double * 3)(*(dm.xmax + [0.0] * (3 - dim)))
        self.boundary = problem.boundary.native_code
        self.gamma = problem.eos.get_gamma()
        self.limiter = recon.slope_limiter
        self.gizmo = int(recon.gizmo_limiter)
        self.solver = problem.riemann.solver
        self.boost = int(problem.riemann.boost)
        self.plan = mesh.plan(pc)

    def prologue(self):
        """Generator and face velocities (mesh.pyx:363-477)."""
        self.p.mesh.assign_generator_velocities(self.p.particles, self.p.eos)
        self.p.mesh.assign_face_velocities(self.p.particles)

    def time_step(self):
        return self.p.riemann.compute_time_step(self.p.particles, self.p.eos)

    def run(self, dt, h=None, write_flux=True, xp_fresh=False):
        """One core step.  write_flux=False skips writing riemann.fluxes (the
        update uses the (dt*area)*flux rows either way); xp_fresh=True when
        the previous epilogue refreshed the plan's packed cell rows."""
        h = nat.handle() if h is None else h
        rc = self.lib.phd_core_step(h, self.plan, ctypes.byref(self.cells),
                                    ctypes.byref(self.faces), self.grad.data.data_ptr(),
                                    self.grad.ld, self.limiter, self.boundary,
                                    ctypes.cast(self.lo, ctypes.c_void_p),
                                    ctypes.cast(self.hi, ctypes.c_void_p), self.gamma, float(dt),
                                    self.boost, self.gizmo, self.solver,
                                    self.flux_ptrs if write_flux else None,
                                    nat.CORE_XP_FRESH if xp_fresh else 0)
        nat.check(rc, "phd_core_step")

    def epilogue(self, h=None):
        """Ghost conserved refresh with the boundary rule + EOS, all cells;
        refreshes the plan's packed cell rows."""
        h = nat.handle() if h is None else h
        rc = self.lib.phd_ghost_refresh_eos(h, self.plan, ctypes.byref(self.cells),
                                            self.boundary,
                                            ctypes.cast(self.lo, ctypes.c_void_p),
                                            ctypes.cast(self.hi, ctypes.c_void_p), self.gamma)
        nat.check(rc, "phd_ghost_refresh_eos")

    def check(self):
        nat.raise_device_fault(nat.handle())


def host_arrays(problem):
    """Every array the core step reads, as host numpy (for the oracle)."""
    pc, mesh = problem.particles, problem.mesh
    g = pc.carray_named_groups
    cpu = lambda t: t.detach().cpu().numpy().copy()
    out = {
        "dim": problem.dim,
        "n": pc.get_carray_size(),
        "nf": mesh.faces.get_carray_size(),
        "tag": cpu(pc["tag"]), "type": cpu(pc["type"]), "map": cpu(pc["map"]),
        "x": [cpu(pc[f]) for f in g["position"]],
        "dcom": [cpu(pc[f]) for f in g["dcom"]],
        "volume": cpu(pc["volume"]),
        "prim": [cpu(pc[f]) for f in g["primitive"]],
        "cons": [cpu(pc[f]) for f in g["conservative"]],
        "w": [cpu(pc[f]) for f in g["w"]],
        "area": cpu(mesh.faces["area"]),
        "com": [cpu(mesh.faces[f]) for f in mesh.faces.carray_named_groups["com"]],
        "normal": [cpu(mesh.faces[f]) for f in mesh.faces.carray_named_groups["normal"]],
        "fv": [cpu(mesh.faces[f]) for f in mesh.faces.carray_named_groups["velocity"]],
        "pair_i": cpu(mesh.faces["pair-i"]), "pair_j": cpu(mesh.faces["pair-j"]),
        "nbr_off": cpu(mesh.neighbors.offsets), "nbr_face": cpu(mesh.neighbors.faces),
        "lo": list(problem.domain_manager.xmin), "hi": list(problem.domain_manager.xmax),
        "boundary": problem.boundary.native_code, "gamma": problem.eos.get_gamma(),
    }
    return out
