"""StaticMeshMUSCLHancock through the plugin API (integrate.py:295-378):
initialize, before_loop, then compute_time_step + evolve_timestep for a few
steps on a 2D Kelvin-Helmholtz set-up, each step compared with the oracle's
core step followed by its ghost refresh + EOS (the same two stages in the
same order, integrate.py:355-377), starting from the GPU's state.

The plugin path is reconstruction.compute_gradients -> compute_states ->
riemann.compute_fluxes -> mesh.update_from_fluxes ->
domain_manager.update_ghost_fields(conservative, True) ->
equation_state.primitive_from_conservative, every stage a library call.
"""
import math

import numpy as np
import pytest
import torch

from parity import assert_bitwise

pytestmark = pytest.mark.gpu


class _View(object):
    """What phd.problems.host_arrays reads of a set-up."""

    def __init__(self, integ, dim):
        self.particles = integ.particles
        self.mesh = integ.mesh
        self.dim = dim
        self.domain_manager = integ.domain_manager
        self.boundary = integ.boundary_condition
        self.eos = integ.equation_state


def _kh_static(n, limiter="arepo", gizmo=True, riemann="hllc"):
    import phd
    from phd import problems
    x = problems.jittered_lattice(n, 2, seed=7)
    pc = phd.HydroParticleCreator(num=x[0].size, dim=2)
    pc["position-x"][:] = torch.as_tensor(x[0], device=pc.device)
    pc["position-y"][:] = torch.as_tensor(x[1], device=pc.device)
    band = (pc["position-y"] - 0.5).abs() < 0.25
    pc["density"][:] = torch.where(band, 2.0, 1.0)
    pc["velocity-x"][:] = torch.where(band, 0.5, -0.5)
    pc["velocity-y"][:] = 0.01 * torch.sin(4.0 * math.pi * pc["position-x"])
    pc["pressure"][:] = 2.5
    integ = phd.StaticMeshMUSCLHancock()
    integ.set_particles(pc)
    integ.set_mesh(phd.Mesh())
    integ.set_equation_state(phd.IdealGas(5.0 / 3.0))
    integ.set_riemann(phd.HLLC(cfl=0.3) if riemann == "hllc" else phd.HLL(cfl=0.3))
    integ.set_reconstruction(phd.PieceWiseLinear(limiter=limiter, gizmo_limiter=gizmo))
    integ.set_domain_manager(phd.DomainManager([0.0, 0.0], [1.0, 1.0], 0.1))
    integ.set_boundary_condition(phd.Periodic())
    integ.initialize()
    integ.before_loop()
    return integ


@pytest.mark.parametrize("limiter,gizmo,riemann", [("arepo", True, "hllc"),
                                                   ("tess", False, "hll")])
def test_static_mesh_plugin_steps_match_oracle(gpu, limiter, gizmo, riemann):
    from oracle import oracle
    from phd import problems
    from phd import _native as nat
    print("case StaticMeshMUSCLHancock KH 24^2, %s limiter, gizmo %s, %s"
          % (limiter, gizmo, riemann), flush=True)
    integ = _kh_static(24, limiter, gizmo, riemann)
    view = _View(integ, 2)
    pc = integ.particles
    # before_loop: static mesh, zero generator and face velocities
    assert all(float(pc[f].abs().max()) == 0.0 for f in pc.carray_named_groups["w"])
    real = None
    for s in range(4):
        a = problems.host_arrays(view)
        real = a["tag"] == 0
        dt = integ.compute_time_step()
        dt_ref = oracle.time_step(2, a["gamma"], integ.riemann.cfl, integ.riemann.boost,
                                  a["tag"], a["prim"], a["volume"])
        assert dt == dt_ref, (s, dt, dt_ref)
        ref = oracle.core_step(a, dt, solver=integ.riemann.solver, boost=integ.riemann.boost,
                               gizmo=gizmo, limiter=integ.reconstruction.slope_limiter)
        assert ref["rc"] == 0
        cons, prim = oracle.ghost_refresh_eos(a, ref["cons"])
        t0, it0 = integ.time, integ.iteration
        integ.evolve_timestep()
        nat.raise_device_fault(nat.handle())
        torch.cuda.synchronize()
        assert integ.iteration == it0 + 1 and integ.time == t0 + dt
        for q, f in enumerate(pc.carray_named_groups["conservative"]):
            assert_bitwise(pc[f].cpu().numpy(), cons[q], "step %d %s" % (s, f))
        for q, f in enumerate(pc.carray_named_groups["primitive"]):
            assert_bitwise(pc[f].cpu().numpy(), prim[q], "step %d %s" % (s, f))
        flux = [integ.riemann.fluxes[f].cpu().numpy()
                for f in integ.riemann.fluxes.carray_named_groups["conservative"]]
        for q in range(4):
            assert_bitwise(flux[q], ref["flux"][q], "step %d flux[%d]" % (s, q))
    assert real.sum() == 24 * 24
