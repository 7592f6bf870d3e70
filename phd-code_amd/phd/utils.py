"""Particle tags, the hydro particle schema and the setter type check.

ParticleTAGS and the field schema follow utils/particle_tags.py and
utils/particle_creator.pyx:4-62; check_class follows utils/tools.py:19-31.
"""
import functools

from .containers import CarrayContainer


class ParticleTAGS:
    Real = 0
    Ghost = 1
    Exterior = 2
    Interior = 3
    Undefined = 4


class SimulationTAGS:
    PRE_EVOLVE = 0
    BEFORE_LOOP = 1
    MAIN_LOOP = 2
    AFTER_LOOP = 3
    POST_EVOLVE = 4


def HydroParticleCreator(num=0, dim=2, parallel=False, device=None):
    """Container with the reference's hydro fields and named groups
    (particle_creator.pyx:4-62): primitive = density, velocity-*, pressure;
    conservative = mass, momentum-*, energy; tag/type int8, ids int32."""
    pc = CarrayContainer(num, device=device)
    axes = "xyz"[:dim]
    groups = {"position": [], "velocity": [], "momentum": []}
    pc.register_carray(num, "density", "double")
    for a in axes:
        pc.register_carray(num, "position-" + a, "double")
        pc.register_carray(num, "velocity-" + a, "double")
        groups["position"].append("position-" + a)
        groups["velocity"].append("velocity-" + a)
    pc.register_carray(num, "pressure", "double")
    pc.register_carray(num, "mass", "double")
    for a in axes:
        pc.register_carray(num, "momentum-" + a, "double")
        groups["momentum"].append("momentum-" + a)
    pc.register_carray(num, "energy", "double")
    pc.register_carray(num, "tag", "int")
    pc.register_carray(num, "type", "int")
    pc.register_carray(num, "ids", "long")
    groups["primitive"] = ["density"] + groups["velocity"] + ["pressure"]
    groups["conservative"] = ["mass"] + groups["momentum"] + ["energy"]
    pc["tag"][:] = ParticleTAGS.Real
    pc["type"][:] = ParticleTAGS.Undefined
    pc.carray_named_groups = groups
    return pc


def check_class(cls):
    """Setter decorator: reject arguments that are not instances of cls."""
    def wrapper(func):
        @functools.wraps(func)
        def checked(self, value):
            if not isinstance(value, cls):
                raise RuntimeError("%s only accepts class type of %s" %
                                   (func.__name__, cls.__name__))
            return func(self, value)
        return checked
    return wrapper
