"""Ideal-gas equation of state (equation_state/equation_state.pyx:5-125)."""
import ctypes
import math

from . import _native as nat
from .views import cells_view


class EquationStateBase(object):
    def conservative_from_primitive(self, particles):
        raise NotImplementedError("EquationStateBase::conservative_from_primitive called!")

    def primitive_from_conservative(self, particles):
        raise NotImplementedError("EquationStateBase::primitive_from_conservative called!")

    def sound_speed(self, density, pressure):
        raise NotImplementedError("EquationStateBase::sound_speed called!")

    def get_gamma(self):
        raise NotImplementedError("EquationStateBase::get_gamma called!")


class IdealGas(EquationStateBase):
    """Ideal gas law; conversions run over all cells (real + ghost) on the
    GPU (equation_state.pyx:43-117)."""

    def __init__(self, gamma=1.4):
        self.gamma = float(gamma)

    def conservative_from_primitive(self, particles):
        lib, h = nat.flux(), nat.handle()
        c = cells_view(particles)
        nat.check(lib.phd_conservative_from_primitive(h, ctypes.byref(c), self.gamma),
                  "conservative_from_primitive")

    def primitive_from_conservative(self, particles):
        lib, h = nat.flux(), nat.handle()
        c = cells_view(particles)
        nat.check(lib.phd_primitive_from_conservative(h, ctypes.byref(c), self.gamma),
                  "primitive_from_conservative")

    def sound_speed(self, density, pressure):
        """equation_state.pyx:119-121 (a host scalar, as in the reference)."""
        return math.sqrt(self.gamma * pressure / density)

    def get_gamma(self):
        return self.gamma
