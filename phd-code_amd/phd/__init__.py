"""phd -- MI355X-native face-flux pipeline behind the plugin API of
rickyfernandez/phd-code (moving-mesh finite-volume hydrodynamics).

Importing the package needs no GPU; every hot-path method runs a HIP kernel
of libphd_flux.so and raises NativeUnavailable when no GPU is visible.
Public names follow the reference's phd/__init__.py.
"""
from .containers import CarrayContainer, GradientContainer
from .utils import ParticleTAGS, SimulationTAGS, HydroParticleCreator, check_class
from .equation_state import EquationStateBase, IdealGas
from .mesh import Mesh
from .domain import DomainManager, BoundaryConditionBase, Reflective, Periodic
from .reconstruction import ReconstructionBase, PieceWiseConstant, PieceWiseLinear
from .riemann import RiemannBase, HLL, HLLC, Exact
from .integrate import IntegrateBase, StaticMeshMUSCLHancock, MovingMeshMUSCLHancock
from .read_write import ReaderWriterBase, Hdf5, NpzTree
from ._native import NativeUnavailable

# single GPU: no MPI communicator (the reference's phd._comm globals)
_comm = None
_rank = 0
_size = 1
_in_parallel = False
