"""easyfea_amd: EasyFEA's finite-element solve path on AMD Instinct MI355X.

Element assembly, the deterministic CSR scatter, the block-Jacobi PCG solve
(FP64 or FP32-matrix with FP64 residual replacement) and the phase-field
staggered update run as hand-written HIP kernels in ``libefea_hip.so``
behind a C ABI (``include/efea.h``); this package is the host side, with the
reference's ``Models`` / ``Simulations`` interface.
"""

from . import Models, Simulations, mesh, mesh_io  # noqa: F401
from ._utils import ElemType, MatrixType  # noqa: F401
from .mesh import Mesh, hexa8_box, hexa8_cube, quad4_rectangle, tri3_notched_square, tri3_rectangle  # noqa: F401
from .mesh_io import load_mesh as Load_Mesh  # noqa: F401
from .mesh_io import read_msh  # noqa: F401
from .Simulations import Load_Simu  # noqa: F401
from .Solvers import AlgoType, SolverType  # noqa: F401

__version__ = "0.1.0"
