"""Solver selection for ``simu.solver`` (``EasyFEA/Simulations/Solvers.py:200-222``).

The device path offers two solvers for the Dirichlet-reduced system of
``__Solver_1`` (``Solvers.py:502-553``):

* ``hip_fp64``  block-Jacobi PCG, FP64 matrix and vectors;
* ``hip_mixed`` block-Jacobi PCG whose Krylov products read an FP32 copy of
  the matrix (FP64 vectors and scalars), with the residual replaced by the
  FP64 ``b - K x`` whenever it falls tenfold (reliable updates), so the
  answer carries FP64 accuracy.

The reference's host backends (pypardiso, PETSc, scipy, ...) are not part
of this library; selecting one prints an error and keeps the current solver,
as the reference's setter does for an unknown value (``_simu.py:1151-1159``).
"""

from enum import Enum


class SolverType(str, Enum):
    hip_fp64 = "hip_fp64"
    hip_mixed = "hip_mixed"

    def __str__(self):
        return self.name


REFERENCE_ONLY = ("pypardiso", "petsc", "scipy", "lsq_linear", "cg", "bicg", "gmres", "lgmres")


class ResolType(str, Enum):
    r1 = "1"


class AlgoType(str, Enum):
    """Time scheme of ``K u^t + C v^t + M a^t = F^t`` (``Simulations/Solvers.py:64-190``):
    ``elliptic`` (``K u = F``), ``parabolic`` (generalized trapezoidal,
    ``alpha``), and the hyperbolic ``newmark``, ``midpoint``, ``hht``,
    ``hht_newmark``, ``euler_implicit`` and ``euler_explicit``."""

    elliptic = "elliptic"
    parabolic = "parabolic"
    newmark = "newmark"
    midpoint = "midpoint"
    hht = "hht"
    hht_newmark = "hht_newmark"
    euler_implicit = "euler_implicit"
    euler_explicit = "euler_explicit"

    def __str__(self) -> str:
        return self.name

    @staticmethod
    def Get_Hyperbolic_Types() -> list:
        return [AlgoType.newmark, AlgoType.midpoint, AlgoType.hht, AlgoType.hht_newmark, AlgoType.euler_implicit,
                AlgoType.euler_explicit]

    @staticmethod
    def Get_Hyperbolic_and_Parabolic_Types() -> list:
        return AlgoType.Get_Hyperbolic_Types() + [AlgoType.parabolic]
