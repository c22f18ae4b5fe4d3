"""Synthetic inputs of the measured workloads (seeded, reproducible).

* ``cantilever``: QUAD4 360x39 on 120x13 mm, plane stress, t = 13, clamped at
  x = 0, surface load -800/13^2 N/mm^2 in y at x = 120 (EasyFEA's
  ``examples/LinearizedElasticity/Elas1.py`` set-up).
* ``sent``: the single-edge-notched square in tension (AT2 phase field,
  ``examples/PhaseField/Tension.py`` boundary conditions).
* ``cube``: HEXA8 unit cube, lognormal i.i.d. Young's modulus per element.
* ``newmark_cube``: the cube in structural dynamics (steel density in
  tonne/mm^3, the load applied suddenly at t = 0), Newmark average
  acceleration with about 25 steps per period of the first shear mode.
* ``thermal_square``: QUAD4 unit square, lognormal conductivity field,
  t = 0 at x = 0, a heat flux of 5 through y = 1, Crank-Nicolson steps
  (``examples/Thermal/Thermal3.py``-like conditions).
* ``serving_elastic``: a stream of QUAD4 plane-stress squares of random side
  with a correlated lognormal Young's modulus field, Poisson arrivals.
* ``serving_phasefield``: phase-field realisations of the notched square with
  a correlated lognormal Gc field and random mesh sizes.

Correlated fields are white noise filtered with a Gaussian kernel in Fourier
space, rescaled to unit variance, then mapped to a lognormal of the given
mean and coefficient of variation: ``m exp(s z - s^2/2)``, ``s^2 = ln(1 + CoV^2)``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import Models, Simulations
from .mesh import Mesh, hexa8_cube, quad4_rectangle, tri3_notched_square


def gaussian_field(shape: tuple, corr_cells: float, rng: np.random.Generator) -> np.ndarray:
    """Unit-variance, zero-mean Gaussian random field on a periodic grid with a
    Gaussian covariance of correlation length ``corr_cells`` (in cells)."""
    w = rng.standard_normal(shape)
    freqs = np.meshgrid(*[np.fft.fftfreq(n) for n in shape], indexing="ij")
    k2 = sum(f * f for f in freqs)
    filt = np.exp(-0.5 * (2 * math.pi) ** 2 * k2 * corr_cells ** 2)
    z = np.real(np.fft.ifftn(np.fft.fftn(w) * filt))
    z -= z.mean()
    sd = z.std()
    return z / sd if sd > 0 else z


def lognormal(mean: float, cov: float, z: np.ndarray) -> np.ndarray:
    s = math.sqrt(math.log(1.0 + cov * cov))
    return mean * np.exp(s * z - 0.5 * s * s)


def cell_of_elements(mesh: Mesh, n: int) -> np.ndarray:
    """Flat index ``i + n*j`` of the cell of an ``n x n`` grid on the unit
    square holding each element's centroid."""
    cache = mesh.__dict__.setdefault("_cell_of_elements", {})
    if n not in cache:
        c = mesh.coord[mesh.connect].mean(axis=1)
        ij = np.clip(np.floor(c[:, :2] * n).astype(np.int64), 0, n - 1)
        cache[n] = ij[:, 0] + n * ij[:, 1]
    return cache[n]


# ------------------------------------------------------------ single solves
def cantilever(solver=None) -> Simulations.Elastic:
    mesh = quad4_rectangle(120.0, 13.0, 360, 39)
    mat = Models.Elastic.Isotropic(2, E=210000.0, v=0.3, planeStress=True, thickness=13.0)
    simu = Simulations.Elastic(mesh, mat)
    if solver:
        simu.solver = solver
    simu.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: x == 0), [0, 0], ["x", "y"])
    simu.add_surfLoad(mesh.Nodes_Conditions(lambda x, y, z: x == 120.0), [-800.0 / 13.0 ** 2], ["y"])
    return simu


def sent(cells: int, split: str = "Miehe", Gc=2.7, l0: float = 0.01, solver=None) -> Simulations.PhaseField:
    mesh = tri3_notched_square(cells)
    mat = Models.Elastic.Isotropic(2, E=210000.0, v=0.3, planeStress=False, thickness=1.0)
    simu = Simulations.PhaseField(mesh, Models.PhaseField(mat, split, "AT2", Gc, l0))
    if solver:
        simu.solver = solver
    return simu


def sent_bc(simu: Simulations.PhaseField, uy: float) -> None:
    """Bottom uy = 0; top ux = 0, uy = ``uy`` (``examples/PhaseField/Tension.py``)."""
    mesh = simu.mesh
    c = simu.__dict__.get("_sent_nodes")
    if c is None:
        c = simu.__dict__["_sent_nodes"] = (mesh.Nodes_Conditions(lambda x, y, z: y == 0.0),
                                            mesh.Nodes_Conditions(lambda x, y, z: y == 1.0))
    bottom, top = c
    simu.Bc_Init()
    simu.add_dirichlet(bottom, [0], ["y"])
    simu.add_dirichlet(top, [0, uy], ["x", "y"])


def sent_top_y_dofs(simu: Simulations.PhaseField) -> np.ndarray:
    top = simu.mesh.Nodes_Conditions(lambda x, y, z: y == 1.0)
    return simu.Bc_dofs_nodes(top, ["y"])


def cube_E(Ne: int, seed: int = 0) -> np.ndarray:
    """I.i.d. lognormal Young's modulus per element (mean 210000, CoV 0.2)."""
    rng = np.random.default_rng(seed)
    s = math.sqrt(math.log(1.04))
    return np.exp(rng.normal(math.log(210000.0) - 0.5 * s * s, s, Ne))


def cube(n: int, seed: int = 0, solver=None) -> Simulations.Elastic:
    mesh = hexa8_cube(n)
    E = cube_E(mesh.Ne, seed)
    simu = Simulations.Elastic(mesh, Models.Elastic.Isotropic(3, E=E, v=0.3))
    if solver:
        simu.solver = solver
    simu.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: z == 0), [0, 0, 0], ["x", "y", "z"])
    simu.add_surfLoad(mesh.Nodes_Conditions(lambda x, y, z: z == 1.0), [-10.0], ["z"])
    return simu


RHO_STEEL = 7.85e-9  # tonne / mm^3 (with MPa and mm)


def newmark_cube(n: int, seed: int = 0, solver=None, dt: float = 5e-8) -> Simulations.Elastic:
    """The cube of ``cube`` with steel density and the Newmark scheme
    (beta 1/4, gamma 1/2).  The first shear period of the clamped unit cube
    is about 4 L / sqrt(G / rho) = 1.2e-6 s, so dt = 5e-8 s takes ~25 steps
    per period."""
    simu = cube(n, seed, solver)
    simu.rho = RHO_STEEL
    simu.Solver_Set_Hyperbolic_Algorithm(dt)
    return simu


def thermal_k(n: int, seed: int = 0) -> np.ndarray:
    """Conductivity per element of the n x n square: a correlated lognormal
    field (mean 1, CoV 0.3, correlation 0.05 of the side)."""
    return lognormal(1.0, 0.3, gaussian_field((n, n), 0.05 * n, np.random.default_rng(seed)).T.ravel())


def thermal_square(n: int, seed: int = 0, solver=None, dt: float = 1e-4, alpha: float = 0.5) -> Simulations.Thermal:
    """QUAD4 n x n unit square, k a correlated lognormal field (mean 1, CoV
    0.3, correlation 0.05 of the side), c = rho = 1; t = 0 at x = 0, heat flux
    5 in through y = 1; the parabolic scheme with ``dt`` and ``alpha``."""
    mesh = quad4_rectangle(1.0, 1.0, n, n)
    simu = Simulations.Thermal(mesh, Models.Thermal(thermal_k(n, seed), 1.0))
    if solver:
        simu.solver = solver
    simu.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: x == 0), [0.0], ["t"])
    simu.add_surfLoad(mesh.Nodes_Conditions(lambda x, y, z: y == 1.0), [5.0], ["t"])
    simu.Solver_Set_Parabolic_Algorithm(dt, alpha)
    return simu


# ------------------------------------------------------------------ serving
@dataclass
class ElasticRequestSpec:
    rid: int
    side: int
    E: np.ndarray  # (side*side,) per element
    arrival: float


def serving_elastic_specs(n: int = 512, qps: float = 100.0, side_lo: int = 32, side_hi: int = 256,
                          seed: int = 0) -> list:
    """Request stream: side ~ U{side_lo..side_hi}, E a correlated lognormal
    field (mean 210000, CoV 0.2, correlation length 0.1 side cells; seed =
    request id), Poisson arrivals at ``qps``."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(side_lo, side_hi + 1, n)
    gaps = rng.exponential(1.0 / qps, n) if qps > 0 else np.zeros(n)
    arrivals = np.cumsum(gaps) - gaps[0]
    out = []
    for rid in range(n):
        s = int(sides[rid])
        z = gaussian_field((s, s), 0.1 * s, np.random.default_rng(rid))
        E = lognormal(210000.0, 0.2, z.T.ravel())  # element id i + s*j <- z[i, j]
        out.append(ElasticRequestSpec(rid, s, E, float(arrivals[rid])))
    return out


class MeshCache:
    """Meshes (and with them their device patterns) shared by requests of the
    same size; building them is set-up, reported apart from request latency."""

    def __init__(self, maker: Callable[[int], Mesh]):
        self.maker = maker
        self.meshes = {}

    def get(self, key: int) -> Mesh:
        m = self.meshes.get(key)
        if m is None:
            m = self.meshes[key] = self.maker(key)
        return m


def square_quad4(side: int) -> Mesh:
    return quad4_rectangle(1.0, 1.0, side, side)


def elastic_request(spec: ElasticRequestSpec, mesh: Mesh, mode: str, precond: str = "auto", max_blocks: int = 0):
    """The request's work as a generator of solve jobs; returns (u, info)."""
    mat = Models.Elastic.Isotropic(2, E=spec.E, v=0.3, planeStress=True, thickness=1.0)
    simu = Simulations.Elastic(mesh, mat)
    simu.solver = "hip_fp64" if mode == "fp64" else "hip_mixed"
    simu.mixed_strategy = "values"  # the batch engine's algorithm, alone or batched
    simu.pipelined = False
    simu.preconditioner = precond
    simu.max_blocks = max_blocks
    left, right = _square_sides(mesh)
    simu.add_dirichlet(left, [0, 0], ["x", "y"])
    simu.add_surfLoad(right, [1.0], ["x"])
    yield from simu._solve_gen()
    return simu._u, simu.last_solve["elastic"]


def _square_sides(mesh: Mesh):
    c = mesh.__dict__.get("_sides")
    if c is None:
        c = mesh.__dict__["_sides"] = (mesh.Nodes_Conditions(lambda x, y, z: x == 0.0),
                                       mesh.Nodes_Conditions(lambda x, y, z: x == 1.0))
    return c


@dataclass
class PhaseFieldRequestSpec:
    rid: int
    cells: int
    Gc: np.ndarray  # per element
    arrival: float
    steps: int = 50
    du: float = 1e-4
    step_latency: list = field(default_factory=list)
    forces: list = field(default_factory=list)


def serving_phasefield_specs(n: int = 64, steps: int = 50, ne_lo: int = 20000, ne_hi: int = 80000,
                             seed: int = 0) -> list:
    """Realisations: Ne ~ U[ne_lo, ne_hi] -> cells = sqrt(Ne/2) rounded to an
    even count (the notch ends at the centre node); Gc a
    correlated lognormal field (mean 2.7, CoV 0.15, correlation length
    0.05 mm); all arrive at t = 0 and are admitted by memory budget."""
    rng = np.random.default_rng(seed)
    ne = rng.uniform(ne_lo, ne_hi, n)
    out = []
    for rid in range(n):
        cells = 2 * int(round(math.sqrt(ne[rid] / 2.0) / 2.0))
        z = gaussian_field((cells, cells), 0.05 * cells, np.random.default_rng(10_000 + rid))
        out.append(PhaseFieldRequestSpec(rid, cells, z.T.ravel(), 0.0, steps))
    return out


def phasefield_request(spec: PhaseFieldRequestSpec, mesh: Mesh, mode: str, clock: Callable[[], float],
                       max_blocks: int = 0):
    """``spec.steps`` increments of the staggered loop (Amor, AT2, one pass per
    increment) as a generator of solve jobs; records per-increment latency and
    the reaction force on the top edge."""
    Gc_e = lognormal(2.7, 0.15, spec.Gc[cell_of_elements(mesh, spec.cells)])
    simu = _sent_on(mesh, Gc_e, mode)
    simu.max_blocks = max_blocks
    dofs = sent_top_y_dofs(simu)
    for n in range(1, spec.steps + 1):
        t = clock()
        sent_bc(simu, n * spec.du)
        yield from simu._staggered(1.0, 500, 1)
        spec.forces.append(float(np.sum(simu.Calc_Reaction(dofs))))
        simu.Save_Iter()
        spec.step_latency.append(clock() - t)
    return simu


def _sent_on(mesh: Mesh, Gc_e: np.ndarray, mode: str) -> Simulations.PhaseField:
    mat = Models.Elastic.Isotropic(2, E=210000.0, v=0.3, planeStress=False, thickness=1.0)
    simu = Simulations.PhaseField(mesh, Models.PhaseField(mat, "Amor", "AT2", Gc_e, 0.01))
    simu.solver = "hip_fp64" if mode == "fp64" else "hip_mixed"
    simu.mixed_strategy = "values"  # the batch engine's algorithm, alone or batched
    simu.pipelined = False
    return simu
