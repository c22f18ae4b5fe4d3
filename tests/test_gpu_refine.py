"""FP64 iterative refinement around an FP32 inner two-level PCG (GPU tests):
r = b - K u and the correction update in FP64, the inner solves with FP32
vectors, values, block Jacobi and coarse inverse.  Results within 1e-8 of the
oracle's direct solves on a 2D cantilever, a cracked phase-field body over
staggered increments and a 3D cube."""

import numpy as np
import pytest

import easyfea_amd as ea
from easyfea_amd import ElemType, Models, Simulations, workloads
from oracle import efea_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _refine(simu):
    simu.solver = "hip_mixed"
    simu.mixed_adaptive = False
    simu.mixed_strategy = "refine"
    simu.preconditioner = "two-level"


def test_refine_cantilever_matches_direct():
    simu = workloads.cantilever()
    _refine(simu)
    u = simu.Solve()
    info = simu.last_solve["elastic"]
    mesh = simu.mesh
    ref = O.ElasticOracle(mesh.connect, mesh.coord, "QUAD4", 210000.0, 0.3, True, 13.0,
                          {"SEG2": mesh.dict_groupElem[ElemType.SEG2].connect})
    ref.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: x == 0), [0, 0], ["x", "y"])
    ref.add_surfLoad(mesh.Nodes_Conditions(lambda x, y, z: x == 120.0), [-800.0 / 13.0 ** 2], ["y"])
    err = _rel(u, ref.Solve())
    print(f"cantilever refine: {info.iterations} inner iterations, {info.replacements} outer steps, "
          f"rel residual {info.rel_residual:.2e}, vs direct {err:.2e}, fallback {info.fallback}")
    assert info.refine and not info.fallback and info.status == 1 and info.replacements >= 2
    assert err < 1e-8


def test_refine_cube16_matches_direct():
    from easyfea_amd.mesh import hexa8_cube

    simu = workloads.cube(16, 0)
    _refine(simu)
    u = simu.Solve()
    info = simu.last_solve["elastic"]
    m = hexa8_cube(16)
    ref = O.ElasticOracle(m.connect, m.coord, "HEXA8", workloads.cube_E(m.Ne, 0), 0.3, False, 1.0,
                          {"QUAD4": m.dict_groupElem[ElemType.QUAD4].connect})
    ref.add_dirichlet(m.Nodes_Conditions(lambda x, y, z: z == 0), [0, 0, 0], ["x", "y", "z"])
    ref.add_surfLoad(m.Nodes_Conditions(lambda x, y, z: z == 1.0), [-10.0], ["z"])
    err = _rel(u, ref.Solve())
    print(f"16^3 refine: {info.iterations} inner iterations, {info.replacements} outer steps, vs direct {err:.2e}")
    assert info.refine and not info.fallback and err < 1e-8


def test_refine_cracked_body_staggered_matches_oracle():
    """Displacement solves of a growing crack (g(d) down to 1e-12 near the
    crack) through refinement; damage solves block Jacobi."""
    mesh = ea.tri3_notched_square(60)
    mat = Models.Elastic.Isotropic(2, 210000.0, 0.3, planeStress=False, thickness=1.0)
    simu = Simulations.PhaseField(mesh, Models.PhaseField(mat, "Miehe", "AT2", 2.7, 0.02))
    _refine(simu)
    ref = O.PhaseFieldOracle(mesh.connect, mesh.coord, "TRI3", 210000.0, 0.3, False, 1.0, "Miehe", "AT2", 2.7, 0.02)
    top = mesh.Nodes_Conditions(lambda x, y, z: y == 1.0)
    bot = mesh.Nodes_Conditions(lambda x, y, z: y == 0.0)
    rows = []
    for dep in [2e-3, 4e-3, 5.5e-3, 6.5e-3, 7.5e-3]:
        for s in (simu, ref):
            s.Bc_Init()
            s.add_dirichlet(top, [0, dep], ["x", "y"])
            s.add_dirichlet(bot, [0], ["y"])
        u, d, _ = simu.Solve(1.0, 10, 1)
        ur, dr, _ = ref.Solve(1.0, 10, 1)
        simu.Save_Iter()
        ref.Save_Iter()
        i = simu.last_solve["elastic"]
        rows.append((dep, i.refine, i.fallback, i.iterations, i.replacements, _rel(u, ur), _rel(d, dr),
                     float(dr.max())))
    print("dep, refine, fallback, inner iterations, outer steps, u err, d err, max d:", rows)
    assert rows[-1][-1] > 0.3  # the damage has grown
    for r in rows:
        assert r[1] and not r[2] and r[5] < 1e-8 and r[6] < 1e-8, rows
