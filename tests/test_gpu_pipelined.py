"""Pipelined two-level PCG (``efea_solve_pcg2p``, GPU tests): the FP64 solve
with two launches per iteration gives the direct solution within 1e-8 on a
2D cantilever, a 3D cube and a cracked phase-field body over staggered
increments; it is deterministic and takes about the classic solve's
iteration count.  With FP32 matrix values in the iterations (the mixed
"values" strategy, forced at these sizes) it reaches the same 1e-8 on the
cantilever and the cracked body."""

import numpy as np
import pytest

import easyfea_amd as ea
from easyfea_amd import ElemType, Models, Simulations, workloads
from oracle import efea_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _cantilever(pipelined):
    simu = workloads.cantilever("hip_fp64")
    simu.preconditioner = "two-level"
    simu.pipelined = pipelined
    return simu


def test_pipelined_cantilever_matches_direct_and_classic():
    s1, s2 = _cantilever(True), _cantilever(False)
    u1, u2 = s1.Solve(), s2.Solve()
    i1, i2 = s1.last_solve["elastic"], s2.last_solve["elastic"]
    mesh = s1.mesh
    ref = O.ElasticOracle(mesh.connect, mesh.coord, "QUAD4", 210000.0, 0.3, True, 13.0,
                          {"SEG2": mesh.dict_groupElem[ElemType.SEG2].connect})
    ref.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: x == 0), [0, 0], ["x", "y"])
    ref.add_surfLoad(mesh.Nodes_Conditions(lambda x, y, z: x == 120.0), [-800.0 / 13.0 ** 2], ["y"])
    ur = ref.Solve()
    print(f"cantilever: pipelined {i1.iterations} it ({_rel(u1, ur):.2e} vs direct), classic {i2.iterations} it "
          f"({_rel(u2, ur):.2e}); pipelined failure: {getattr(s1.last_job, 'pipeline_failure', None)}")
    assert i1.pipelined and not i2.pipelined and i1.status == 1
    assert _rel(u1, ur) < 1e-8
    assert abs(i1.iterations - i2.iterations) <= max(3, 0.05 * i2.iterations)
    # deterministic
    s1._u.zero_()
    s1.Need_Update()
    assert np.array_equal(s1.Solve(), u1)


@pytest.fixture
def _fp32_values(monkeypatch):
    """FP32 matrix values in the Krylov products at every size."""
    import easyfea_amd.device as dv

    monkeypatch.setattr(dv, "MIXED_MIN_MATRIX_BYTES", 0)


def test_pipelined_mixed_cantilever_matches_direct(_fp32_values):
    s1 = workloads.cantilever("hip_mixed")
    s1.preconditioner = "two-level"
    s1.pipelined = True
    u1 = s1.Solve()
    i1 = s1.last_solve["elastic"]
    mesh = s1.mesh
    ref = O.ElasticOracle(mesh.connect, mesh.coord, "QUAD4", 210000.0, 0.3, True, 13.0,
                          {"SEG2": mesh.dict_groupElem[ElemType.SEG2].connect})
    ref.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: x == 0), [0, 0], ["x", "y"])
    ref.add_surfLoad(mesh.Nodes_Conditions(lambda x, y, z: x == 120.0), [-800.0 / 13.0 ** 2], ["y"])
    err = _rel(u1, ref.Solve())
    print(f"cantilever, FP32 values: pipelined {i1.iterations} it, mode {i1.mode}, fallback {i1.fallback}, "
          f"vs direct {err:.2e}")
    assert i1.pipelined and i1.mode == "mixed" and not i1.fallback and i1.status == 1
    assert err < 1e-8


def test_pipelined_cube_matches_direct():
    from easyfea_amd.mesh import hexa8_cube

    simu = workloads.cube(12, 0, "hip_fp64")
    simu.preconditioner = "two-level"
    simu.coarse_target = 128
    simu.pipelined = True
    u = simu.Solve()
    info = simu.last_solve["elastic"]
    m = hexa8_cube(12)
    ref = O.ElasticOracle(m.connect, m.coord, "HEXA8", workloads.cube_E(m.Ne, 0), 0.3, False, 1.0,
                          {"QUAD4": m.dict_groupElem[ElemType.QUAD4].connect})
    ref.add_dirichlet(m.Nodes_Conditions(lambda x, y, z: z == 0), [0, 0, 0], ["x", "y", "z"])
    ref.add_surfLoad(m.Nodes_Conditions(lambda x, y, z: z == 1.0), [-10.0], ["z"])
    err = _rel(u, ref.Solve())
    print(f"12^3 cube pipelined: {info.iterations} it, vs direct {err:.2e}")
    assert info.pipelined and err < 1e-8


@pytest.mark.parametrize("solver", ["hip_fp64", "hip_mixed"])
def test_pipelined_cracked_body_staggered_matches_oracle(solver, _fp32_values):
    mesh = ea.tri3_notched_square(60)
    mat = Models.Elastic.Isotropic(2, 210000.0, 0.3, planeStress=False, thickness=1.0)
    simu = Simulations.PhaseField(mesh, Models.PhaseField(mat, "Miehe", "AT2", 2.7, 0.02))
    simu.solver = solver
    simu.pipelined = True
    ref = O.PhaseFieldOracle(mesh.connect, mesh.coord, "TRI3", 210000.0, 0.3, False, 1.0, "Miehe", "AT2", 2.7, 0.02)
    top = mesh.Nodes_Conditions(lambda x, y, z: y == 1.0)
    bot = mesh.Nodes_Conditions(lambda x, y, z: y == 0.0)
    rows = []
    for dep in [2e-3, 4e-3, 5.5e-3, 6.5e-3, 7.5e-3, 8.5e-3]:
        for s in (simu, ref):
            s.Bc_Init()
            s.add_dirichlet(top, [0, dep], ["x", "y"])
            s.add_dirichlet(bot, [0], ["y"])
        u, d, _ = simu.Solve(1.0, 10, 1)
        ur, dr, _ = ref.Solve(1.0, 10, 1)
        simu.Save_Iter()
        ref.Save_Iter()
        i = simu.last_solve["elastic"]
        rows.append((dep, i.pipelined, i.iterations, _rel(u, ur), _rel(d, dr), float(dr.max()), i.mode, i.fallback))
    print(f"{solver}: dep, pipelined, iterations, u err, d err, max d, mode, fallback:", rows)
    for r in rows:
        assert r[1] and r[3] < 1e-8 and r[4] < 1e-8 and not r[7], rows
        assert r[6] == ("mixed" if solver == "hip_mixed" else "fp64"), rows
