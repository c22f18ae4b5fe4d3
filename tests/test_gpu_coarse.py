"""Two-level preconditioner (GPU tests): the device coarse matrix equals
P^T A P of the host prolongator, the device inverse equals numpy's, and
two-level solves meet the same bar as the one-level ones (u within 1e-8 of
the oracle's sparse direct solve) with fewer iterations."""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import easyfea_amd as ea  # noqa: E402
from easyfea_amd import Models, Simulations, _lib, workloads  # noqa: E402
from easyfea_amd.coarse import prolongator  # noqa: E402
from easyfea_amd.device import DeviceMesh, ptr, stream_ptr, to_dev  # noqa: E402
from oracle import efea_oracle as O  # noqa: E402


@pytest.fixture(autouse=True)
def _force_fp32_krylov(monkeypatch):
    import easyfea_amd.device as dv

    monkeypatch.setattr(dv, "MIXED_MIN_MATRIX_BYTES", 0)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("maker,dof_n,target,known_frac", [
    (lambda: ea.tri3_notched_square(24), 2, 48, 0.1),
    (lambda: ea.quad4_rectangle(3.0, 1.0, 30, 10), 2, 40, 0.1),
    (lambda: ea.hexa8_box(1.0, 1.0, 1.0, 6, 5, 4), 3, 40, 0.1),
    (lambda: ea.tri3_notched_square(24), 1, 48, 0.1),
    # > 32 Gauss-Jordan blocks (the three-launch step with a pre-scaled row
    # panel); small aggregates, so no known DOFs that could make modes dependent
    (lambda: ea.quad4_rectangle(4.0, 2.0, 80, 40), 2, 8, 0.0),
])
def test_coarse_matrix_and_inverse(maker, dof_n, target, known_frac):
    mesh = maker()
    dev = DeviceMesh(mesh)
    A = dev.csr_map(dof_n)
    rng = np.random.default_rng(5)
    # SPD values on the pattern: a random element-wise SPD operator
    Ne, nPe = mesh.connect.shape
    nd = nPe * dof_n
    G = rng.standard_normal((Ne, nd, nd))
    Ke = np.einsum("eij,ekj->eik", G, G) + nd * np.eye(nd)
    A.scatter(to_dev(Ke.ravel()))
    known = np.zeros(mesh.Nn * dof_n, dtype=np.uint8)
    known[rng.choice(known.size, int(known.size * known_frac), replace=False)] = 1
    cs = A.coarse(target)
    Ainv = cs.setup(A, to_dev(known))
    K = A.to_scipy()
    P = prolongator(cs.ag, mesh.dim, dof_n, known.astype(bool))
    Ac = (P.T @ K @ P).toarray()
    nc = Ac.shape[0]
    d0 = np.diag(Ac) == 0
    Ac[d0, d0] = 1.0
    X = Ainv.cpu().numpy().reshape(cs.ld, cs.ld)[:nc, :nc]
    assert np.linalg.cond(Ac) < 1e10
    if known_frac == 0.0:
        assert cs.ld // 32 > 32, "case meant to exercise the three-launch Gauss-Jordan step"
    ref = np.linalg.inv(Ac)
    assert np.abs(X - ref).max() <= 1e-9 * np.abs(ref).max()
    assert np.array_equal(X, X.T)


def _cantilever(nx, ny, solver, precond):
    L, h = 120.0, 13.0
    mesh = ea.quad4_rectangle(L, h, nx, ny)
    simu = Simulations.Elastic(mesh, Models.Elastic.Isotropic(2, 210000.0, 0.3, True, h))
    simu.solver = solver
    simu.preconditioner = precond
    n0 = mesh.Nodes_Conditions(lambda x, y, z: x == 0)
    nL = mesh.Nodes_Conditions(lambda x, y, z: x == L)
    simu.add_dirichlet(n0, [0, 0], ["x", "y"])
    simu.add_surfLoad(nL, [-800 / h**2], ["y"])
    ref = O.ElasticOracle(mesh.connect, mesh.coord, "QUAD4", 210000.0, 0.3, True, h,
                          {"SEG2": mesh.dict_groupElem[ea.ElemType.SEG2].connect})
    ref.add_dirichlet(n0, [0, 0], ["x", "y"])
    ref.add_surfLoad(nL, [-800 / h**2], ["y"])
    return simu, ref


@pytest.mark.parametrize("solver", ["hip_fp64", "hip_mixed"])
def test_two_level_cantilever(solver):
    simu, ref = _cantilever(360, 39, solver, "two-level")
    u = simu.Solve()
    info = simu.last_solve["elastic"]
    uref = ref.Solve()
    print(f"two-level {solver}: {info.iterations} iterations")
    assert info.precond == "two-level"
    assert _rel(u, uref) < 1e-8
    simu1, _ = _cantilever(360, 39, solver, "bj")
    simu1.Solve()
    assert info.iterations < simu1.last_solve["elastic"].iterations / 3


@pytest.mark.parametrize("solver", ["hip_fp64", "hip_mixed"])
def test_two_level_hexa8(solver):
    mesh = ea.hexa8_cube(10)
    rng = np.random.default_rng(0)
    s = np.sqrt(np.log(1.04))
    E = np.exp(np.log(210000.0) - s**2 / 2 + s * rng.standard_normal(mesh.Ne))
    simu = Simulations.Elastic(mesh, Models.Elastic.Isotropic(3, E, 0.3))
    simu.solver = solver
    simu.preconditioner = "two-level"
    bot = mesh.Nodes_Conditions(lambda x, y, z: z == 0)
    top = mesh.Nodes_Conditions(lambda x, y, z: z == 1)
    simu.add_dirichlet(bot, [0, 0, 0], ["x", "y", "z"])
    simu.add_surfLoad(top, [-10.0], ["z"])
    u = simu.Solve()
    ref = O.ElasticOracle(mesh.connect, mesh.coord, "HEXA8", E, 0.3, False, 1.0,
                          {"QUAD4": mesh.dict_groupElem[ea.ElemType.QUAD4].connect})
    ref.add_dirichlet(bot, [0, 0, 0], ["x", "y", "z"])
    ref.add_surfLoad(top, [-10.0], ["z"])
    assert _rel(u, ref.Solve()) < 1e-8
    assert simu.last_solve["elastic"].precond == "two-level"


@pytest.mark.parametrize("solver", ["hip_fp64", "hip_mixed"])
def test_two_level_staggered_matches_oracle(solver):
    """Five load steps of the notched-square tension test (Miehe, AT2) with
    the two-level preconditioner on the displacement solves."""
    mesh = ea.tri3_notched_square(40)
    mat = Models.Elastic.Isotropic(2, 210000.0, 0.3, planeStress=False, thickness=1.0)
    simu = Simulations.PhaseField(mesh, Models.PhaseField(mat, "Miehe", "AT2", 2.7, 0.05))
    simu.solver = solver
    simu.preconditioner = "two-level"
    ref = O.PhaseFieldOracle(mesh.connect, mesh.coord, "TRI3", 210000.0, 0.3, False, 1.0, "Miehe", "AT2", 2.7, 0.05)
    top = mesh.Nodes_Conditions(lambda x, y, z: y == 1.0)
    bot = mesh.Nodes_Conditions(lambda x, y, z: y == 0.0)
    dofsY = simu.Bc_dofs_nodes(top, ["y"])
    for it in range(1, 6):
        for s_ in (simu, ref):
            s_.Bc_Init()
            s_.add_dirichlet(top, [0, it * 2e-3], ["x", "y"])
            s_.add_dirichlet(bot, [0], ["y"])
        u, d, _ = simu.Solve(1.0, 500, 1)
        ur, dr, _ = ref.Solve(1.0, 500, 1)
        simu.Save_Iter()
        ref.Save_Iter()
        assert _rel(u, ur) < 1e-8 and _rel(d, dr) < 1e-8
        f, fr = np.sum(simu.Calc_Reaction(dofsY, "elastic")), np.sum(ref.Calc_Reaction(dofsY))
        assert abs(f - fr) <= 1e-6 * abs(fr)
        assert simu.last_solve["elastic"].precond == "two-level"


def test_fp32_split_keeps_block_row_sums():
    """efea_split_f32: FP32 values are the rounded FP64 ones and K32 + corr32
    has K64's block row sums to FP32 accuracy of the (tiny) correction."""
    mesh = ea.tri3_notched_square(30)
    dev = DeviceMesh(mesh)
    A = dev.csr_map(2)
    rng = np.random.default_rng(3)
    Ne, nPe = mesh.connect.shape
    G = rng.standard_normal((Ne, 6, 6))
    A.scatter(to_dev((np.einsum("eij,ekj->eik", G, G) * 1e5).ravel()))
    A.make_fp32()
    v64 = A.values.cpu().numpy()
    v32 = A.values32.cpu().numpy()
    assert np.array_equal(v32, v64.astype(np.float32))
    corr = A.corr32.cpu().numpy().reshape(mesh.Nn, 2, 2)
    K64 = A.to_scipy().tocoo()
    d = (K64.data - K64.data.astype(np.float32).astype(np.float64))
    ref = np.zeros((mesh.Nn, 2, 2))
    np.add.at(ref, (K64.row // 2, K64.row % 2, K64.col % 2), d)
    assert np.abs(corr - ref).max() <= 1e-6 * np.abs(ref).max() + 1e-30


def test_mixed_two_level_on_cracked_body_reaches_fp64_accuracy():
    """The notched square's displacement system (194 x 194 cells, first
    increment): FP32 values without the row-sum correction stall near 5e-7
    error and then diverge; with it the mixed solve matches FP64 to 1e-8."""
    from easyfea_amd.device import SolveJob

    mesh = ea.tri3_notched_square(194)
    mat = Models.Elastic.Isotropic(2, 210000.0, 0.3, planeStress=False, thickness=1.0)
    out = {}
    for solver in ("hip_fp64", "hip_mixed"):
        simu = Simulations.PhaseField(mesh, Models.PhaseField(mat, "Amor", "AT2", 2.7, 0.01))
        simu.solver = solver
        simu.preconditioner = "two-level"
        top = mesh.Nodes_Conditions(lambda x, y, z: y == 1.0)
        bot = mesh.Nodes_Conditions(lambda x, y, z: y == 0.0)
        simu.add_dirichlet(bot, [0], ["y"])
        simu.add_dirichlet(top, [0, 1e-4], ["x", "y"])
        simu.Solve(1.0, 500, 1)
        out[solver] = (simu.displacement, simu.last_solve["elastic"])
    u64, i64 = out["hip_fp64"]
    u32, i32 = out["hip_mixed"]
    print(f"fp64 {i64.iterations} it, mixed {i32.iterations} it ({i32.mode}), rel res {i32.rel_residual:.1e}")
    assert i32.mode == "mixed" and not i32.fallback and i32.status == 1
    assert i32.iterations <= 1.3 * i64.iterations
    assert _rel(u32, u64) < 1e-8
