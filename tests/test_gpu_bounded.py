"""BoundConstrain on the device (GPU tests).

``efea_solve_bounded_lsq`` returns the minimiser of ``|A x - b|^2`` over the
box; the oracle's ``bounded_lsq_exact`` (scipy's bounded-variable
least-squares active-set method on the dense matrix) gives the same point.
The reference's own call (``lsq_linear``, trust-region, tol 1e-10) stops
short of it on ill-conditioned systems, so parity is with the minimiser:
x within 1e-7 (damage units), the first-order conditions within 1e-9, and
the reference call's cost never below the device's.
"""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import easyfea_amd as ea  # noqa: E402
from easyfea_amd import Models, Simulations, _lib  # noqa: E402
from easyfea_amd.device import ptr, stream_ptr, to_dev  # noqa: E402
from oracle import efea_oracle as O  # noqa: E402


def _device_solve(A, b, lb, ub, x0=None, tol=1e-10, kkt=1e-9, max_outer=200):
    A = sp.csr_matrix(A)
    A.sort_indices()
    N = A.shape[0]
    rp = to_dev(A.indptr.astype(np.int32))
    ci = to_dev(A.indices.astype(np.int32))
    va = to_dev(A.data.astype(np.float64))
    x = to_dev(np.zeros(N) if x0 is None else np.asarray(x0, float))
    nbytes = int(_lib.load().efea_bounded_lsq_workspace_bytes(N))
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    info = np.zeros(4)
    db, dl, du = to_dev(b), to_dev(lb), to_dev(ub)  # kept alive through the call
    print(f"efea_solve_bounded_lsq N={N}", flush=True)
    rc = _lib.call("efea_solve_bounded_lsq", N, ptr(rp), ptr(ci), ptr(va), ptr(db), ptr(dl), ptr(du), ptr(x),
                   ptr(ws), nbytes, tol, kkt, 100000, max_outer, info.ctypes.data, stream_ptr())
    torch.cuda.synchronize()
    return rc, x.cpu().numpy(), info


def _damage_1d(n=200):
    h, k = 1.0 / (n - 1), 0.75 * 2.7 * 0.02
    main = np.full(n, 2.0)
    main[[0, -1]] = 1.0
    L = sp.diags([main, -np.ones(n - 1), -np.ones(n - 1)], [0, 1, -1]) / h
    s = np.abs(np.linspace(0, 1, n) - 0.5)
    r = 2 * np.where(s < 0.1, 40.0, 0.5)
    A = sp.csr_matrix(k * L + sp.diags(r * h))
    f = np.maximum(1.2 * r - 3 * 2.7 / (8 * 0.02), 0) * h
    lb = np.where(s < 0.05, 0.6, np.where(s > 0.4, 0.2, 0.0))
    return A, f, lb, np.ones(n)


def _damage_2d(seed=0):
    """Q1 Laplacian + lumped reaction on a 24x24 QUAD4 square, random bounds."""
    m = ea.quad4_rectangle(1.0, 1.0, 24, 24)
    rng = np.random.default_rng(seed)
    Nn = m.Nn
    Ke = np.array([[4, -1, -2, -1], [-1, 4, -1, -2], [-2, -1, 4, -1], [-1, -2, -1, 4]]) / 6.0
    rows = np.repeat(m.connect, 4, axis=1).ravel()
    cols = np.tile(m.connect, (1, 4)).ravel()
    K = sp.csr_matrix((np.tile(Ke.ravel(), m.Ne), (rows, cols)), shape=(Nn, Nn))
    A = sp.csr_matrix(0.05 * K + sp.diags(rng.uniform(0.01, 2.0, Nn) / Nn))
    f = rng.uniform(-0.5, 1.5, Nn) / Nn
    lb = np.where(rng.random(Nn) < 0.3, rng.uniform(0.0, 0.5, Nn), 0.0)
    return A, f, lb, np.ones(Nn)


@pytest.mark.parametrize("case", ["1d", "2d-0", "2d-1", "loose", "all-upper"])
def test_bounded_lsq_is_the_minimiser(case):
    if case == "1d":
        A, b, lb, ub = _damage_1d()
    elif case.startswith("2d"):
        A, b, lb, ub = _damage_2d(int(case[-1]))
    elif case == "loose":
        A, b, lb, ub = _damage_2d(2)
        lb, ub = np.full_like(lb, -1e3), np.full_like(ub, 1e3)
    else:  # a load pushing every variable past 1
        A, _, lb, ub = _damage_2d(3)
        b = A @ np.full(A.shape[0], 2.0)
    rc, x, info = _device_solve(A, b, lb, ub)
    assert rc == 0, (_lib.last_error(), info)
    xe = np.linalg.solve(A.toarray(), b) if case == "loose" else O.bounded_lsq_exact(A, b, lb, ub)
    assert np.abs(x - xe).max() < 1e-7, (case, np.abs(x - xe).max(), info)
    assert np.all(x >= lb) and np.all(x <= ub)
    if case == "all-upper":
        assert (x == ub).any()
    g = A @ (A @ x - b)
    d = np.asarray(A.multiply(A).sum(axis=1)).ravel()
    assert np.abs(x - np.clip(x - g / d, lb, ub)).max() <= 1e-9
    xr = O.bounded_lsq(A, b, lb, ub)  # the reference's call
    cost = lambda y: 0.5 * np.sum((A @ y - b) ** 2)  # noqa: E731
    assert cost(x) <= cost(xr) * (1 + 1e-9)
    # bitwise repeatable, from any start
    rc2, x2, _ = _device_solve(A, b, lb, ub)
    assert np.array_equal(x, x2)
    rc3, x3, _ = _device_solve(A, b, lb, ub, x0=np.clip(xe + 0.1, lb, ub))
    assert np.abs(x3 - xe).max() < 1e-7


def test_bounded_lsq_fine_1d_band():
    """A finer 1D damage system (n = 800, cond(A) ~ 6.6e4) whose damage band
    needs about 270 active-set updates: with the update cap raised, x within
    1e-7 of the minimiser and inside the bounds exactly."""
    A, b, lb, ub = _damage_1d(800)
    rc, x, info = _device_solve(A, b, lb, ub, max_outer=2000)
    assert rc == 0, (_lib.last_error(), info)
    xe = O.bounded_lsq_exact(A, b, lb, ub)
    print(f"n=800: {int(info[0])} active-set updates, {int(info[1])} CGLS iterations, "
          f"max |x - x*| {np.abs(x - xe).max():.2e}")
    assert np.abs(x - xe).max() < 1e-7
    assert np.all(x >= lb) and np.all(x <= ub)


def test_bounded_lsq_reports_non_convergence():
    A, b, lb, ub = _damage_1d()
    N = A.shape[0]
    rp, ci, va = (to_dev(A.indptr.astype(np.int32)), to_dev(A.indices.astype(np.int32)), to_dev(A.data))
    x = to_dev(np.zeros(N))
    nbytes = int(_lib.load().efea_bounded_lsq_workspace_bytes(N))
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    db, dl, du = to_dev(b), to_dev(lb), to_dev(ub)
    # one CGLS block per pass and one active-set update cannot reach the minimiser
    rc = _lib.call("efea_solve_bounded_lsq", N, ptr(rp), ptr(ci), ptr(va), ptr(db), ptr(dl), ptr(du), ptr(x),
                   ptr(ws), nbytes, 1e-10, 1e-12, 1, 1, None, stream_ptr())
    assert rc == _lib.ERR_NOCONV and "violate" in _lib.last_error()


def test_boundconstrain_staggered_matches_oracle():
    """AT1 BoundConstrain tension steps of a small notched square: u and d
    within 1e-6 of the oracle's staggered loop with the exact minimiser."""
    n = 16
    mesh = ea.tri3_notched_square(n)
    mat = Models.Elastic.Isotropic(2, 210000.0, 0.3, planeStress=False)
    pfm = Models.PhaseField(mat, "Amor", "AT1", 2.7, 0.08, solver="BoundConstrain")
    simu = Simulations.PhaseField(mesh, pfm)
    simu.solver = "hip_fp64"
    ref = O.PhaseFieldOracle(mesh.connect, mesh.coord, "TRI3", 210000.0, 0.3, False, 1.0, "Amor", "AT1", 2.7, 0.08,
                             solver="BoundConstrain")
    top = mesh.Nodes_Conditions(lambda x, y, z: y == 1.0)
    bot = mesh.Nodes_Conditions(lambda x, y, z: y == 0.0)
    errs = []
    # four steps: the fifth breaks the specimen through, after which the
    # displacement is fixed only by the 1e-12 residual stiffness
    for it in range(1, 5):
        dep = it * 6e-3
        for s in (simu, ref):
            s.Bc_Init()
            s.add_dirichlet(top, [0, dep], ["x", "y"])
            s.add_dirichlet(bot, [0], ["y"])
        u, d, _ = simu.Solve(1.0, 10, 1)
        ur, dr, _ = ref.Solve(1.0, 10, 1)
        simu.Save_Iter()
        ref.Save_Iter()
        errs.append((it, float(np.abs(u - ur).max() / np.abs(ur).max()), float(np.abs(d - dr).max()),
                     float(d.max()), simu.last_bounded))
    print("BoundConstrain steps (it, u, d, dmax, info):", errs)
    assert errs[-1][3] > 0.5  # the damage has grown
    for it, eu, ed, _, _ in errs:
        assert eu < 1e-6 and ed < 1e-6, errs
