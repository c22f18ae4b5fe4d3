"""The BoundConstrain damage solve of the oracle (``bounded_lsq``, the
reference's ``lsq_linear`` call) pinned by brute force and by the
first-order conditions (CPU tests).

* tiny systems: every assignment of each variable to free / lower / upper
  is solved exactly and the feasible one of least residual kept;
* loose bounds: the unconstrained solution ``A^-1 b``;
* a damage-like sparse system: the KKT conditions of the exact variant
  (``bounded_lsq_exact``) for
  ``min 1/2 |A x - b|^2`` (gradient ``A (A x - b)`` zero on free variables,
  >= 0 at the lower bound, <= 0 at the upper bound).

* an AT1 damage system of 637 unknowns (increment 13 of
  ``tests/golden/make_at1_broken_history.py``) where BVLS alone stops 8 %
  above the minimum cost: the exact variant reaches the cost of a
  trust-region solve run to 1e-15 and meets the first-order conditions.

The brute force also shows that the energy minimiser
``min 1/2 x.A.x - b.x`` over the same box (a plausible mistake) is a
different point when bounds are active.
"""

import itertools
import os

import numpy as np
import scipy.sparse as sp

from oracle import efea_oracle as O


def _spd(n, rng):
    Q = rng.standard_normal((n, n))
    return Q @ Q.T + n * np.eye(n)


def _brute(A, b, lb, ub, objective):
    n = len(b)
    best, xbest = np.inf, None
    for pat in itertools.product((0, 1, 2), repeat=n):
        x = np.where(np.array(pat) == 1, lb, np.where(np.array(pat) == 2, ub, 0.0))
        F = np.array(pat) == 0
        if F.any():
            if objective == "lsq":
                r = b - A[:, ~F] @ x[~F]
                x[F] = np.linalg.lstsq(A[:, F], r, rcond=None)[0]
            else:  # energy: A_FF x_F = b_F - A_FC x_C
                x[F] = np.linalg.solve(A[np.ix_(F, F)], b[F] - A[np.ix_(F, ~F)] @ x[~F])
        if np.any(x < lb - 1e-12) or np.any(x > ub + 1e-12):
            continue
        f = 0.5 * np.sum((A @ x - b) ** 2) if objective == "lsq" else 0.5 * x @ A @ x - b @ x
        if f < best - 1e-14:
            best, xbest = f, x
    return xbest


def test_bounded_lsq_matches_brute_force():
    rng = np.random.default_rng(4)
    differs = 0
    for trial in range(6):
        n = 5
        A = _spd(n, rng)
        b = 3 * n * rng.standard_normal(n)
        lb = rng.uniform(-1.0, 0.0, n)
        ub = lb + rng.uniform(0.2, 1.5, n)
        x = O.bounded_lsq_exact(sp.csr_matrix(A), b, lb, ub)
        xb = _brute(A, b, lb, ub, "lsq")
        assert np.allclose(x, xb, atol=1e-7), (trial, x, xb)
        xe = _brute(A, b, lb, ub, "energy")
        differs += not np.allclose(xe, xb, atol=1e-6)
    assert differs >= 3  # the energy minimiser is another point


def test_bounded_lsq_with_loose_bounds_is_the_solve():
    rng = np.random.default_rng(5)
    A = _spd(12, rng)
    b = rng.standard_normal(12)
    x = O.bounded_lsq(sp.csr_matrix(A), b, np.full(12, -1e3), np.full(12, 1e3))
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)


def test_bounded_lsq_first_order_conditions_on_a_damage_system():
    # 1D AT1-like damage operator: k * Laplacian + r * mass, source in the middle
    n, h, k = 200, 1.0 / 199, 0.75 * 2.7 * 0.02
    main = np.full(n, 2.0)
    main[[0, -1]] = 1.0
    L = sp.diags([main, -np.ones(n - 1), -np.ones(n - 1)], [0, 1, -1]) / h
    r = 2 * np.where(np.abs(np.linspace(0, 1, n) - 0.5) < 0.1, 40.0, 0.5)
    M = sp.diags(r * h)
    A = sp.csr_matrix(k * L + M)
    f = np.maximum(2 * r / 2 * 1.2 - 3 * 2.7 / (8 * 0.02), 0) * h
    s = np.abs(np.linspace(0, 1, n) - 0.5)
    lb = np.where(s < 0.05, 0.6, np.where(s > 0.4, 0.2, 0.0))  # above the free solution in both zones
    ub = np.ones(n)
    x = O.bounded_lsq_exact(A, f, lb, ub)
    g = A @ (A @ x - f)
    scale = np.abs(A @ f).max()
    at_lb = x <= lb + 1e-9
    at_ub = x >= ub - 1e-9
    free = ~(at_lb | at_ub)
    assert at_lb.any() and free.any()
    assert np.all(x >= lb - 1e-12) and np.all(x <= ub + 1e-12)
    assert np.abs(g[free]).max() <= 1e-6 * scale
    assert g[at_lb].min() >= -1e-6 * scale
    assert (not at_ub.any()) or g[at_ub].max() <= 1e-6 * scale
    # the reference's trust-region call stops short of it: feasible, its cost
    # above the minimum (1.4e-3 relative here)
    xr = O.bounded_lsq(A, f, lb, ub)
    cost = lambda y: 0.5 * np.sum((A @ y - f) ** 2)  # noqa: E731
    assert np.all(xr >= lb - 1e-12) and np.all(xr <= ub + 1e-12)
    assert cost(x) <= cost(xr) <= cost(x) * (1 + 1e-2)


def test_bounded_lsq_exact_where_bvls_stops_short():
    from scipy import optimize

    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "at1_damage_system_13.npz"))
    A = sp.csr_matrix((z["data"], z["indices"], z["indptr"]), shape=tuple(z["shape"]))
    b, lb, ub = z["b"], z["lb"], z["ub"]
    cost = lambda y: 0.5 * np.sum((A @ y - b) ** 2)  # noqa: E731
    x = O.bounded_lsq_exact(A, b, lb, ub)
    Ad = A.toarray()
    xb = optimize.lsq_linear(Ad, b, bounds=(lb, ub), method="bvls", tol=1e-14).x
    xt = optimize.lsq_linear(Ad, b, bounds=(lb, ub), method="trf", tol=1e-15, max_iter=10000).x
    assert np.all(x >= lb) and np.all(x <= ub)
    assert O._lsq_kkt(Ad.T @ Ad, Ad.T @ b, lb, ub, x) <= 1e-11
    assert cost(x) <= cost(xt) * (1 + 1e-9)
    assert cost(xb) > 1.05 * cost(x)  # BVLS alone: 8 % above
    assert np.abs(x - xt).max() < 1e-4
