"""Pins the oracle's time schemes, mass / capacity operators and thermal
simulation against closed forms and invariants (CPU only).

* element mass matrices: the textbook consistent masses of a straight TRI3,
  a rectangular QUAD4 and a TETRA4, total mass on curved/distorted meshes,
  the dof_n = dim matrix as kron(M, I);
* steady conduction: a linear temperature is reproduced exactly on distorted
  meshes and the boundary flux equals k dT/L times the section (no thickness
  in the conduction matrix);
* Newmark average acceleration on a structural mode: the exact rotation by
  2 atan(omega dt / 2) per step, undamped and (with Rayleigh damping) the
  Cayley transform of the first-order system;
* backward / forward Euler on a damped mode against the first-order Euler
  steps; the parabolic scheme on a thermal mode against its amplification
  factor;
* reductions written in the reference's docstrings: hht(alpha = 0) and
  hht_newmark(alpha = 0) are newmark, hht(alpha = 1/2) is the midpoint rule;
* every scheme satisfies K u_t + C v_t + M a_t = F at its evaluation states
  (a formula set of its own in the reference) at the free DOFs.
"""

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

import easyfea_amd as ea
from oracle import efea_oracle as O

ALGOS = ["newmark", "midpoint", "hht", "hht_newmark", "euler_implicit", "euler_explicit"]


def _distort(m, amp=0.04, seed=1):
    rng = np.random.default_rng(seed)
    c = m.coord.copy()
    dim = m.dim
    inner = np.ones(m.Nn, bool)
    for k in range(dim):
        inner &= (c[:, k] > c[:, k].min() + 1e-12) & (c[:, k] < c[:, k].max() - 1e-12)
    c[inner, :dim] += amp * (rng.random((inner.sum(), dim)) - 0.5)
    return c


# ------------------------------------------------------------- mass operator
def test_element_mass_textbook_matrices():
    # straight TRI3 (area A): A/12 [[2,1,1],[1,2,1],[1,1,2]]
    X = np.array([[0.0, 0, 0], [2.0, 0.3, 0], [0.4, 1.5, 0]])
    area = 0.5 * abs(2.0 * 1.5 - 0.3 * 0.4)
    g = O.Geometry("TRI3", np.array([[0, 1, 2]]), X, "mass", 2)
    assert np.allclose(O.UV(g, 3.0)[0], 3.0 * area / 12 * (np.ones((3, 3)) + np.eye(3)), rtol=1e-14, atol=0)
    # rectangular QUAD4 a x b (nodes counter-clockwise): ab/36 [[4,2,1,2],...]
    a, b = 1.5, 0.8
    X = np.array([[0.0, 0, 0], [a, 0, 0], [a, b, 0], [0, b, 0]])
    g = O.Geometry("QUAD4", np.array([[0, 1, 2, 3]]), X, "mass", 2)
    ref = a * b / 36 * np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]])
    assert np.allclose(O.UV(g, 1.0)[0], ref, rtol=1e-14, atol=0)
    # TETRA4 (volume V): V/20 (1 + delta_ab)
    X = np.array([[0.0, 0, 0], [1.2, 0, 0], [0.1, 0.9, 0], [0.2, 0.3, 1.1]])
    V = abs(np.linalg.det(X[1:] - X[0])) / 6
    g = O.Geometry("TETRA4", np.array([[0, 1, 2, 3]]), X, "mass", 3)
    assert np.allclose(O.UV(g, 1.0)[0], V / 20 * (np.ones((4, 4)) + np.eye(4)), rtol=1e-14, atol=0)


@pytest.mark.parametrize("et", ["TRI3", "QUAD4", "HEXA8", "TETRA4", "TETRA10"])
def test_mass_total_and_vector_structure(et):
    makers = {
        "TRI3": lambda: ea.tri3_rectangle(2.0, 1.5, 5, 4),
        "QUAD4": lambda: ea.quad4_rectangle(2.0, 1.5, 5, 4),
        "HEXA8": lambda: ea.hexa8_box(2.0, 1.5, 1.0, 3, 3, 2),
        "TETRA4": lambda: ea.mesh.tetra4_box(2.0, 1.5, 1.0, 3, 2, 2),
        "TETRA10": lambda: ea.mesh.tetra10_from_tetra4(ea.mesh.tetra4_box(2.0, 1.5, 1.0, 2, 2, 2)),
    }
    m = makers[et]()
    dim = m.dim
    # interior moves keep the outer measure (straight-sided TETRA10 only: its
    # 4-point rule does not integrate the Jacobian of curved edges exactly)
    c = _distort(m, 0.0 if et == "TETRA10" else 0.05)
    g = O.Geometry(et, m.connect, c, "mass", dim)
    rng = np.random.default_rng(3)
    rho = 1 + rng.random(m.Ne)
    Ms = O.UV(g, rho)
    # sum_ab M_ab = int rho; the element measures tile the box
    vol_e = g.wJ.sum(1)
    assert abs(vol_e.sum() - 3.0) < 1e-12
    assert np.allclose(Ms.sum((1, 2)), rho * vol_e, rtol=1e-13, atol=0)
    Mv = O.UV(g, rho, dof_n=dim)
    assert np.allclose(Mv, np.stack([np.kron(M, np.eye(dim)) for M in Ms]), rtol=1e-14, atol=0)


# ------------------------------------------------------------- thermal steady
@pytest.mark.parametrize("et", ["TRI3", "QUAD4", "HEXA8", "TETRA4", "TETRA10"])
def test_thermal_steady_linear_field_and_flux(et):
    L, H, W = 2.0, 1.5, 1.0
    makers = {
        "TRI3": lambda: ea.tri3_rectangle(L, H, 6, 4),
        "QUAD4": lambda: ea.quad4_rectangle(L, H, 6, 4),
        "HEXA8": lambda: ea.hexa8_box(L, H, W, 4, 3, 2),
        "TETRA4": lambda: ea.mesh.tetra4_box(L, H, W, 3, 2, 2),
        "TETRA10": lambda: ea.mesh.tetra10_from_tetra4(ea.mesh.tetra4_box(L, H, W, 2, 2, 2)),
    }
    m = makers[et]()
    c = _distort(m, 0.0 if et == "TETRA10" else 0.04)  # the TETRA10 patch test holds on straight edges
    k, th = 1.7, 3.0
    sim = O.ThermalOracle(m.connect, c, et, k, 0.0, thickness=th)
    x0 = np.nonzero(np.abs(c[:, 0]) < 1e-12)[0]
    xL = np.nonzero(np.abs(c[:, 0] - L) < 1e-12)[0]
    sim.add_dirichlet(x0, [0.0], ["t"], "thermal")
    sim.add_dirichlet(xL, [40.0], ["t"], "thermal")
    t = sim.Step()
    assert np.max(np.abs(t - 40.0 * c[:, 0] / L)) < 1e-11
    K, _ = sim.Get_K_C()
    section = H if m.dim == 2 else H * W  # no thickness on K
    assert abs(np.sum((K @ t)[xL]) - k * 40.0 / L * section) < 1e-10


# ------------------------------------------------------------- modal closed forms
def _structure(rayleigh=(0.0, 0.0), et="QUAD4"):
    m = ea.quad4_rectangle(2.0, 1.0, 6, 3) if et == "QUAD4" else ea.tri3_rectangle(2.0, 1.0, 6, 3)
    c = _distort(m, 0.03, seed=4)
    sim = O.DynamicElasticOracle(m.connect, c, et, 200.0, 0.3, True, 0.5, rho=2.0, rayleigh=rayleigh)
    clamp = np.nonzero(np.abs(c[:, 0]) < 1e-12)[0]
    sim.add_dirichlet(clamp, [0.0, 0.0], ["x", "y"])
    K, C, M = sim.Get_K_C_M()
    free = np.setdiff1d(np.arange(sim.Ndof), O.dofs_nodes(clamp, ["x", "y"], ["x", "y"]))
    w2, phi = sla.eigh(K[free][:, free].toarray(), M[free][:, free].toarray())
    return sim, free, w2, phi


def _mode(sim, free, phi, j):
    f = np.zeros(sim.Ndof)
    f[free] = phi[:, j]
    return f


def test_newmark_mode_rotates_by_the_trapezoidal_angle():
    sim, free, w2, phi = _structure()
    j = 2
    w = np.sqrt(w2[j])
    f = _mode(sim, free, phi, j)
    dt = 0.7 / w
    theta = 2 * np.arctan(w * dt / 2)
    sim.speed = f.copy()  # u0 = 0, a0 = 0 = -w^2 u0: consistent
    p = O.hyperbolic_params(dt)
    for n in range(1, 9):
        sim.Step("newmark", p)
        assert np.allclose(sim.u, f * np.sin(n * theta) / w, rtol=0, atol=1e-10 * np.abs(f).max() / w)
        assert np.allclose(sim.speed, f * np.cos(n * theta), rtol=0, atol=1e-10 * np.abs(f).max())
        assert np.allclose(sim.accel, -w2[j] * sim.u, rtol=0, atol=1e-9 * np.abs(f).max() * w)


def _damped(sim, free, w2, phi, j, rayleigh):
    cM, cK = rayleigh
    w = np.sqrt(w2[j])
    two_zeta_w = cM + cK * w2[j]
    J = np.array([[0.0, 1.0], [-w2[j], -two_zeta_w]])  # (u, v)' = J (u, v) for the mode
    return w, J, _mode(sim, free, phi, j)


def test_newmark_damped_mode_is_the_cayley_transform():
    ray = (0.3, 0.02)
    sim, free, w2, phi = _structure(ray)
    w, J, f = _damped(sim, free, w2, phi, 1, ray)
    dt = 0.5 / w
    I = np.eye(2)
    R = np.linalg.solve(I - dt / 2 * J, I + dt / 2 * J)
    y = np.array([0.0, 1.0])
    sim.speed = f.copy()
    sim.accel = (J @ y)[1] * f  # consistent initial acceleration
    for _ in range(6):
        y = R @ y
        sim.Step("newmark", O.hyperbolic_params(dt))
        assert np.allclose(sim.u, y[0] * f, rtol=0, atol=1e-10 * np.abs(f).max())
        assert np.allclose(sim.speed, y[1] * f, rtol=0, atol=1e-10 * np.abs(f).max())


@pytest.mark.parametrize("algo", ["euler_implicit", "euler_explicit"])
def test_euler_schemes_on_a_damped_mode(algo):
    ray = (0.2, 0.01)
    sim, free, w2, phi = _structure(ray)
    w, J, f = _damped(sim, free, w2, phi, 0, ray)
    dt = 0.3 / w if algo == "euler_implicit" else 0.05 / w
    y = np.array([0.4, 1.0])
    sim.u, sim.speed = y[0] * f, y[1] * f
    for _ in range(5):
        if algo == "euler_implicit":
            # u1 = u0 + dt v1, v1 = v0 + dt a1, a1 = (J y1)[1]
            A = np.array([[1.0, -dt], [-dt * J[1, 0], 1.0 - dt * J[1, 1]]])
            y = np.linalg.solve(A, y)
        else:
            a0 = (J @ y)[1]
            y = np.array([y[0] + dt * y[1], y[1] + dt * a0])
        sim.Step(algo, O.hyperbolic_params(dt, algo))
        assert np.allclose(sim.u, y[0] * f, rtol=0, atol=1e-10 * np.abs(f).max())
        assert np.allclose(sim.speed, y[1] * f, rtol=0, atol=1e-10 * np.abs(f).max())


@pytest.mark.parametrize("alpha", [0.5, 0.75, 1.0])
def test_parabolic_thermal_mode_amplification(alpha):
    m = ea.quad4_rectangle(2.0, 1.0, 6, 3)
    c = _distort(m, 0.03, seed=5)
    sim = O.ThermalOracle(m.connect, c, "QUAD4", 1.3, 0.8, rho=1.5)
    x0 = np.nonzero(np.abs(c[:, 0]) < 1e-12)[0]
    sim.add_dirichlet(x0, [0.0], ["t"], "thermal")
    K, C = sim.Get_K_C()
    free = np.setdiff1d(np.arange(m.Nn), x0)
    lam, phi = sla.eigh(K[free][:, free].toarray(), C[free][:, free].toarray())
    j = 3
    f = np.zeros(m.Nn)
    f[free] = phi[:, j]
    dt = 0.8 / lam[j]
    sim.u = f.copy()
    amp = 1 / (1 + alpha * lam[j] * dt)  # first step from v0 = 0
    r = (1 - (1 - alpha) * lam[j] * dt) / (1 + alpha * lam[j] * dt)
    for n in range(1, 7):
        sim.Step("parabolic", (dt, alpha))
        expect = amp * r ** (n - 1)
        assert np.allclose(sim.u, expect * f, rtol=0, atol=1e-11 * np.abs(f).max())
        assert np.allclose(sim.speed, -lam[j] * expect * f, rtol=0, atol=1e-10 * lam[j] * np.abs(f).max())


# ------------------------------------------------------------- reductions
def _loaded(ray=(0.1, 0.003)):
    sim, free, _, _ = _structure(ray)
    right = np.nonzero(np.abs(sim.coord[:, 0] - 2.0) < 1e-12)[0]
    sim.neu["elastic"] = (O.dofs_nodes(right, ["y"], ["x", "y"]), np.full(right.size, -0.4))
    return sim


def _run(algo, params, steps=5):
    sim = _loaded()
    for _ in range(steps):
        sim.Step(algo, params)
    return sim.u, sim.speed, sim.accel


@pytest.mark.parametrize("other,params", [
    ("hht", (0.05, 0.25, 0.5, 0.0)),
    ("hht_newmark", O.hyperbolic_params(0.05, "hht_newmark", alpha=0.0)),
])
def test_hht_variants_at_alpha_zero_are_newmark(other, params):
    ref = _run("newmark", (0.05, 0.25, 0.5, 0.0))
    got = _run(other, params)
    for r, g in zip(ref, got):
        assert np.allclose(g, r, rtol=1e-9, atol=1e-12 * np.abs(r).max())


def test_hht_at_one_half_is_the_midpoint_rule():
    ref = _run("midpoint", O.hyperbolic_params(0.05, "midpoint"))
    got = _run("hht", (0.05, 0.25, 0.5, 0.5))
    for r, g in zip(ref, got):
        assert np.allclose(g, r, rtol=1e-9, atol=1e-12 * np.abs(r).max())


@pytest.mark.parametrize("algo", ALGOS + ["parabolic"])
def test_every_scheme_balances_at_its_evaluation_states(algo):
    sim = _loaded()
    rng = np.random.default_rng(8)
    sim.u, sim.speed, sim.accel = (1e-2 * rng.standard_normal(sim.Ndof) for _ in range(3))
    dofs, _ = sim.dir["elastic"]
    sim.u[dofs] = sim.speed[dofs] = sim.accel[dofs] = 0.0
    params = (0.04, 0.6) if algo == "parabolic" else O.hyperbolic_params(0.04, algo, alpha=0.2)
    K, C, M = sim.Get_K_C_M()
    F = sim.neumann_vector("elastic", sim.Ndof)
    u_n, v_n, a_n = sim.u.copy(), sim.speed.copy(), sim.accel.copy()
    sim.Step(algo, params)
    x = sim.accel if algo == "euler_explicit" else sim.u
    u_t, v_t, a_t = O.time_evaluate(algo, params, u_n, v_n, a_n, x)
    r = K @ u_t + C @ v_t + (M @ (x if algo == "euler_explicit" else a_t) if algo != "parabolic" else 0.0) - F
    free = np.setdiff1d(np.arange(sim.Ndof), dofs)
    scale = np.abs(K @ u_t).max() + np.abs(F).max()
    assert np.abs(r[free]).max() < 1e-10 * scale, algo


def test_time_update_matches_a_plain_restatement_of_newmark():
    # the reference's Newmark corrector written from Hughes (1987) ch. 9
    rng = np.random.default_rng(2)
    u, v, a, u1 = rng.standard_normal((4, 50))
    dt, beta, gamma = 0.03, 0.3, 0.6
    a1 = (u1 - u - dt * v - dt ** 2 * (0.5 - beta) * a) / (beta * dt ** 2)
    v1 = v + dt * ((1 - gamma) * a + gamma * a1)
    _, gv, ga = O.time_update("newmark", (dt, beta, gamma, 0.0), u, v, a, u1)
    assert np.allclose(ga, a1, rtol=1e-10) and np.allclose(gv, v1, rtol=1e-10)
