"""The oracle's anisotropic law (``C_anisotropic``, ``rotation_mandel``;
``Models/Elastic/_laws.py:1111-1259``, ``Models/_utils.py:433-660``), CPU
tests.  Pinned by what the mathematics fixes: the Kelvin-Mandel rotation is
orthogonal; it maps the Kelvin-Mandel vector of a strain tensor to the vector
of the tensor rotated directly (eps_g = p eps_m p^T, p = [axis_1 axis_2
axis_3]), which fixes every entry of P (a transposed p or a misplaced
direction cosine fails); the strain energy is invariant; an isotropic C is
unchanged by any rotation; a quarter turn about z swaps the xx and yy
stiffnesses; a 6 x 6 input enters as given whatever useVoigtNotation says
(the reference's behaviour)."""

import numpy as np

from oracle import efea_oracle as O


def _axes(rng):
    a1 = rng.standard_normal(3)
    a1 /= np.linalg.norm(a1)
    a2 = np.cross(a1, rng.standard_normal(3))
    a2 /= np.linalg.norm(a2)
    return a1, a2


def _km(t):
    r2 = np.sqrt(2)
    return np.array([t[0, 0], t[1, 1], t[2, 2], r2 * t[1, 2], r2 * t[0, 2], r2 * t[0, 1]])


def _spd(rng):
    Q = rng.standard_normal((6, 6))
    return Q @ Q.T + 6 * np.eye(6)


def test_rotation_is_orthogonal_and_rotates_strain_vectors():
    rng = np.random.default_rng(0)
    for _ in range(20):
        a1, a2 = _axes(rng)
        P = O.rotation_mandel(a1, a2)
        assert np.abs(P @ P.T - np.eye(6)).max() < 1e-14
        p = np.stack([a1, a2, np.cross(a1, a2)], axis=1)
        S = rng.standard_normal((3, 3))
        eps_m = S + S.T
        assert np.abs(P @ _km(eps_m) - _km(p @ eps_m @ p.T)).max() < 1e-13


def test_energy_invariance_and_isotropic_invariance():
    rng = np.random.default_rng(1)
    Ci = O.C_isotropic(210000.0, 0.3, 3, False)
    for _ in range(10):
        a1, a2 = _axes(rng)
        Cm = _spd(rng)
        Cg = O.C_anisotropic(Cm, False, a1, a2)
        P = O.rotation_mandel(a1, a2)
        e = rng.standard_normal(6)
        assert abs((P @ e) @ Cg @ (P @ e) - e @ Cm @ e) < 1e-10 * abs(e @ Cm @ e)
        assert np.abs(O.C_anisotropic(Ci, True, a1, a2) - Ci).max() < 1e-10 * np.abs(Ci).max()


def test_quarter_turn_swaps_in_plane_stiffnesses():
    Cm = np.diag([200.0, 50.0, 80.0, 30.0, 20.0, 10.0])
    Cg = O.C_anisotropic(Cm, False, (0, 1, 0), (-1, 0, 0))
    assert np.allclose(np.diag(Cg)[:3], [50.0, 200.0, 80.0])
    assert np.allclose(np.diag(Cg)[3:], [20.0, 30.0, 10.0])


def test_six_by_six_input_enters_as_given():
    rng = np.random.default_rng(2)
    Cm = _spd(rng)
    a1, a2 = _axes(rng)
    assert np.array_equal(O.C_anisotropic(Cm, True, a1, a2), O.C_anisotropic(Cm, False, a1, a2))
    # the Kelvin-Mandel scaling itself (applied to 3 x 3 inputs)
    M = np.ones((3, 3))
    assert np.allclose(O.kelvin_mandel(2, M), [[1, 1, np.sqrt(2)], [1, 1, np.sqrt(2)], [np.sqrt(2)] * 2 + [2]])


def test_device_model_builds_the_same_global_C():
    """``Models.Elastic.Anisotropic`` (host side of the device path) forms
    the rotation its own way (rotated basis tensors); it agrees with the
    oracle's restatement of the reference's direction-cosine tables."""
    from easyfea_amd import Models

    rng = np.random.default_rng(3)
    for _ in range(5):
        a1, a2 = _axes(rng)
        Cm = _spd(rng)
        m = Models.Elastic.Anisotropic(3, Cm, True, a1, a2)
        assert np.abs(m.C - O.C_anisotropic(Cm, True, a1, a2)).max() < 1e-12 * np.abs(Cm).max()
        assert np.abs(m.S @ m.C - np.eye(6)).max() < 1e-10
    Ce = np.stack([_spd(rng) for _ in range(4)])
    m = Models.Elastic.Anisotropic(3, Ce, False, (0, 0, 1), (1, 0, 0))
    assert m.isHeterogeneous and np.abs(m.C - O.C_anisotropic(Ce, False, (0, 0, 1), (1, 0, 0))).max() < 1e-10
