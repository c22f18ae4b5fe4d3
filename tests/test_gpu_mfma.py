"""Dense 3D element stiffness on the FP64 matrix cores
(``efea_elastic_ke_dense``, GPU tests): K_e = sum_p wJ B^T D B for HEXA8 and
TETRA10 on distorted meshes against the oracle's ``K_elastic``
(``FEM/Operators/Bilinear.py:62-79``), with the isotropic D of a
heterogeneous E, with a random symmetric positive definite D per element
and with one shared D; and against the closed-form isotropic kernel.
Bars: 1e-12 relative per element (Frobenius) and per entry for every entry
above 1e-2 of its element's largest; every entry within 1e-14 of the
element's largest (smaller entries are sums that cancel, in the oracle's
einsum as on the device: on the distorted TETRA10 mesh entries near 1e-3 of
the largest differ by 1.1e-12 relative, 1e-15 of the largest, while the
matrix-core and closed-form device kernels agree to 3e-16)."""

import numpy as np
import pytest

import easyfea_amd as ea
from easyfea_amd import _lib
from easyfea_amd._utils import MatrixType
from easyfea_amd.device import DeviceMesh, ptr, stream_ptr, to_dev
from oracle import efea_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

MESHES = [
    (lambda: ea.hexa8_box(1.0, 2.0, 1.5, 4, 3, 5), "HEXA8"),
    (lambda: ea.mesh.tetra10_from_tetra4(ea.mesh.tetra4_box(1.0, 2.0, 1.5, 3, 2, 4)), "TETRA10"),
]


def _distorted(mesh, amp, seed=0):
    rng = np.random.default_rng(seed)
    c = mesh.coord.copy()
    interior = np.ones(mesh.Nn, dtype=bool)
    for k in range(3):
        interior &= (c[:, k] > c[:, k].min()) & (c[:, k] < c[:, k].max())
    span = c.max(0) - c.min(0)
    h = span.min() / max(10, round(mesh.Ne ** (1 / 3)))
    c[interior] += amp * h * (rng.random((interior.sum(), 3)) - 0.5)
    return ea.Mesh({et: g.connect for et, g in mesh.dict_groupElem.items()}, c)


def _dense(mesh, D=None, E=None, nu=0.3):
    dev = DeviceMesh(mesh)
    tab, nPg = dev.tables(MatrixType.rigi)
    nd = dev.nPe * 3
    Ke = torch.empty(dev.Ne * nd * nd, dtype=torch.float64, device="cuda")
    Dd = None if D is None else to_dev(np.ascontiguousarray(D, dtype=np.float64).reshape(-1))
    Ed = None if E is None else to_dev(np.atleast_1d(np.asarray(E, float)))
    _lib.call("efea_elastic_ke_dense", dev.code, dev.Ne, ptr(dev.connect), ptr(dev.coord), ptr(tab), nPg, ptr(Dd),
              int(D is not None and np.ndim(D) == 3), ptr(Ed), int(E is not None and np.ndim(E) > 0), nu, ptr(Ke),
              stream_ptr())
    return Ke.cpu().numpy().reshape(dev.Ne, nd, nd)


def _check(K, ref):
    per_el = np.max(np.linalg.norm(K - ref, axis=(1, 2)) / np.linalg.norm(ref, axis=(1, 2)))
    top = np.abs(ref).max(axis=(1, 2), keepdims=True)
    big = np.abs(ref) >= 1e-2 * top
    per_entry = np.max(np.abs(K - ref)[big] / np.abs(ref)[big])
    assert np.max(np.abs(K - ref) / top) < 1e-14
    return per_el, per_entry


def _spd_D(rng, n):
    Q = rng.standard_normal((n, 6, 6))
    return 1e4 * (Q @ np.swapaxes(Q, 1, 2) + 6 * np.eye(6))


@pytest.mark.parametrize("maker,et", MESHES)
def test_dense_isotropic_heterogeneous_E(maker, et):
    mesh = _distorted(maker(), 0.4)
    E = 210000.0 * np.exp(0.2 * np.random.default_rng(1).standard_normal(mesh.Ne))
    K = _dense(mesh, E=E)
    geo = O.Geometry(et, mesh.connect, mesh.coord, "rigi", 3)
    per_el, per_entry = _check(K, O.K_elastic(geo, O.C_isotropic(E, 0.3, 3, False)))
    print(f"{et} isotropic: per element {per_el:.2e}, per entry {per_entry:.2e}")
    assert per_el < 1e-12 and per_entry < 1e-12


@pytest.mark.parametrize("maker,et", MESHES)
def test_dense_general_D_per_element_and_shared(maker, et):
    mesh = _distorted(maker(), 0.4, seed=3)
    rng = np.random.default_rng(2)
    geo = O.Geometry(et, mesh.connect, mesh.coord, "rigi", 3)
    D = _spd_D(rng, mesh.Ne)
    per_el, per_entry = _check(_dense(mesh, D=D), O.K_elastic(geo, D))
    D1 = _spd_D(rng, 1)[0]
    per_el1, per_entry1 = _check(_dense(mesh, D=D1), O.K_elastic(geo, D1))
    print(f"{et} general D: per element {per_el:.2e} / {per_el1:.2e}, per entry {per_entry:.2e} / {per_entry1:.2e}")
    assert max(per_el, per_el1) < 1e-12 and max(per_entry, per_entry1) < 1e-12


@pytest.mark.parametrize("maker,et", MESHES)
def test_dense_equals_closed_form(maker, et):
    mesh = _distorted(maker(), 0.3, seed=5)
    E = 210000.0 * np.exp(0.2 * np.random.default_rng(4).standard_normal(mesh.Ne))
    K = _dense(mesh, E=E)
    dev = DeviceMesh(mesh)
    tab, nPg = dev.tables(MatrixType.rigi)
    nd = dev.nPe * 3
    Kc = torch.empty(dev.Ne * nd * nd, dtype=torch.float64, device="cuda")
    _lib.call("efea_elastic_ke", dev.code, 64, dev.Ne, ptr(dev.connect), ptr(dev.coord), ptr(tab), nPg,
              ptr(to_dev(E)), 1, 0.3, 0, 1.0, ptr(Kc), stream_ptr())
    Kc = Kc.cpu().numpy().reshape(dev.Ne, nd, nd)
    err = np.max(np.linalg.norm(K - Kc, axis=(1, 2)) / np.linalg.norm(Kc, axis=(1, 2)))
    print(f"{et}: matrix cores vs closed form {err:.2e}")
    assert err < 1e-13


@pytest.mark.parametrize("maker,et", MESHES)
def test_anisotropic_solve_matches_direct(maker, et):
    """A rotated orthotropic law (``Models.Elastic.Anisotropic``) through
    ``Simulations.Elastic``: the matrix-core element matrices, the scatter
    and the two-level / block-Jacobi PCG against the oracle's direct solve."""
    from easyfea_amd import Models, Simulations

    mesh = _distorted(maker(), 0.3, seed=7)
    E1, E2, E3, G23, G13, G12 = 150e3, 10e3, 12e3, 4e3, 5e3, 6e3
    nu12, nu13, nu23 = 0.3, 0.25, 0.4
    S = np.zeros((6, 6))
    S[:3, :3] = [[1 / E1, -nu12 / E1, -nu13 / E1], [-nu12 / E1, 1 / E2, -nu23 / E2], [-nu13 / E1, -nu23 / E2, 1 / E3]]
    S[3, 3], S[4, 4], S[5, 5] = 1 / (2 * G23), 1 / (2 * G13), 1 / (2 * G12)  # Kelvin-Mandel compliance
    Cm = np.linalg.inv(S)
    a1 = np.array([1.0, 1.0, 0.5]) / np.linalg.norm([1.0, 1.0, 0.5])
    a2 = np.cross(a1, [0.0, 0.0, 1.0])
    a2 /= np.linalg.norm(a2)
    law = Models.Elastic.Anisotropic(3, Cm, False, a1, a2)
    simu = Simulations.Elastic(mesh, law)
    simu.solver = "hip_fp64"
    allnodes = np.arange(mesh.Nn)
    simu.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: z == 0), [0, 0, 0], ["x", "y", "z"])
    simu.add_volumeLoad(allnodes, [-10.0, 3.0], ["z", "x"])
    u = simu.Solve()
    ref = O.ElasticOracle(mesh.connect, mesh.coord, et, 1.0, 0.3, False, 1.0, {}, C=law.C)
    ref.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: z == 0), [0, 0, 0], ["x", "y", "z"])
    ref.add_volumeLoad(allnodes, [-10.0, 3.0], ["z", "x"])
    ur = ref.Solve()
    err = np.linalg.norm(u - ur) / np.linalg.norm(ur)
    info = simu.last_solve["elastic"]
    print(f"{et} anisotropic solve: {info.iterations} it ({info.precond}), vs direct {err:.2e}")
    assert err < 1e-8
