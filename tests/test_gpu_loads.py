"""Body and pressure loads through the device integration vs the CPU oracle
(GPU tests).

``add_volumeLoad`` integrates over the main elements with the device load
kernel (TRI3, QUAD4, TETRA4, TETRA10, HEXA8); ``add_pressureLoad`` takes the
host nodal normals and integrates them over the boundary elements on the
device.  Tolerance: Neumann vectors within 1e-12 relative of the oracle's
(both are FP64 sums of a few products per entry); solutions within 1e-8 of
the oracle's sparse direct solve.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import easyfea_amd as ea  # noqa: E402
from easyfea_amd import Models, Simulations  # noqa: E402
from oracle import efea_oracle as O  # noqa: E402


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _moved(m, amp=0.2, seed=3):
    """Interior nodes moved so that every element has its own Jacobian."""
    rng = np.random.default_rng(seed)
    c = m.coord.copy()
    inner = np.ones(m.Nn, bool)
    for k in range(m.dim):
        inner &= (c[:, k] > c[:, k].min()) & (c[:, k] < c[:, k].max())
    h = (c.max(0) - c.min(0))[: m.dim].min() / 6
    c[inner, : m.dim] += amp * h * (rng.random((inner.sum(), m.dim)) - 0.5)
    return ea.Mesh({t: g.connect for t, g in m.dict_groupElem.items()}, c)


MAKERS = {
    "TRI3": lambda: _moved(ea.mesh.tri3_rectangle(3.0, 2.0, 6, 5)),
    "QUAD4": lambda: _moved(ea.quad4_rectangle(3.0, 2.0, 6, 5)),
    "TETRA4": lambda: _moved(ea.mesh.tetra4_box(3.0, 2.0, 1.0, 3, 3, 2)),
    "TETRA10": lambda: ea.mesh.tetra10_from_tetra4(ea.mesh.tetra4_box(3.0, 2.0, 1.0, 2, 2, 1)),
    "HEXA8": lambda: _moved(ea.hexa8_box(3.0, 2.0, 1.0, 3, 3, 2)),
}


def _pair(mesh, th=0.5, boundary=None):
    dim = mesh.dim
    mat = Models.Elastic.Isotropic(dim, 210000.0, 0.3, planeStress=True, thickness=th) if dim == 2 \
        else Models.Elastic.Isotropic(3, 210000.0, 0.3)
    simu = Simulations.Elastic(mesh, mat)
    simu.solver = "hip_fp64"
    ref = O.ElasticOracle(mesh.connect, mesh.coord, mesh.elemType.value, 210000.0, 0.3, True,
                          th if dim == 2 else 1.0, boundary or {})
    return simu, ref


@pytest.mark.parametrize("kind", ["constant", "callable", "nodal"])
@pytest.mark.parametrize("et", list(MAKERS))
def test_volume_load_matches_oracle(et, kind):
    mesh = MAKERS[et]()
    dim = mesh.dim
    simu, ref = _pair(mesh)
    th = 0.5 if dim == 2 else 1.0
    # a subset of the nodes: only elements whose every node is in it count
    nodes = mesh.Nodes_Conditions(lambda x, y, z: x <= 2.0)
    if kind == "constant":
        values, unk = [-7.85e-9 * 9.81, 2.5], ["y", "x"]
    elif kind == "callable":
        values, unk = [lambda x, y, z: 1.0 + x * y - z, lambda x, y, z: x * x], ["x", "y"]
    else:
        c = mesh.coord[nodes]
        values, unk = [1.0 + c[:, 0] - 2 * c[:, 1], c[:, 1] ** 2], ["y", "x"]
    simu.add_volumeLoad(nodes, values, unk)
    ref.add_volumeLoad(nodes, values, unk, thickness=th)
    F = simu._neumann_vector("elastic")
    Fr = ref.neumann_vector("elastic", mesh.Nn * dim)
    assert _rel(F, Fr) < 1e-12, (et, kind)


def _quarter_annulus(a=1.0, b=2.0, nr=4, nt=8):
    """QUAD4 quarter annulus: the (r, theta) rectangle mapped to the plane
    (orientation-preserving, so the boundary stays counter-clockwise)."""
    m = ea.quad4_rectangle(b - a, np.pi / 2, nr, nt, x0=a)
    r, t = m.coord[:, 0].copy(), m.coord[:, 1].copy()
    c = m.coord.copy()
    c[:, 0], c[:, 1] = r * np.cos(t), r * np.sin(t)
    c[np.isclose(t, np.pi / 2), 0] = 0.0
    return ea.Mesh({tt: g.connect for tt, g in m.dict_groupElem.items()}, c), r


@pytest.mark.parametrize("case", ["annulus-bore", "rectangle-corner", "box-faces"])
def test_pressure_load_matches_oracle(case):
    if case == "annulus-bore":
        mesh, r = _quarter_annulus()
        nodes = np.nonzero(np.isclose(r, 1.0))[0]
        bt = ea.ElemType.SEG2
    elif case == "rectangle-corner":
        mesh = _moved(ea.quad4_rectangle(3.0, 2.0, 6, 5))
        nodes = mesh.Nodes_Conditions(lambda x, y, z: (x == 3.0) | (y == 2.0))
        bt = ea.ElemType.SEG2
    else:
        mesh = _moved(ea.hexa8_box(3.0, 2.0, 1.0, 3, 3, 2))
        nodes = mesh.Nodes_Conditions(lambda x, y, z: (z == 1.0) | (x == 0.0))
        bt = ea.ElemType.QUAD4
    simu, ref = _pair(mesh, boundary={bt.value: mesh.dict_groupElem[bt].connect})
    th = 0.5 if mesh.dim == 2 else 1.0
    simu.add_pressureLoad(nodes, 3.0)
    ref.add_pressureLoad(nodes, 3.0, thickness=th)
    F = simu._neumann_vector("elastic")
    Fr = ref.neumann_vector("elastic", mesh.Nn * mesh.dim)
    assert np.abs(F).max() > 0 and _rel(F, Fr) < 1e-12
    if case == "box-faces":  # the flat top face alone: p * area * (+z), area 6
        simu.Bc_Init()
        simu.add_pressureLoad(mesh.Nodes_Conditions(lambda x, y, z: z == 1.0), 3.0)
        tot = simu._neumann_vector("elastic").reshape(-1, 3).sum(0)
        assert np.allclose(tot, [0.0, 0.0, 3.0 * 6.0], rtol=1e-12, atol=1e-12)


def test_pressure_is_refused_on_scalar_problems(capsys):
    mesh = ea.quad4_rectangle(1.0, 1.0, 4, 4)
    simu = Simulations.Thermal(mesh, Models.Thermal(1.0, 1.0))
    simu.add_pressureLoad(mesh.Nodes_Conditions(lambda x, y, z: x == 1.0), 2.0)
    assert "scalar" in capsys.readouterr().out
    assert simu._neumann_vector("thermal").size == 0 or not simu._neumann_vector("thermal").any()


def test_self_weight_cantilever_matches_oracle():
    """HEXA8 beam clamped at x = 0 under its own weight (add_volumeLoad) and a
    pressure on its free end: u within 1e-8 of the oracle's direct solve."""
    mesh = ea.hexa8_box(8.0, 1.0, 1.0, 16, 2, 2)
    simu, ref = _pair(mesh, boundary={"QUAD4": mesh.dict_groupElem[ea.ElemType.QUAD4].connect})
    clamp = mesh.Nodes_Conditions(lambda x, y, z: x == 0)
    end = mesh.Nodes_Conditions(lambda x, y, z: x == 8.0)
    rho_g = -7.85e-9 * 9.81
    for s in (simu, ref):
        s.add_dirichlet(clamp, [0, 0, 0], ["x", "y", "z"])
        s.add_volumeLoad(mesh.nodes, [rho_g], ["z"])
        s.add_pressureLoad(end, -0.5)
    u = simu.Solve()
    uref = ref.Solve()
    assert _rel(u, uref) < 1e-8
    tot = simu._neumann_vector("elastic").reshape(-1, 3).sum(0)
    assert np.allclose(tot, [-0.5, 0.0, rho_g * 8.0], rtol=1e-12, atol=1e-15)
