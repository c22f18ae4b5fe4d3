"""Body and pressure loads: the oracle's restatement pinned by closed forms,
and the host-side nodal normals (``Mesh.Get_normals``) against the same
closed forms (CPU tests).

Pins: on straight boundaries the nodal normals are the edge/face normals
(a corner node shared by two edges gets their normalised sum); a pressure on
a flat face adds up to ``p * area * n``; a body load adds up to ``f * V``
and its first moment to ``f * V * x_c`` (a transposed or mis-indexed
shape-function table fails the moment); nodal-array values interpolate
exactly, so ``x`` as nodal values integrates to ``V * x_c``.
"""

import numpy as np
import pytest

import easyfea_amd as ea
from oracle import efea_oracle as O


def _rect(nx=6, ny=4, Lx=3.0, Ly=2.0):
    m = ea.quad4_rectangle(Lx, Ly, nx, ny)
    return m, m.dict_groupElem[ea.ElemType.SEG2].connect


@pytest.mark.parametrize("who", ["oracle", "mesh"])
def test_rectangle_normals_in_closed_form(who):
    """Edges run counter-clockwise, so ``e_z x t`` points into the domain."""
    m, seg = _rect()
    nodes = m.Nodes_Conditions(lambda x, y, z: (x == 0) | (y == 0))
    if who == "oracle":
        n, used = O.node_normals("SEG2", seg, m.coord, nodes)
    else:
        n, used = m.Get_normals(nodes)
    assert np.array_equal(used, np.sort(nodes))
    c = m.coord[used]
    expect = np.zeros((used.size, 3))
    expect[c[:, 0] == 0, 0] += 1.0
    expect[c[:, 1] == 0, 1] += 1.0
    expect /= np.linalg.norm(expect, axis=1, keepdims=True)
    assert np.allclose(n, expect, atol=1e-15)


def test_box_face_normals_point_out():
    m = ea.hexa8_box(1.0, 2.0, 1.5, 3, 4, 2)
    for k, val, sign in ((2, 1.5, 1.0), (2, 0.0, -1.0), (0, 0.0, -1.0), (1, 2.0, 1.0)):
        nodes = m.Nodes_Conditions(lambda x, y, z, k=k, val=val: np.stack([x, y, z])[k] == val)
        n, used = m.Get_normals(nodes)
        e = np.zeros(3)
        e[k] = sign
        assert used.size == nodes.size and np.allclose(n, e, atol=1e-15)
        no, uo = O.node_normals("QUAD4", m.dict_groupElem[ea.ElemType.QUAD4].connect, m.coord, nodes)
        assert np.array_equal(uo, used) and np.allclose(no, n, atol=1e-15)


def test_in_dim():
    assert ea.quad4_rectangle(1, 1, 2, 2).inDim == 2
    assert ea.hexa8_cube(2).inDim == 3


def test_pressure_on_flat_faces_adds_up():
    # 2D: right edge x = Lx, normal -x (into the domain), times thickness
    m, seg = _rect()
    th, p = 0.7, 5.0
    ref = O.ElasticOracle(m.connect, m.coord, "QUAD4", 1.0, 0.3, True, th, {"SEG2": seg})
    ref.add_pressureLoad(m.Nodes_Conditions(lambda x, y, z: x == 3.0), p, thickness=th)
    F = ref.neumann_vector("elastic", m.Nn * 2).reshape(-1, 2)
    assert np.isclose(F[:, 0].sum(), -p * 2.0 * th, rtol=1e-13) and abs(F[:, 1].sum()) < 1e-13
    # 3D: top face z = Lz, outward +z
    b = ea.hexa8_box(1.0, 2.0, 1.5, 3, 4, 2)
    ref = O.ElasticOracle(b.connect, b.coord, "HEXA8", 1.0, 0.3, False, 1.0,
                          {"QUAD4": b.dict_groupElem[ea.ElemType.QUAD4].connect})
    ref.add_pressureLoad(b.Nodes_Conditions(lambda x, y, z: z == 1.5), p)
    F = ref.neumann_vector("elastic", b.Nn * 3).reshape(-1, 3)
    assert np.allclose(F.sum(0), [0, 0, p * 2.0], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("et", ["QUAD4", "TRI3", "HEXA8", "TETRA4", "TETRA10"])
def test_volume_load_sum_and_moment(et):
    if et == "QUAD4":
        m = ea.quad4_rectangle(3.0, 2.0, 5, 4)
    elif et == "TRI3":
        m = ea.mesh.tri3_rectangle(3.0, 2.0, 5, 4)
    elif et == "HEXA8":
        m = ea.hexa8_box(3.0, 2.0, 1.0, 3, 2, 2)
    elif et == "TETRA4":
        m = ea.mesh.tetra4_box(3.0, 2.0, 1.0, 3, 2, 2)
    else:
        m = ea.mesh.tetra10_from_tetra4(ea.mesh.tetra4_box(3.0, 2.0, 1.0, 2, 2, 1))
    dim = m.dim
    th = 0.5 if dim == 2 else 1.0
    V = 6.0 * (th if dim == 2 else 1.0)
    xc = np.array([1.5, 1.0, 0.5])[:dim]
    f = -9.81
    unk = ["x", "y", "z"][:dim]
    ref = O.ElasticOracle(m.connect, m.coord, et, 1.0, 0.3, True, th)
    ref.add_volumeLoad(m.nodes, [f], ["y"], thickness=th)
    F = ref.neumann_vector("elastic", m.Nn * dim).reshape(-1, dim)
    assert np.isclose(F[:, 1].sum(), f * V, rtol=1e-13)
    assert np.allclose(F[:, 1] @ m.coord[:, :dim], f * V * xc, rtol=1e-12)
    assert np.abs(np.delete(F, 1, axis=1)).max() == 0.0
    # nodal values x interpolate exactly: sum = V * x_c
    ref.Bc_Init()
    ref.add_volumeLoad(m.nodes, [m.coord[:, 0]], [unk[0]], thickness=th)
    F = ref.neumann_vector("elastic", m.Nn * dim).reshape(-1, dim)
    assert np.isclose(F[:, 0].sum(), V * xc[0], rtol=1e-12)
    # a callable of the Gauss-point coordinates gives the same for f = x
    ref.Bc_Init()
    ref.add_volumeLoad(m.nodes, [lambda x, y, z: x], [unk[0]], thickness=th)
    assert np.isclose(ref.neumann_vector("elastic", m.Nn * dim).sum(), V * xc[0], rtol=1e-12)
