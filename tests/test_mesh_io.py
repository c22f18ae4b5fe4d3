"""gmsh import / export and the mesh file (CPU only).

The reference reads gmsh files through the gmsh API (absent here); the
import follows its renumbering (node tags sorted, compacted), its element
filtering (degenerate elements dropped) and its tag naming, pinned on files
written by hand from the gmsh format specification (2.2 and 4.1 ASCII).
"""

import numpy as np
import pytest

import easyfea_amd as ea
from easyfea_amd import ElemType
from easyfea_amd.mesh_io import load_mesh, read_msh, write_msh

# unit square split into two triangles, node tags 1, 2, 5, 7 (a gap, out of
# order in the file), a left-edge SEG2 in physical group 1 "left", the
# triangles in group 2 "domain", plus a degenerate triangle (three
# collinear nodes) that the import drops
MSH22 = """$MeshFormat
2.2 0 8
$EndMeshFormat
$PhysicalNames
2
1 1 "left"
2 2 "domain"
$EndPhysicalNames
$Nodes
5
5 1.0 1.0 0.0
1 0.0 0.0 0.0
2 1.0 0.0 0.0
7 0.0 1.0 0.0
9 0.5 0.0 0.0
$EndNodes
$Elements
5
1 15 2 0 1 1
2 1 2 1 4 7 1
3 2 2 2 1 1 2 5
4 2 2 2 1 1 5 7
5 2 2 2 1 1 9 2
$EndElements
"""

MSH41 = """$MeshFormat
4.1 0 8
$EndMeshFormat
$PhysicalNames
2
1 1 "left"
2 2 "domain"
$EndPhysicalNames
$Entities
1 1 1 0
1 0 0 0 0
4 0 0 0 0 1 0 1 1 2 1 -1
1 0 0 0 1 1 0 1 2 0
$EndEntities
$Nodes
1 5 1 9
2 1 0 5
5
1
2
7
9
1.0 1.0 0.0
0.0 0.0 0.0
1.0 0.0 0.0
0.0 1.0 0.0
0.5 0.0 0.0
$EndNodes
$Elements
3 5 1 5
0 1 15 1
1 1
1 4 1 1
2 7 1
2 1 2 3
3 1 2 5
4 1 5 7
5 1 9 2
$EndElements
"""


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


@pytest.mark.parametrize("text", [MSH22, MSH41], ids=["2.2", "4.1"])
def test_gmsh_import_renumbers_filters_and_tags(tmp_path, text):
    m = read_msh(_write(tmp_path, "sq.msh", text), coef=2.0)
    # tags 1, 2, 5, 7, 9 -> 0..4 in tag order
    assert np.array_equal(m.coord, 2.0 * np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0, 0]]))
    assert m.dim == 2 and m.elemType == ElemType.TRI3
    assert np.array_equal(m.connect, [[0, 1, 2], [0, 2, 3]])  # the degenerate triangle is gone
    assert np.array_equal(m.dict_groupElem[ElemType.SEG2].connect, [[3, 0]])
    assert np.array_equal(m.Nodes_Tags(["left"]), [0, 3])
    assert np.array_equal(m.Nodes_Tags("domain"), [0, 1, 2, 3, 4])  # nodes of the group's elements
    assert np.array_equal(m.orphanNodes, [4])


def test_gmsh_round_trip_of_generated_meshes(tmp_path):
    for mesh in (ea.quad4_rectangle(2.0, 1.0, 5, 3), ea.hexa8_box(1.0, 2.0, 1.5, 3, 2, 2),
                 ea.mesh.tetra10_from_tetra4(ea.mesh.tetra4_box(1.0, 1.0, 1.0, 2, 2, 2))):
        p = write_msh(mesh, str(tmp_path / "m.msh"))
        back = read_msh(p)
        assert np.array_equal(back.coord, mesh.coord)
        for et, g in mesh.dict_groupElem.items():
            assert np.array_equal(back.dict_groupElem[et].connect, g.connect)


def test_mesh_file_round_trip(tmp_path):
    m = read_msh(_write(tmp_path, "sq.msh", MSH41))
    path = m.Save(str(tmp_path / "out"), "square")
    back = load_mesh(path)
    assert np.array_equal(back.coord, m.coord)
    for et, g in m.dict_groupElem.items():
        assert np.array_equal(back.dict_groupElem[et].connect, g.connect)
    assert set(back.tags) == {"left", "domain"}
    assert np.array_equal(back.Nodes_Tags(["left", "domain"]), m.Nodes_Tags(["left", "domain"]))


def test_binary_and_unknown_files_are_refused(tmp_path):
    with pytest.raises(ValueError, match="binary"):
        read_msh(_write(tmp_path, "b.msh", "$MeshFormat\n4.1 1 8\n$EndMeshFormat\n"))
    with pytest.raises(AssertionError):
        read_msh(_write(tmp_path, "x.msh", "hello\n"))


def test_model_headers_round_trip():
    """Model parameters (scalars and per-element arrays) survive the
    simulation file's header."""
    import json

    from easyfea_amd import Models
    from easyfea_amd.sim_io import _model_from, _model_header

    E = np.linspace(1.0, 2.0, 7)
    mats = [Models.Elastic.Isotropic(2, E, 0.25, planeStress=False, thickness=3.0),
            Models.PhaseField(Models.Elastic.Isotropic(3, 5.0, 0.2), "Bourdin", "AT1", np.arange(1.0, 8.0), 0.1,
                              "HistoryDamage"),
            Models.Thermal(np.full(7, 2.0), 0.5, thickness=2.0)]
    for m in mats:
        arrays = {}
        h = json.loads(json.dumps(_model_header(m, arrays)))
        back = _model_from(h, arrays)
        assert json.dumps(_model_header(back, {}), sort_keys=True) == json.dumps(h, sort_keys=True)
    pf = _model_from(json.loads(json.dumps(_model_header(mats[1], {}))), {"p/Gc": np.arange(1.0, 8.0)})
    assert pf.split.value == "Bourdin" and pf.regularization.value == "AT1" and pf.solver.value == "HistoryDamage"
    assert np.array_equal(pf.Gc, np.arange(1.0, 8.0)) and pf.material.dim == 3
