"""Mesh files: gmsh ``.msh`` import / export and the library's own mesh file.

``read_msh`` builds a ``Mesh`` from an ASCII gmsh file (format 2.2 or 4.1)
the way ``Mesher.Mesh_Import_mesh`` / ``_Mesh_Get_Mesh`` build one through the
gmsh API (``EasyFEA/FEM/_mesher.py:1129-1186, 2227-2301, 2395-2475``):

* node tags sorted ascending and renumbered 0..Nn-1 in that order (the
  ``changes`` map of ``__Get_coordinates_and_changes``), coordinates times
  ``coef``;
* one element group per supported element type, elements in file order,
  node order as gmsh writes it (EasyFEA's element definitions use gmsh's);
  elements of zero measure dropped (the reference drops those whose sampled
  minimal scaled Jacobian is 0);
* each physical group becomes a node tag: the nodes of its elements
  (``getNodesForPhysicalGroup``), named by ``$PhysicalNames`` or, without a
  name, ``P|L|S|V{tag}`` after the entity names the reference gives.

Supported element types: SEG2 (gmsh 1), TRI3 (2), QUAD4 (3), TETRA4 (4),
HEXA8 (5), TETRA10 (11); points and other types are skipped.  Binary files
are refused.

``Mesh.Save`` / ``Load_Mesh`` store coordinates, connectivities and tags in
an ``.npz`` file (no pickle): the reference pickles its own classes
(``FEM/_mesh.py:1309-1400``), which only EasyFEA itself can load.
"""

from __future__ import annotations

import os

import numpy as np

from ._utils import ElemType
from .mesh import Mesh

GMSH_TYPES = {1: ElemType.SEG2, 2: ElemType.TRI3, 3: ElemType.QUAD4, 4: ElemType.TETRA4, 5: ElemType.HEXA8,
              11: ElemType.TETRA10}
GMSH_IDS = {et: gid for gid, et in GMSH_TYPES.items()}
_DIM_PREFIX = {0: "P", 1: "L", 2: "S", 3: "V"}


def _sections(text: str) -> dict:
    """{section name: [lines]} of a gmsh file ($Name ... $EndName)."""
    out: dict = {}
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        ln = lines[i].strip()
        if ln.startswith("$") and not ln.startswith("$End"):
            name = ln[1:]
            j = i + 1
            while j < len(lines) and lines[j].strip() != "$End" + name:
                j += 1
            out.setdefault(name, []).append(lines[i + 1:j])
            i = j
        i += 1
    return out


def _measure(et: ElemType, X: np.ndarray) -> np.ndarray:
    """|Jacobian| at the reference centroid (a zero means a degenerate element)."""
    if et == ElemType.SEG2:
        return np.linalg.norm(X[:, 1] - X[:, 0], axis=1)
    if et == ElemType.TRI3:
        return np.linalg.norm(np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), axis=1)
    if et == ElemType.QUAD4:  # the diagonals' cross product (twice the area of a planar quad)
        return np.linalg.norm(np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 1]), axis=1)
    if et in (ElemType.TETRA4, ElemType.TETRA10):
        return np.abs(np.einsum("ei,ei->e", X[:, 1] - X[:, 0], np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0])))
    if et == ElemType.HEXA8:  # Jacobian at the centre, gmsh node order
        jx = (X[:, 1] + X[:, 2] + X[:, 5] + X[:, 6]) - (X[:, 0] + X[:, 3] + X[:, 4] + X[:, 7])
        jy = (X[:, 2] + X[:, 3] + X[:, 6] + X[:, 7]) - (X[:, 0] + X[:, 1] + X[:, 4] + X[:, 5])
        jz = (X[:, 4] + X[:, 5] + X[:, 6] + X[:, 7]) - (X[:, 0] + X[:, 1] + X[:, 2] + X[:, 3])
        return np.abs(np.einsum("ei,ei->e", jx, np.cross(jy, jz)))
    raise ValueError(et)


def _parse_v2(sec: dict):
    nodes = sec["Nodes"][0]
    n = int(nodes[0])
    arr = np.array([ln.split() for ln in nodes[1:n + 1]], dtype=float)
    tags, xyz = arr[:, 0].astype(np.int64), arr[:, 1:4]
    elems = []  # (gmsh type, physical tag, node tags)
    lines = sec["Elements"][0]
    for ln in lines[1:int(lines[0]) + 1]:
        f = [int(v) for v in ln.split()]
        gtype, ntags = f[1], f[2]
        phys = f[3] if ntags > 0 else 0
        elems.append((gtype, phys, f[3 + ntags:]))
    by_type: dict = {}
    phys_nodes: dict = {}
    for gtype, phys, nt in elems:
        by_type.setdefault(gtype, []).append(nt)
        if phys:
            phys_nodes.setdefault(phys, set()).update(nt)
    return tags, xyz, by_type, {k: (None, v) for k, v in phys_nodes.items()}


def _parse_v4(sec: dict):
    # $Entities: entity (dim, tag) -> physical tags
    ent_phys: dict = {}
    if "Entities" in sec:
        L = sec["Entities"][0]
        counts = [int(v) for v in L[0].split()]
        k = 1
        for dim, cnt in enumerate(counts):
            for _ in range(cnt):
                f = L[k].split()
                k += 1
                tag = int(f[0])
                # points: tag x y z nphys phys...; others: tag minx miny minz maxx maxy maxz nphys phys... nbound ...
                off = 4 if dim == 0 else 7
                nph = int(f[off])
                ent_phys[(dim, tag)] = [int(v) for v in f[off + 1:off + 1 + nph]]
    L = sec["Nodes"][0]
    nblocks = int(L[0].split()[0])
    k = 1
    tags_l, xyz_l = [], []
    for _ in range(nblocks):
        f = L[k].split()
        k += 1
        parametric, cnt = int(f[2]), int(f[3])
        t = [int(L[k + i]) for i in range(cnt)]
        k += cnt
        c = np.array([L[k + i].split()[:3] for i in range(cnt)], dtype=float) if cnt else np.zeros((0, 3))
        k += cnt
        assert parametric == 0 or c.shape[1] >= 3
        tags_l += t
        xyz_l.append(c.reshape(-1, 3))
    tags = np.array(tags_l, dtype=np.int64)
    xyz = np.concatenate(xyz_l) if xyz_l else np.zeros((0, 3))
    L = sec["Elements"][0]
    nblocks = int(L[0].split()[0])
    k = 1
    by_type: dict = {}
    phys_nodes: dict = {}
    for _ in range(nblocks):
        f = [int(v) for v in L[k].split()]
        k += 1
        edim, etag, gtype, cnt = f
        rows = [[int(v) for v in L[k + i].split()[1:]] for i in range(cnt)]
        k += cnt
        by_type.setdefault(gtype, []).extend(rows)
        for ph in ent_phys.get((edim, etag), []):
            s = phys_nodes.setdefault((edim, ph), set())
            for r in rows:
                s.update(r)
    return tags, xyz, by_type, {ph: (dim, v) for (dim, ph), v in phys_nodes.items()}


def read_msh(path: str, coef: float = 1.0) -> Mesh:
    """``Mesher.Mesh_Import_mesh(path, coef=coef)`` for an ASCII gmsh file."""
    with open(path, "rb") as fh:
        raw = fh.read()
    text = raw.decode("latin-1")
    sec = _sections(text)
    assert "MeshFormat" in sec, f"{path}: not a gmsh file"
    version, ftype = sec["MeshFormat"][0][0].split()[:2]
    if int(ftype) != 0:
        raise ValueError(f"{path}: binary gmsh files are not supported (save as ASCII)")
    major = int(float(version))
    if major == 2:
        tags, xyz, by_type, phys = _parse_v2(sec)
    elif major == 4:
        tags, xyz, by_type, phys = _parse_v4(sec)
    else:
        raise ValueError(f"{path}: gmsh format {version} is not supported (2.2 or 4.1)")
    names: dict = {}
    for lines in sec.get("PhysicalNames", []):
        for ln in lines[1:int(lines[0]) + 1]:
            f = ln.split(maxsplit=2)
            names[(int(f[0]), int(f[1]))] = f[2].strip().strip('"')
    # nodes sorted by tag, renumbered in that order (__Get_coordinates_and_changes)
    order = np.argsort(tags, kind="stable")
    s_tags = tags[order]
    coord = xyz[order] * coef
    changes = np.full(int(s_tags.max()) + 1 if s_tags.size else 1, -1, dtype=np.int64)
    changes[s_tags] = np.arange(s_tags.size)
    groups = {}
    for gtype, rows in by_type.items():
        if gtype not in GMSH_TYPES:
            continue
        et = GMSH_TYPES[gtype]
        connect = changes[np.asarray(rows, dtype=np.int64).reshape(len(rows), et.nPe)]
        keep = _measure(et, coord[connect]) != 0
        groups[et] = connect[keep]
    if not groups:
        raise ValueError(f"{path}: no element of a supported type")
    mesh = Mesh(groups, coord)
    tags_out = {}
    for ph, (dim, nodes) in sorted(phys.items(), key=lambda kv: str(kv[0])):
        name = names.get((dim, ph)) if dim is not None else next((v for (d, t), v in names.items() if t == ph), None)
        if name is None:
            name = f"{_DIM_PREFIX.get(dim, 'P')}{ph}"
        nd = changes[np.fromiter(nodes, dtype=np.int64)]
        tags_out[name] = np.unique(nd[nd >= 0])
    mesh.tags = tags_out
    return mesh


def write_msh(mesh: Mesh, path: str) -> str:
    """ASCII gmsh 2.2 file of ``mesh``: nodes 1..Nn, every element group,
    one physical group per node tag is not representable (gmsh groups are
    element sets), so tags are not written."""
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(mesh.Nn)]
    lines += [f"{i + 1} {x!r} {y!r} {z!r}" for i, (x, y, z) in enumerate(mesh.coord.tolist())]
    lines += ["$EndNodes", "$Elements"]
    rows = []
    k = 1
    for et, g in mesh.dict_groupElem.items():
        gid = GMSH_IDS[ElemType(et)]
        for conn in (g.connect + 1).tolist():
            rows.append(f"{k} {gid} 2 0 {g.dim} " + " ".join(map(str, conn)))
            k += 1
    lines += [str(len(rows))] + rows + ["$EndElements"]
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path


def save_mesh(mesh: Mesh, folder: str, filename: str = "mesh") -> str:
    """``Mesh.Save``: ``folder/filename.npz`` with the coordinates, each
    group's connectivity and the node tags."""
    os.makedirs(folder, exist_ok=True)
    path = os.path.join(folder, f"{filename}.npz")
    arrays = {"coord": mesh.coord}
    for et, g in mesh.dict_groupElem.items():
        arrays[f"connect:{ElemType(et).value}"] = g.connect
    for name, nodes in getattr(mesh, "tags", {}).items():
        arrays[f"tag:{name}"] = np.asarray(nodes, dtype=np.int64)
    np.savez(path, **arrays)
    return path


def load_mesh(path: str) -> Mesh:
    """``Load_Mesh`` of a file written by ``save_mesh`` (no pickle)."""
    assert os.path.exists(path), f"The file {path} cannot be found."
    with np.load(path, allow_pickle=False) as z:
        coord = z["coord"]
        groups = {k.split(":", 1)[1]: z[k] for k in z.files if k.startswith("connect:")}
        tags = {k.split(":", 1)[1]: z[k] for k in z.files if k.startswith("tag:")}
    mesh = Mesh(groups, coord)
    mesh.tags = tags
    return mesh
