"""Mesh containers and structured mesh generators.

The reference builds meshes through gmsh (``EasyFEA/FEM/_mesher.py``), which
is not on the solve path.  This module provides the element groups the solve
path consumes (``Mesh`` / ``GroupElem``, after ``EasyFEA/FEM/_mesh.py:49-146``
and ``EasyFEA/FEM/_group_elem.py``) and deterministic generators for the
workloads: organised QUAD4 rectangles, the TRI3 single-edge-notched square and
HEXA8 boxes, each with its boundary groups (SEG2 edges in 2D, QUAD4 faces in
3D) so that surface loads integrate over boundary elements as in
``_simu.py:2770-2842``.

Numbering rule: node ``(i, j[, k])`` of an ``nx x ny [x nz]`` grid has id
``i + (nx+1)*j [+ (nx+1)*(ny+1)*k]``; element ids run ``x`` fastest.
"""

from typing import Callable, Optional

import numpy as np

from ._utils import ElemType


class GroupElem:
    """A group of elements of one type sharing the mesh coordinates."""

    def __init__(self, elemType: ElemType, connect: np.ndarray, coord: np.ndarray):
        self.elemType = ElemType(elemType)
        connect = np.ascontiguousarray(connect, dtype=np.int64)
        assert connect.ndim == 2 and connect.shape[1] == self.elemType.nPe
        self.connect = connect
        self.coord = coord
        self._node_elem = None

    @property
    def dim(self) -> int:
        return self.elemType.dim

    @property
    def nPe(self) -> int:
        return self.elemType.nPe

    @property
    def Ne(self) -> int:
        return self.connect.shape[0]

    @property
    def nodes(self) -> np.ndarray:
        c = self.connect
        cached = self.__dict__.get("_nodes_cache")
        if cached is None or cached[0] is not c:  # sorted unique nodes, kept per connectivity array
            cached = self.__dict__["_nodes_cache"] = (c, np.unique(c))
        return cached[1].copy()

    def Get_assembly_e(self, dof_n: int) -> np.ndarray:
        """Element DOF rows ``[n0*dof_n+0, n0*dof_n+1, ...]`` (``_group_elem.py:358-380``)."""
        Ne, nPe = self.connect.shape
        out = np.empty((Ne, nPe * dof_n), dtype=np.int64)
        for d in range(dof_n):
            out[:, d::dof_n] = self.connect * dof_n + d
        return out

    def Get_Nodes_Conditions(self, func: Callable) -> np.ndarray:
        nodes = self.nodes
        x, y, z = self.coord[nodes, 0], self.coord[nodes, 1], self.coord[nodes, 2]
        return nodes[np.asarray(func(x, y, z), dtype=bool)]

    def Get_Elements_Nodes(self, nodes: np.ndarray, exclusively=True) -> np.ndarray:
        """Elements using the nodes; with ``exclusively`` only elements whose
        every node is in ``nodes`` (``_group_elem.py:1500-1537``).  Returned
        sorted ascending."""
        nodes = np.asarray(nodes, dtype=np.int64).ravel()
        Nmax = max(int(self.connect.max()) + 1, int(nodes.max()) + 1 if nodes.size else 0)
        mark = np.zeros(Nmax, dtype=bool)
        mark[nodes] = True
        inside = mark[self.connect]
        sel = inside.all(axis=1) if exclusively else inside.any(axis=1)
        return np.nonzero(sel)[0].astype(np.int64)


class Mesh:
    """Mesh: coordinates ``(Nn, 3)`` and element groups keyed by type."""

    def __init__(self, groups: dict, coord: np.ndarray):
        coord = np.ascontiguousarray(coord, dtype=np.float64)
        assert coord.ndim == 2 and coord.shape[1] == 3
        self.coord = coord
        self.dict_groupElem = {}
        for elemType, connect in groups.items():
            self.dict_groupElem[ElemType(elemType)] = GroupElem(elemType, connect, coord)
        self.dim = max(g.dim for g in self.dict_groupElem.values())
        self.tags: dict = {}  # node sets by name (gmsh physical groups)
        used = np.zeros(self.Nn, dtype=bool)
        for g in self.Get_list_groupElem(self.dim):
            used[g.connect.ravel()] = True
        self.orphanNodes = np.nonzero(~used)[0]

    @property
    def Nn(self) -> int:
        return self.coord.shape[0]

    @property
    def nodes(self) -> np.ndarray:
        return np.arange(self.Nn)

    @property
    def groupElem(self) -> GroupElem:
        return self.Get_list_groupElem(self.dim)[0]

    @property
    def Ne(self) -> int:
        return self.groupElem.Ne

    @property
    def nPe(self) -> int:
        return self.groupElem.nPe

    @property
    def elemType(self) -> ElemType:
        return self.groupElem.elemType

    @property
    def connect(self) -> np.ndarray:
        return self.groupElem.connect

    @property
    def inDim(self) -> int:
        """Dimension of the space the mesh lies in: 3 if any z, else 2 if any
        y, else 1 (``FEM/_group_elem.py:155-175``)."""
        if self.dim == 3:
            return 3
        _, y, z = np.abs(self.coord.T)
        return 3 if z.max() > 0 else (2 if y.max() > 0 else 1)

    def Get_normals(self, nodes: Optional[np.ndarray] = None) -> tuple:
        """Unit normals at the boundary nodes (``FEM/_mesh.py:557-615``).

        For each group of dimension ``dim - 1``: the unit normal at every
        mass-rule point (``e_z x dx/dr`` on lines, ``dx/dr x dx/ds`` on faces,
        ``FEM/_group_elem.py:497-539``) summed over the element, averaged
        over the elements around each node whose every node is in ``nodes``,
        then normalised.  Which side the normal points to follows the node
        order of the boundary elements.  Returns ``(normals (n, 3), nodes)``,
        the nodes of each group sorted."""
        from ._utils import MatrixType
        from .quadrature import dN_pg

        nodes = self.nodes if nodes is None else np.asarray(nodes, dtype=np.int64)
        normals, used_all = [], []
        for g in self.Get_list_groupElem(self.dim - 1):
            elements = g.Get_Elements_Nodes(nodes, exclusively=True)
            if elements.size == 0:
                continue
            connect = g.connect[elements]
            dN = dN_pg(g.elemType, MatrixType.mass)  # (nPg, dim, nPe)
            X = self.coord[connect]
            dxdr = np.einsum("pn,end->epd", dN[:, 0], X)
            if g.dim == 1:
                n_ep = np.cross(np.array([0.0, 0.0, 1.0]), dxdr)
            else:
                n_ep = np.cross(dxdr, np.einsum("pn,end->epd", dN[:, 1], X))
            n_e = (n_ep / np.linalg.norm(n_ep, axis=2, keepdims=True)).sum(axis=1)
            used = np.unique(connect)
            acc = np.zeros((self.Nn, 3))
            cnt = np.zeros(self.Nn)
            for a in range(connect.shape[1]):
                np.add.at(acc, connect[:, a], n_e)
                np.add.at(cnt, connect[:, a], 1.0)
            n = acc[used] / cnt[used, None]
            normals.append(n / np.linalg.norm(n, axis=1, keepdims=True))
            used_all.append(used)
        if not normals:
            return np.zeros((0, 3)), np.array([], dtype=np.int64)
        return np.concatenate(normals), np.concatenate(used_all)

    def Get_list_groupElem(self, dim: Optional[int] = None) -> list:
        if dim is None:
            dim = self.dim
        return [g for g in self.dict_groupElem.values() if g.dim == dim]

    def Nodes_Tags(self, tags: list) -> np.ndarray:
        """Sorted nodes of the named node sets (``FEM/_mesh.py`` ``Nodes_Tags``)."""
        tags = [tags] if isinstance(tags, str) else list(tags)
        missing = [t for t in tags if t not in self.tags]
        assert not missing, f"unknown tags {missing}; available: {sorted(self.tags)}"
        if not tags:
            return np.array([], dtype=np.int64)
        return np.unique(np.concatenate([np.asarray(self.tags[t], dtype=np.int64) for t in tags]))

    def Save(self, folder: str, filename: str = "mesh") -> str:
        """Writes ``folder/filename.npz`` (coordinates, connectivities, tags);
        ``mesh_io.load_mesh`` reads it back."""
        from .mesh_io import save_mesh

        return save_mesh(self, folder, filename)

    def Nodes_Conditions(self, func: Callable) -> np.ndarray:
        """Nodes of the main-dimension groups meeting ``func(x, y, z)``
        (``EasyFEA/FEM/_mesh.py:619-643``)."""
        groups = self.Get_list_groupElem(self.dim)
        if len(groups) == 1:
            return groups[0].Get_Nodes_Conditions(func)
        nodes = set()
        for g in groups:
            nodes.update(g.Get_Nodes_Conditions(func).tolist())
        return np.array(sorted(nodes), dtype=np.int64)


# ---------------------------------------------------------------------------
# Structured generators
# ---------------------------------------------------------------------------


def _grid2d(x0, y0, Lx, Ly, nx, ny):
    xs = x0 + Lx * np.arange(nx + 1) / nx
    ys = y0 + Ly * np.arange(ny + 1) / ny
    X, Y = np.meshgrid(xs, ys)  # (ny+1, nx+1), x fastest when raveled
    coord = np.zeros(((nx + 1) * (ny + 1), 3))
    coord[:, 0] = X.ravel()
    coord[:, 1] = Y.ravel()
    # exact boundary values so that x == L style selections hit every node
    coord[X.ravel() == X.max(), 0] = x0 + Lx
    coord[Y.ravel() == Y.max(), 1] = y0 + Ly
    return coord


def _rect_edges(nx, ny, node):
    """SEG2 edges of the rectangle boundary (bottom, right, top, left)."""
    i = np.arange(nx)
    j = np.arange(ny)
    bottom = np.stack([node(i, 0), node(i + 1, 0)], 1)
    right = np.stack([node(nx, j), node(nx, j + 1)], 1)
    top = np.stack([node(i + 1, ny), node(i, ny)], 1)
    left = np.stack([node(0, j + 1), node(0, j)], 1)
    return np.concatenate([bottom, right, top, left])


def quad4_rectangle(Lx: float, Ly: float, nx: int, ny: int, x0=0.0, y0=0.0) -> Mesh:
    """Organised QUAD4 mesh of ``[x0, x0+Lx] x [y0, y0+Ly]``, ``nx x ny`` cells."""
    coord = _grid2d(x0, y0, Lx, Ly, nx, ny)

    def node(i, j):
        return np.asarray(i) + (nx + 1) * np.asarray(j)

    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    I, J = I.ravel(), J.ravel()
    connect = np.stack([node(I, J), node(I + 1, J), node(I + 1, J + 1), node(I, J + 1)], 1)
    edges = _rect_edges(nx, ny, node)
    return Mesh({ElemType.QUAD4: connect, ElemType.SEG2: edges}, coord)


def tri3_rectangle(Lx: float, Ly: float, nx: int, ny: int) -> Mesh:
    """TRI3 mesh: each cell split along its ``(i,j)-(i+1,j+1)`` diagonal."""
    return tri3_notched_square(nx, ny, Lx, Ly, notch=False)


def tri3_notched_square(
    nx: int, ny: Optional[int] = None, Lx: float = 1.0, Ly: float = 1.0, notch=True
) -> Mesh:
    """TRI3 single-edge-notched rectangle.

    Cells are split along the ``(i,j)-(i+1,j+1)`` diagonal.  With ``notch``
    the nodes on ``y = Ly/2`` with ``x < Lx/2`` are duplicated (the tip at
    ``x = Lx/2`` is not): elements below the line keep the original nodes,
    elements above use the copies, which are numbered after the grid nodes
    in increasing ``x``.  This is an open crack from the left edge to the
    centre, as the reference's ``Line(..., isOpen=True)`` crack produces.
    """
    ny = nx if ny is None else ny
    coord = _grid2d(0.0, 0.0, Lx, Ly, nx, ny)

    def node(i, j):
        return np.asarray(i) + (nx + 1) * np.asarray(j)

    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    I, J = I.ravel(), J.ravel()
    n00, n10, n11, n01 = node(I, J), node(I + 1, J), node(I + 1, J + 1), node(I, J + 1)
    lower = np.stack([n00, n10, n11], 1)
    upper = np.stack([n00, n11, n01], 1)
    # element order: per cell, lower then upper triangle
    connect = np.empty((2 * I.size, 3), dtype=np.int64)
    connect[0::2] = lower
    connect[1::2] = upper
    cellj = np.repeat(J, 2)
    edges = _rect_edges(nx, ny, node)

    if notch:
        assert nx % 2 == 0 and ny % 2 == 0, "notch needs even cell counts"
        jc, ic = ny // 2, nx // 2
        crack_nodes = node(np.arange(ic), jc)  # x < Lx/2 on y = Ly/2
        Ngrid = coord.shape[0]
        dup = Ngrid + np.arange(ic)
        coord = np.concatenate([coord, coord[crack_nodes]])
        remap = np.arange(Ngrid + ic)
        remap[crack_nodes] = dup
        above = cellj >= jc
        connect[above] = remap[connect[above]]
        # crack-lip edges: lower lip on the original nodes, upper lip on the copies
        i = np.arange(ic)
        lip_lo = np.stack([node(i, jc), node(i + 1, jc)], 1)
        lip_hi = remap[np.stack([node(i + 1, jc), node(i, jc)], 1)]
        # the tip node keeps its id on both lips
        edges = np.concatenate([edges, lip_lo, lip_hi])
        # top/left boundary edges above the crack use the copies
        left_above = (edges[:, 0] % (nx + 1) == 0) & (edges[:, 1] % (nx + 1) == 0)
        yl = np.minimum(edges[:, 0], edges[:, 1]) // (nx + 1)
        fix = left_above & (yl >= jc) & (edges[:, 0] < Ngrid) & (edges[:, 1] < Ngrid)
        edges[fix] = remap[edges[fix]]
    return Mesh({ElemType.TRI3: connect, ElemType.SEG2: edges}, coord)


def hexa8_box(
    Lx: float, Ly: float, Lz: float, nx: int, ny: int, nz: int
) -> Mesh:
    """Organised HEXA8 mesh of ``[0,Lx] x [0,Ly] x [0,Lz]`` with QUAD4 faces."""
    xs = Lx * np.arange(nx + 1) / nx
    ys = Ly * np.arange(ny + 1) / ny
    zs = Lz * np.arange(nz + 1) / nz
    xs[-1], ys[-1], zs[-1] = Lx, Ly, Lz
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    coord = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)

    def node(i, j, k):
        return np.asarray(i) + (nx + 1) * (np.asarray(j) + (ny + 1) * np.asarray(k))

    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    connect = np.stack(
        [
            node(I, J, K),
            node(I + 1, J, K),
            node(I + 1, J + 1, K),
            node(I, J + 1, K),
            node(I, J, K + 1),
            node(I + 1, J, K + 1),
            node(I + 1, J + 1, K + 1),
            node(I, J + 1, K + 1),
        ],
        1,
    )

    faces = []

    def quads(a, b, na, nb, f):
        A, B = np.meshgrid(np.arange(na), np.arange(nb), indexing="ij")
        A, B = A.ravel(), B.ravel()
        return np.stack([f(A, B), f(A + 1, B), f(A + 1, B + 1), f(A, B + 1)], 1)

    faces.append(quads(0, 0, nx, ny, lambda a, b: node(a, b, 0))[:, ::-1])  # z = 0
    faces.append(quads(0, 0, nx, ny, lambda a, b: node(a, b, nz)))  # z = Lz
    faces.append(quads(0, 0, nx, nz, lambda a, c: node(a, 0, c)))  # y = 0
    faces.append(quads(0, 0, nx, nz, lambda a, c: node(a, ny, c))[:, ::-1])  # y = Ly
    faces.append(quads(0, 0, ny, nz, lambda b, c: node(0, b, c))[:, ::-1])  # x = 0
    faces.append(quads(0, 0, ny, nz, lambda b, c: node(nx, b, c)))  # x = Lx
    return Mesh({ElemType.HEXA8: connect, ElemType.QUAD4: np.concatenate(faces)}, coord)


def hexa8_cube(n: int, L: float = 1.0) -> Mesh:
    return hexa8_box(L, L, L, n, n, n)


def tetra4_box(Lx: float, Ly: float, Lz: float, nx: int, ny: int, nz: int) -> Mesh:
    """TETRA4 mesh of ``[0,Lx] x [0,Ly] x [0,Lz]``: every cell of the
    ``nx x ny x nz`` grid split into the six tetrahedra along its main
    diagonal (the same diagonal direction in every cell, so faces conform),
    nodes numbered as in ``hexa8_box``, positively oriented."""
    grid = hexa8_box(Lx, Ly, Lz, nx, ny, nz)
    h = grid.dict_groupElem[ElemType.HEXA8].connect
    # cell corners v[a + 2b + 4c] at offsets (a, b, c)
    v = h[:, [0, 1, 3, 2, 4, 5, 7, 6]]
    tets = []
    for order in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        path, idx = [0], 0
        for ax in order:
            idx += 1 << ax
            path.append(idx)
        tets.append(v[:, path])
    connect = np.concatenate(tets)
    X = grid.coord[connect]
    vol = np.einsum("ei,ei->e", np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), X[:, 3] - X[:, 0])
    flip = vol < 0
    connect[flip] = connect[flip][:, [0, 2, 1, 3]]
    return Mesh({ElemType.TETRA4: connect}, grid.coord)


def tetra10_from_tetra4(mesh: Mesh) -> Mesh:
    """Quadratic tetrahedra on the same vertices: mid-edge nodes appended
    (one per edge), element order [v0, v1, v2, v3, m01, m12, m02, m03, m23, m13]
    (``EasyFEA/FEM/Elems/_tetra.py``)."""
    c4 = mesh.dict_groupElem[ElemType.TETRA4].connect
    pairs = [(0, 1), (1, 2), (0, 2), (0, 3), (2, 3), (1, 3)]
    e = np.stack([np.sort(c4[:, list(p)], axis=1) for p in pairs], 1).reshape(-1, 2)
    uniq, inv = np.unique(e, axis=0, return_inverse=True)
    mid = 0.5 * (mesh.coord[uniq[:, 0]] + mesh.coord[uniq[:, 1]])
    coord = np.concatenate([mesh.coord, mid])
    connect = np.concatenate([c4, mesh.Nn + inv.reshape(-1, 6)], 1)
    return Mesh({ElemType.TETRA10: connect}, coord)
