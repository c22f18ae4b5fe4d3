"""CPU oracle: an FP64 numpy/scipy restatement of EasyFEA's solve path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import this module, and only as the
checker or the timed CPU baseline.  The product package never imports it.

Every function restates the reference's algorithm from its source text
(citations are to files of the EasyFEA v3.5.1 tree); inputs are plain arrays
(``connect``, ``coord``) so that both this oracle and the device path consume
the same mesh.  Pinning (see ``tests/test_oracle_*.py``): closed forms,
invariants and the reference's own known-answer tests; the staggered loop
(``PhaseFieldOracle.Solve``, every convergence option) by per-pass residual
identities of independently assembled systems, with planted slips shown to
fail them (``tests/test_oracle_staggered.py``).  The 3D spectral
decomposition at exactly repeated principal strains is pinned against numpy's
eigh where the reference's case masks give the spectral split (all equal; the
two smallest equal and not positive); where they do not (the two smallest
equal and positive, the two largest equal: half-split or misshifted
projectors, ``Models/_phasefield.py:854-893, 942-948``) the oracle keeps the
reference's formulas and only the eigenvalues are pinned: parity unpinned
there (the device has no 3D spectral kernel).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

SQRT2 = np.sqrt(2.0)

# ---------------------------------------------------------------------------
# a1-a2: quadrature and shape functions
# ---------------------------------------------------------------------------


def gauss(elemType: str, matrixType: str):
    """Points/weights, ``EasyFEA/FEM/_gauss.py:364-516`` (SEG2 :373-382,
    TRI3 :423-430 with ``_Triangle`` :55-70, QUAD4 :449-451 with
    ``_Quadrangle`` :111-121, TETRA4/10 :467-476 with ``_Tetrahedron``
    :142-163, HEXA8 :478-480 with ``_Hexahedron`` :201-213)."""
    rigi = matrixType == "rigi"
    if elemType == "SEG2":
        x, w = np.polynomial.legendre.leggauss(1 if rigi else 2)
        return x.reshape(-1, 1), w
    if elemType == "TRI3":
        if rigi:
            return np.array([[1 / 3, 1 / 3]]), np.array([1 / 2])
        ksis = [1 / 6, 2 / 3, 1 / 6]
        etas = [1 / 6, 1 / 6, 2 / 3]
        return np.array([ksis, etas]).T, np.array([1 / 6] * 3)
    if elemType == "QUAD4":
        a = 1 / np.sqrt(3)
        return np.array([[-a, a, a, -a], [-a, -a, a, a]]).T, np.ones(4)
    if elemType in ("TETRA4", "TETRA10"):
        if elemType == "TETRA4" and rigi:
            return np.array([[1 / 4, 1 / 4, 1 / 4]]), np.array([1 / 6])
        a = (5 - np.sqrt(5)) / 20
        b = (5 + 3 * np.sqrt(5)) / 20
        return np.array([[a, a, a, b], [a, a, b, a], [a, b, a, a]]).T, np.full(4, 1 / 24)
    if elemType == "HEXA8":
        a = 1 / np.sqrt(3)
        x = [-a, -a, -a, -a, a, a, a, a]
        y = [-a, -a, a, a, -a, -a, a, a]
        z = [-a, a, -a, a, -a, a, -a, a]
        return np.array([x, y, z]).T, np.ones(8)
    raise NotImplementedError(elemType)


def shape_functions(elemType: str, pts: np.ndarray):
    """``(N (nPg, nPe), dN (nPg, dim, nPe))`` from the element files
    ``EasyFEA/FEM/Elems/_seg.py:46-61``, ``_tri.py:53-69``, ``_quad.py:56-75``,
    ``_tetra.py:178-222`` (TETRA10), ``_hexa.py:86-147``."""
    Ns, dNs = [], []
    for p in pts:
        if elemType == "SEG2":
            (r,) = p
            N = [-(r - 1) / 2, (r + 1) / 2]
            dN = [[-1 / 2, 1 / 2]]
        elif elemType == "TRI3":
            r, s = p
            N = [-r - s + 1, r, s]
            dN = [[-1, 1, 0], [-1, 0, 1]]
        elif elemType == "QUAD4":
            r, s = p
            N = [(r - 1) * (s - 1) / 4, -(r + 1) * (s - 1) / 4, (r + 1) * (s + 1) / 4, -(r - 1) * (s + 1) / 4]
            dN = [
                [s / 4 - 1 / 4, 1 / 4 - s / 4, s / 4 + 1 / 4, -s / 4 - 1 / 4],
                [r / 4 - 1 / 4, -r / 4 - 1 / 4, r / 4 + 1 / 4, 1 / 4 - r / 4],
            ]
        elif elemType == "TETRA4":
            r, s, t = p
            N = [-r - s - t + 1, r, s, t]
            dN = [[-1, 1, 0, 0], [-1, 0, 1, 0], [-1, 0, 0, 1]]
        elif elemType == "TETRA10":
            r, s, t = p
            N = [
                (r + s + t - 1) * (2 * r + 2 * s + 2 * t - 1), r * (2 * r - 1), s * (2 * s - 1),
                t * (2 * t - 1), -4 * r * (r + s + t - 1), 4 * r * s, -4 * s * (r + s + t - 1),
                -4 * t * (r + s + t - 1), 4 * s * t, 4 * r * t,
            ]
            c = 4 * r + 4 * s + 4 * t - 3
            dN = np.array([
                [c, c, c], [4 * r - 1, 0, 0], [0, 4 * s - 1, 0], [0, 0, 4 * t - 1],
                [-8 * r - 4 * s - 4 * t + 4, -4 * r, -4 * r], [4 * s, 4 * r, 0],
                [-4 * s, -4 * r - 8 * s - 4 * t + 4, -4 * s], [-4 * t, -4 * t, -4 * r - 4 * s - 8 * t + 4],
                [0, 4 * t, 4 * s], [4 * t, 0, 4 * r],
            ]).T
        elif elemType == "HEXA8":
            r, s, t = p
            N = [
                -(r - 1) * (s - 1) * (t - 1) / 8, (r + 1) * (s - 1) * (t - 1) / 8,
                -(r + 1) * (s + 1) * (t - 1) / 8, (r - 1) * (s + 1) * (t - 1) / 8,
                (r - 1) * (s - 1) * (t + 1) / 8, -(r + 1) * (s - 1) * (t + 1) / 8,
                (r + 1) * (s + 1) * (t + 1) / 8, -(r - 1) * (s + 1) * (t + 1) / 8,
            ]
            dN = np.array([
                [-(s - 1) * (t - 1) / 8, -(r - 1) * (t - 1) / 8, -(r - 1) * (s - 1) / 8],
                [(s - 1) * (t - 1) / 8, (r + 1) * (t - 1) / 8, (r + 1) * (s - 1) / 8],
                [-(s + 1) * (t - 1) / 8, -(r + 1) * (t - 1) / 8, -(r + 1) * (s + 1) / 8],
                [(s + 1) * (t - 1) / 8, (r - 1) * (t - 1) / 8, (r - 1) * (s + 1) / 8],
                [(s - 1) * (t + 1) / 8, (r - 1) * (t + 1) / 8, (r - 1) * (s - 1) / 8],
                [-(s - 1) * (t + 1) / 8, -(r + 1) * (t + 1) / 8, -(r + 1) * (s - 1) / 8],
                [(s + 1) * (t + 1) / 8, (r + 1) * (t + 1) / 8, (r + 1) * (s + 1) / 8],
                [-(s + 1) * (t + 1) / 8, -(r - 1) * (t + 1) / 8, -(r - 1) * (s + 1) / 8],
            ]).T
        else:
            raise NotImplementedError(elemType)
        Ns.append(np.asarray(N, dtype=float))
        dNs.append(np.asarray(dN, dtype=float))
    return np.array(Ns), np.array(dNs)


ELEM_DIM = {"SEG2": 1, "TRI3": 2, "QUAD4": 2, "TETRA4": 3, "TETRA10": 3, "HEXA8": 3}

# ---------------------------------------------------------------------------
# a3-a4: Jacobian, dN_x, B (Kelvin-Mandel)
# ---------------------------------------------------------------------------


def _det(F):
    """``EasyFEA/FEM/_linalg.py:533-573``."""
    d = F.shape[-1]
    if d == 1:
        return F[..., 0, 0]
    if d == 2:
        return F[..., 0, 0] * F[..., 1, 1] - F[..., 1, 0] * F[..., 0, 1]
    a11, a12, a13 = F[..., 0, 0], F[..., 0, 1], F[..., 0, 2]
    a21, a22, a23 = F[..., 1, 0], F[..., 1, 1], F[..., 1, 2]
    a31, a32, a33 = F[..., 2, 0], F[..., 2, 1], F[..., 2, 2]
    return (a11 * (a22 * a33 - a32 * a23) - a12 * (a21 * a33 - a31 * a23)
            + a13 * (a21 * a32 - a31 * a22))


def _inv(F):
    """Adjugate inverse, ``EasyFEA/FEM/_linalg.py:576-656``."""
    d = F.shape[-1]
    det = _det(F)
    if d == 1:
        return 1 / F
    adj = np.zeros_like(F)
    if d == 2:
        adj[..., 0, 0] = F[..., 1, 1]
        adj[..., 0, 1] = -F[..., 0, 1]
        adj[..., 1, 0] = -F[..., 1, 0]
        adj[..., 1, 1] = F[..., 0, 0]
    else:
        T = np.swapaxes(F, -1, -2)
        a = [[T[..., i, j] for j in range(3)] for i in range(3)]
        adj[..., 0, 0] = a[1][1] * a[2][2] - a[2][1] * a[1][2]
        adj[..., 0, 1] = -(a[1][0] * a[2][2] - a[2][0] * a[1][2])
        adj[..., 0, 2] = a[1][0] * a[2][1] - a[2][0] * a[1][1]
        adj[..., 1, 0] = -(a[0][1] * a[2][2] - a[2][1] * a[0][2])
        adj[..., 1, 1] = a[0][0] * a[2][2] - a[2][0] * a[0][2]
        adj[..., 1, 2] = -(a[0][0] * a[2][1] - a[2][0] * a[0][1])
        adj[..., 2, 0] = a[0][1] * a[1][2] - a[1][1] * a[0][2]
        adj[..., 2, 1] = -(a[0][0] * a[1][2] - a[1][0] * a[0][2])
        adj[..., 2, 2] = a[0][0] * a[1][1] - a[1][0] * a[0][1]
    return np.einsum("...,...ij->...ij", 1 / det, adj)


def _local_frame_coords(coord_e: np.ndarray, dim: int) -> np.ndarray:
    """Coordinates of lower-dimensional elements in their own plane/line.

    ``_group_elem.py:841-865`` rebases a boundary element into an orthonormal
    frame before forming the Jacobian; for straight SEG2 and planar QUAD4
    faces this preserves lengths/areas, which is what the load integral uses.
    """
    if dim == 1:
        t = coord_e[:, 1] - coord_e[:, 0]
        t = t / np.linalg.norm(t, axis=1, keepdims=True)
        return np.einsum("end,ed->en", coord_e, t)[..., None]
    # dim == 2 in 3D: i along edge 0-1, k normal, j = k x i
    i = coord_e[:, 1] - coord_e[:, 0]
    i /= np.linalg.norm(i, axis=1, keepdims=True)
    k = np.cross(i, coord_e[:, -1] - coord_e[:, 0])
    k /= np.linalg.norm(k, axis=1, keepdims=True)
    j = np.cross(k, i)
    return np.stack([np.einsum("end,ed->en", coord_e, i), np.einsum("end,ed->en", coord_e, j)], -1)


class Geometry:
    """Per-group geometric blocks for one integration rule.

    ``F`` (``Get_F_e_pg`` :832-869), ``detJ`` (``Get_jacobian_e_pg``
    :871-888), ``wJ`` (:890-900), ``invF`` (:902-915), ``dN_x``
    (``Get_dN_e_pg`` :1083-1105), ``B`` (``Get_B_e_pg`` :1241-1312) of
    ``EasyFEA/FEM/_group_elem.py``.
    """

    def __init__(self, elemType: str, connect: np.ndarray, coord: np.ndarray, matrixType: str, inDim: int):
        self.elemType = elemType
        self.dim = dim = ELEM_DIM[elemType]
        pts, w = gauss(elemType, matrixType)
        self.N_pg, self.dN_pg = shape_functions(elemType, pts)
        self.weights = w
        coord_e = coord[connect][..., :3]
        if dim == inDim:
            X = coord_e[..., :dim]
        else:
            X = _local_frame_coords(coord_e, dim)
        self.F = np.einsum("pdn,enk->epdk", self.dN_pg, X)
        self.detJ = np.abs(_det(self.F))
        self.wJ = self.detJ * w
        self.invF = _inv(self.F)
        self.dN_x = np.einsum("epij,pjn->epin", self.invF, self.dN_pg)
        self.Ne, self.nPg = self.wJ.shape
        self.nPe = connect.shape[1]

    def B(self) -> np.ndarray:
        dN = self.dN_x
        Ne, nPg, dim, nPe = dN.shape
        c = 1 / np.sqrt(2)
        if dim == 2:
            B = np.zeros((Ne, nPg, 3, nPe * 2))
            B[:, :, 0, 0::2] = dN[:, :, 0]
            B[:, :, 1, 1::2] = dN[:, :, 1]
            B[:, :, 2, 0::2] = dN[:, :, 1] * c
            B[:, :, 2, 1::2] = dN[:, :, 0] * c
        else:
            B = np.zeros((Ne, nPg, 6, nPe * 3))
            dx, dy, dz = dN[:, :, 0], dN[:, :, 1], dN[:, :, 2]
            B[:, :, 0, 0::3] = dx
            B[:, :, 1, 1::3] = dy
            B[:, :, 2, 2::3] = dz
            B[:, :, 3, 1::3] = dz * c
            B[:, :, 3, 2::3] = dy * c
            B[:, :, 4, 0::3] = dz * c
            B[:, :, 4, 2::3] = dx * c
            B[:, :, 5, 0::3] = dy * c
            B[:, :, 5, 1::3] = dx * c
        return B


# ---------------------------------------------------------------------------
# a5: isotropic Hooke tensor in Kelvin-Mandel notation
# ---------------------------------------------------------------------------


def lame(E, v, dim: int, planeStress: bool):
    """``get_lambda``/``get_mu``/``get_bulk``, ``Models/Elastic/_laws.py:376-405``."""
    E = np.asarray(E, dtype=float)
    lmbda = E * v / ((1 + v) * (1 - 2 * v))
    if dim == 2 and planeStress:
        lmbda = E * v / (1 - v**2)
    mu = E / (2 * (1 + v))
    bulk = lmbda + 2 * mu / dim
    return lmbda, mu, bulk


def C_isotropic(E, v, dim: int, planeStress: bool) -> np.ndarray:
    """``Isotropic._Behavior`` (``_laws.py:407-482``) followed by
    ``KelvinMandel_Matrix`` (``Models/_utils.py:111-152``).  ``E`` scalar or
    ``(Ne,)`` -> ``(3,3)``/``(6,6)`` or ``(Ne, 3, 3)``/``(Ne, 6, 6)``."""
    lmbda, mu, _ = lame(E, v, dim, planeStress)
    shape = np.shape(lmbda)
    if dim == 2:
        C = np.zeros(shape + (3, 3))
        C[..., 0, 0] = C[..., 1, 1] = lmbda + 2 * mu
        C[..., 0, 1] = C[..., 1, 0] = lmbda
        C[..., 2, 2] = mu
        T = np.array([[1, 1, SQRT2], [1, 1, SQRT2], [SQRT2, SQRT2, 2]])
    else:
        C = np.zeros(shape + (6, 6))
        for i in range(3):
            for j in range(3):
                C[..., i, j] = lmbda
            C[..., i, i] = lmbda + 2 * mu
            C[..., 3 + i, 3 + i] = mu
        T = np.ones((6, 6))
        T[:3, 3:] = T[3:, :3] = SQRT2
        T[3:, 3:] = 2
    return C * T


# ---------------------------------------------------------------------------
# a6-a9: element operators
# ---------------------------------------------------------------------------


def _bcast(coef, Ne, nPg):
    """``FeArray.broadcast`` (``FEM/_linalg.py:426-476``) for scalar,
    ``(Ne,)`` and ``(Ne, nPg)`` coefficients."""
    coef = np.asarray(coef, dtype=float)
    if coef.ndim == 0:
        return np.full((Ne, nPg), float(coef))
    if coef.shape == (Ne, nPg):
        return coef
    if coef.shape == (Ne,):
        return np.broadcast_to(coef[:, None], (Ne, nPg))
    if coef.shape == (nPg,):
        return np.broadcast_to(coef[None, :], (Ne, nPg))
    raise ValueError(coef.shape)


def kelvin_mandel(dim: int, M: np.ndarray) -> np.ndarray:
    """``KelvinMandel_Matrix`` (``Models/_utils.py:111-152``): entries of the
    normal-shear blocks times sqrt(2), of the shear-shear block times 2."""
    r2 = np.sqrt(2)
    n = 3 if dim == 2 else 6
    k = 2 if dim == 2 else 3
    T = np.ones((n, n))
    T[:k, k:] = T[k:, :k] = r2
    T[k:, k:] = 2
    return M * T


def rotation_mandel(axis_1, axis_2) -> np.ndarray:
    """``Get_Pmat(axis_1, axis_2, useMandel=True)`` (``Models/_utils.py:
    433-589``) for one pair of 3D axes: ``C_global = P C_material P^T``.
    p holds the axes as columns (axis_3 = axis_1 x axis_2); P = [[p**2,
    sqrt2 A], [sqrt2 B, D2]] with A, B, D2 the reference's products of the
    direction cosines (Chevalier 1988).  The axes are scaled by their norm as
    the reference's einsum does (the caller passes unit vectors)."""
    a1 = np.asarray(axis_1, float)
    a2 = np.asarray(axis_2, float)
    a1 = a1 * np.linalg.norm(a1)
    a2 = a2 * np.linalg.norm(a2)
    assert abs(a1 @ a2) <= 1e-12, "Must give perpendicular axes"
    a3 = np.cross(a1, a2)
    p = np.stack([a1, a2, a3], axis=1)
    (p11, p12, p13), (p21, p22, p23), (p31, p32, p33) = a1, a2, a3
    A = np.array([[p21 * p31, p11 * p31, p11 * p21],
                  [p22 * p32, p12 * p32, p12 * p22],
                  [p23 * p33, p13 * p33, p13 * p23]])
    B = np.array([[p12 * p13, p22 * p23, p32 * p33],
                  [p11 * p13, p21 * p23, p31 * p33],
                  [p11 * p12, p21 * p22, p31 * p32]])
    D2 = np.array([[p22 * p33 + p32 * p23, p12 * p33 + p32 * p13, p12 * p23 + p22 * p13],
                   [p21 * p33 + p31 * p23, p11 * p33 + p31 * p13, p11 * p23 + p21 * p13],
                   [p21 * p32 + p31 * p22, p11 * p32 + p31 * p12, p11 * p22 + p21 * p12]])
    r2 = np.sqrt(2)
    return np.block([[p ** 2, r2 * A], [r2 * B, D2]])


def C_anisotropic(C, useVoigtNotation: bool, axis_1=(1, 0, 0), axis_2=(0, 1, 0)) -> np.ndarray:
    """``Anisotropic._Behavior`` (``Models/Elastic/_laws.py:1190-1259``) for
    a 3D law (``C`` of shape (6, 6) or (Ne, 6, 6)): the symmetry check, then
    ``P C P^T`` with P of the normalised axes.  As the reference does, a 6 x 6
    ``C`` enters the rotation as given: its Kelvin-Mandel scaling
    (``useVoigtNotation``) only reaches 3 x 3 inputs (``:1230-1246``)."""
    C = np.asarray(C, float)
    assert C.shape[-2:] == (6, 6), "a 3D law takes a (6, 6) matrix"
    Ct = np.swapaxes(C, -1, -2)
    assert np.max(np.linalg.norm(Ct - C, axis=(-2, -1)) / np.linalg.norm(C, axis=(-2, -1))) <= 1e-12
    a1 = np.asarray(axis_1, float) / np.linalg.norm(axis_1)
    a2 = np.asarray(axis_2, float) / np.linalg.norm(axis_2)
    P = rotation_mandel(a1, a2)
    return np.einsum("ij,...jk,lk->...il", P, C, P)


def K_elastic(geo: Geometry, C: np.ndarray) -> np.ndarray:
    """``Bilinear.LinearizedElasticity`` (``FEM/Operators/Bilinear.py:62-79``):
    ``einsum('epij,epjk->eik', (wJ*B^T) @ C, B)``.  ``C`` may be
    ``(ns,ns)``, ``(Ne,ns,ns)`` or ``(Ne,nPg,ns,ns)``."""
    B = geo.B()
    left = geo.wJ[..., None, None] * np.swapaxes(B, -1, -2)
    C = np.asarray(C, dtype=float)
    if C.ndim == 3:
        C = C[:, None]
    return np.einsum("epij,epjk->eik", left @ C, B, optimize=True)


def UV(geo: Geometry, coef=1.0, dof_n: int = 1) -> np.ndarray:
    """``Bilinear.UV`` (:42-59) with ``Get_ReactionPart_e_pg``
    (``_group_elem.py:1337-1360``): ``sum_p coef*wJ*N^T N``."""
    N = geo.N_pg
    nPg, nPe = N.shape
    if dof_n > 1:
        Nr = np.zeros((nPg, dof_n, nPe * dof_n))
        for r in range(dof_n):
            Nr[:, r, r::dof_n] = N
    else:
        Nr = N[:, None, :]
    part = geo.wJ[..., None, None] * np.einsum("pki,pkj->pij", Nr, Nr)[None]
    c = _bcast(coef, geo.Ne, geo.nPg)
    return (c[..., None, None] * part).sum(axis=1)


def GradU_A_GradV(geo: Geometry, A: np.ndarray, coef=1.0) -> np.ndarray:
    """``Bilinear.GradU_A_GradV`` (:229-249) with ``Get_DiffusePart_e_pg``
    (``_group_elem.py:1362-1380``)."""
    dN = geo.dN_x
    part = geo.wJ[..., None, None] * np.swapaxes(dN, -1, -2)
    c = _bcast(coef, geo.Ne, geo.nPg)
    return np.einsum("epij,epjk->eik", c[..., None, None] * part, np.asarray(A) @ dN, optimize=True)


def V(geo: Geometry, f=1.0) -> np.ndarray:
    """``Linear.V`` (``FEM/Operators/Linear.py:18-35``) with
    ``Get_SourcePart_e_pg`` (``_group_elem.py:1382-1407``), ``dof_n = 1``."""
    part = geo.wJ[..., None] * geo.N_pg[None]
    c = _bcast(f, geo.Ne, geo.nPg)
    return (c[..., None] * part).sum(axis=1)


# ---------------------------------------------------------------------------
# a10-a12: DOF maps and the CSR map / bincount scatter
# ---------------------------------------------------------------------------


def assembly_e(connect: np.ndarray, dof_n: int) -> np.ndarray:
    """``_GroupElem._Get_assembly_e`` (``_group_elem.py:358-380``)."""
    Ne, nPe = connect.shape
    out = np.zeros((Ne, nPe * dof_n), dtype=np.int64)
    for d in range(dof_n):
        out[:, np.arange(d, nPe * dof_n, dof_n)] = np.asarray(connect) * dof_n + d
    return out


def rows_cols_e(connect: np.ndarray, dof_n: int):
    """``Get_rows_e``/``Get_columns_e`` (``_group_elem.py:382-402``)."""
    a = assembly_e(connect, dof_n)
    Ne, nd = a.shape
    rows = np.repeat(a, nd).reshape(Ne, nd * nd)
    cols = np.repeat(a, nd, axis=0).reshape(Ne, nd * nd)
    return rows, cols


def csr_map(connect: np.ndarray, dof_n: int, Ndof: int, isMatrix=True):
    """``_Simu.__Get_csr_map`` (``Simulations/_simu.py:1062-1102``): scipy's
    canonical pattern and ``inv[k]`` = slot of element entry ``k``."""
    shape = (Ndof, Ndof) if isMatrix else (Ndof, 1)
    if isMatrix:
        rows, cols = rows_cols_e(connect, dof_n)
        rows, cols = rows.ravel(), cols.ravel()
    else:
        rows = assembly_e(connect, dof_n).ravel()
        cols = np.zeros_like(rows)
    m = sp.csr_matrix((np.ones(rows.size), (rows, cols)), shape=shape)
    m.sort_indices()
    ncol = Ndof if isMatrix else 1
    canon = np.repeat(np.arange(shape[0]), np.diff(m.indptr)) * ncol + m.indices
    inv = np.searchsorted(canon, rows.astype(np.int64) * ncol + cols).astype(np.int32)
    return inv, m.indices, m.indptr, m.nnz


def assemble_csr(values_e: np.ndarray, cmap, Ndof: int, isMatrix=True) -> sp.csr_matrix:
    """``_Simu.__Assemble_csr`` (``_simu.py:989-1060``): one ``bincount`` in
    element-entry order, canonical CSR out."""
    inv, indices, indptr, nnz = cmap
    shape = (Ndof, Ndof) if isMatrix else (Ndof, 1)
    data = np.bincount(inv, weights=np.asarray(values_e).ravel(), minlength=nnz)
    m = sp.csr_matrix((data, indices, indptr), shape=shape)
    m.has_canonical_format = True
    return m


# ---------------------------------------------------------------------------
# a13-a16: boundary conditions and the r1 solve
# ---------------------------------------------------------------------------


def elements_using_nodes(connect: np.ndarray, nodes: np.ndarray) -> np.ndarray:
    """Elements whose every node lies in ``nodes``
    (``Get_Elements_Nodes(exclusively=True)``, ``_group_elem.py:1500-1537``)."""
    mark = np.zeros(max(connect.max(), np.max(nodes)) + 1, dtype=bool)
    mark[nodes] = True
    return np.nonzero(mark[connect].all(axis=1))[0]


def dofs_nodes(nodes, unknowns, available) -> np.ndarray:
    """``BoundaryCondition.Get_dofs_nodes`` (``FEM/_boundary_conditions.py:160-201``)."""
    nodes = np.asarray(nodes, dtype=int).ravel()
    dim = len(available)
    out = np.zeros((nodes.size, len(unknowns)), dtype=int)
    for d, u in enumerate(unknowns):
        out[:, d] = nodes * dim + available.index(u)
    return out.ravel()


def integrate_load(elemType, connect_b, coord, inDim, nodes, values, unknowns, available):
    """``_Simu.__Bc_Integration_Dim`` (``_simu.py:2770-2842``): ``sum_p wJ *
    f * N`` per element node, mass points.  A constant or callable value is
    evaluated at the Gauss points; an array holds one value per node of
    ``nodes`` and is interpolated with ``N`` (``:2818-2826``)."""
    elements = elements_using_nodes(connect_b, nodes)
    if elements.size == 0:
        return np.array([]), np.array([], dtype=int)
    connect = connect_b[elements]
    geo = Geometry(elemType, connect, coord, "mass", inDim)
    coord_e_p = np.einsum("pn,end->epd", geo.N_pg, coord[connect])
    vals = np.zeros((connect.shape[0] * connect.shape[1], len(unknowns)))
    dofs = np.zeros_like(vals, dtype=int)
    for u, unk in enumerate(unknowns):
        val = values[u]
        if callable(val):
            ev = val(coord_e_p[..., 0], coord_e_p[..., 1], coord_e_p[..., 2]) * np.ones(geo.wJ.shape)
        elif np.ndim(val) == 0:
            ev = np.full(geo.wJ.shape, float(val))
        else:
            ev_n = np.zeros(coord.shape[0])
            ev_n[np.asarray(nodes)] = val
            ev = np.einsum("pn,en->ep", geo.N_pg, ev_n[connect])
        v_e_p = np.einsum("ep,ep,pn->epn", geo.wJ, ev, geo.N_pg)
        vals[:, u] = v_e_p.sum(axis=1).ravel()
        dofs[:, u] = dofs_nodes(connect.ravel(), [unk], available)
    return vals.ravel(), dofs.ravel()


def node_normals(elemType, connect_b, coord, nodes):
    """``Mesh.Get_normals`` (``FEM/_mesh.py:557-615``) over one boundary
    group: unit normals at the mass points (``Get_normals_e_pg``,
    ``FEM/_group_elem.py:497-539``: ``e_z x dx/dr`` on lines, ``dx/dr x
    dx/ds`` on faces), summed per element, averaged over the selected
    elements around each node and normalised.  Returns ``(normals (n, 3),
    nodes)`` with the nodes sorted."""
    elements = elements_using_nodes(connect_b, nodes)
    if elements.size == 0:
        return np.zeros((0, 3)), np.array([], dtype=int)
    connect = connect_b[elements]
    pts, _ = gauss(elemType, "mass")
    _, dN = shape_functions(elemType, pts)
    X = coord[connect][..., :3]
    dxdr = np.einsum("pn,end->epd", dN[:, 0], X)
    if ELEM_DIM[elemType] == 1:
        n_ep = np.cross(np.array([0.0, 0.0, 1.0]), dxdr)
    else:
        n_ep = np.cross(dxdr, np.einsum("pn,end->epd", dN[:, 1], X))
    n_ep = n_ep / np.linalg.norm(n_ep, axis=2, keepdims=True)
    n_e = n_ep.sum(axis=1)
    used = np.unique(connect)
    acc = np.zeros((coord.shape[0], 3))
    cnt = np.zeros(coord.shape[0])
    for a in range(connect.shape[1]):
        np.add.at(acc, connect[:, a], n_e)
        np.add.at(cnt, connect[:, a], 1.0)
    n = acc[used] / cnt[used, None]
    return n / np.linalg.norm(n, axis=1, keepdims=True), used


def known_unknown(dirichlet_dofs: np.ndarray, Ndof: int):
    """``Bc_dofs_known_unknown`` (``_simu.py:2295-2316``)."""
    mask = np.ones(Ndof, dtype=bool)
    mask[np.asarray(dirichlet_dofs, dtype=int)] = False
    return np.nonzero(~mask)[0], np.nonzero(mask)[0]


def bounded_lsq(A: sp.csr_matrix, b: np.ndarray, lb: np.ndarray, ub: np.ndarray) -> np.ndarray:
    """The BoundConstrain damage solve (``Simulations/Solvers.py:376-383``):
    ``scipy.optimize.lsq_linear(A, b, bounds=(lb, ub), tol=1e-10,
    method="trf")``, the reference's own call of its scipy dependency."""
    from scipy import optimize

    res = optimize.lsq_linear(sp.csr_matrix(A), np.asarray(b, float).ravel(), bounds=(lb, ub), tol=1e-10,
                              method="trf", verbose=0)
    return np.array(res["x"])


def _lsq_kkt(H: np.ndarray, g: np.ndarray, lb: np.ndarray, ub: np.ndarray, x: np.ndarray) -> float:
    """Largest violation of the first-order conditions of
    ``min 1/2 |A x - b|^2`` over the box (``H = A^T A``, ``g = A^T b``):
    gradient zero on free variables, >= 0 at the lower bound, <= 0 at the
    upper bound; relative to ``max |g|``."""
    grad = H @ x - g
    at_lb, at_ub = x <= lb, x >= ub
    free = ~(at_lb | at_ub)
    v = max(np.abs(grad[free]).max(initial=0.0), (-grad[at_lb]).max(initial=0.0), grad[at_ub].max(initial=0.0))
    return v / max(np.abs(g).max(), 1e-300)


def bounded_lsq_exact(A, b: np.ndarray, lb: np.ndarray, ub: np.ndarray) -> np.ndarray:
    """The minimiser the reference's call approximates: ``lsq_linear`` with
    the bounded-variable least-squares active-set method (``method="bvls"``,
    Stark and Parker 1995) on the dense matrix, tolerance 1e-14, then
    primal-dual active-set steps (Hintermueller, Ito and Kunisch 2002) on the
    normal equations, each solving the free variables exactly, until the
    active sets repeat; the first-order conditions are checked to 1e-11.
    BVLS alone can stop on a small cost change short of the minimum (an AT1
    damage system of 637 unknowns: cost 8 % above it, d off by 0.55).  The
    reference's trust-region call stops on a relative cost change of 1e-10,
    which on ill-conditioned damage systems leaves it measurably short of
    this point (percent-level in x, 1e-3 in the cost on a 1D AT1-like system)."""
    from scipy import optimize

    Ad = A.toarray() if sp.issparse(A) else np.asarray(A, float)
    b = np.asarray(b, float).ravel()
    lb, ub = np.asarray(lb, float), np.asarray(ub, float)
    res = optimize.lsq_linear(Ad, b, bounds=(lb, ub), method="bvls", tol=1e-14, verbose=0)
    x = np.clip(np.array(res["x"]), lb, ub)
    H, g = Ad.T @ Ad, Ad.T @ b
    hd = np.diag(H)
    lam = H @ x - g
    sets = None
    for _ in range(500):
        t = x - lam / hd
        AL, AU = t <= lb, t >= ub
        F = ~(AL | AU)
        new = (AL.tobytes(), AU.tobytes())
        if new == sets:
            break
        sets = new
        x = np.where(AL, lb, np.where(AU, ub, 0.0))
        if F.any():
            x[F] = np.linalg.solve(H[np.ix_(F, F)], g[F] - H[np.ix_(F, ~F)] @ x[~F])
        lam = H @ x - g
        lam[F] = 0.0
    x = np.clip(x, lb, ub)
    kkt = _lsq_kkt(H, g, lb, ub, x)
    if kkt > 1e-11:
        raise RuntimeError(f"bounded_lsq_exact: first-order conditions violated by {kkt:.2e}")
    return x


def solve_r1(K: sp.csr_matrix, b: np.ndarray, dir_dofs, dir_vals, Ndof, orphan_dofs=()):
    """``__Solver_1`` (``Simulations/Solvers.py:502-553``) with
    ``__Solver_Get_Dirichlet_A_x`` (``_simu.py:1916-1984``): reduce to the
    unknown DOFs and solve with ``scipy.sparse.linalg.spsolve``
    (``Solvers.py:361-362``)."""
    A = K
    if len(orphan_dofs):
        diag = np.zeros(Ndof)
        diag[np.asarray(orphan_dofs)] = 1.0
        A = A + sp.diags(diag, format="csr")
    dir_dofs = np.asarray(dir_dofs, dtype=int)
    x = sp.csr_matrix((np.asarray(dir_vals, float), (dir_dofs, np.zeros_like(dir_dofs))), shape=(Ndof, 1))
    known, unknown = known_unknown(dir_dofs, Ndof)
    Ai = A[unknown, :].tocsc()
    Aii = Ai[:, unknown]
    Aic = Ai[:, known]
    bs = sp.csr_matrix(np.asarray(b, float).reshape(-1, 1))
    bi = bs[unknown, 0]
    xc = x[known, 0]
    bi = bi - Aic @ xc
    xi = sla.spsolve(Aii.tocsr(), bi)
    xo = x.toarray().reshape(Ndof)
    xo[unknown] = xi
    return xo


# ---------------------------------------------------------------------------
# a19-a23: phase-field model
# ---------------------------------------------------------------------------


def IxI(dim):
    """``PhaseField.__Build_IxI`` (``Models/_phasefield.py:218-230``)."""
    I = np.array([1.0, 1.0, 0.0]) if dim == 2 else np.array([1.0, 1, 1, 0, 0, 0])
    return np.outer(I, I)


def Rp_Rm(eps: np.ndarray, dim: int):
    """``__Rp_Rm`` (:485-501)."""
    tr = eps[..., 0] + eps[..., 1]
    if dim == 3:
        tr = tr + eps[..., 2]
    return (1 + np.sign(tr)) / 2, (1 + np.sign(-tr)) / 2


def split_amor(eps, E, v, dim, planeStress):
    """``__Split_Amor`` (:451-483).  ``E`` scalar or broadcastable to the
    leading ``(Ne, nPg)`` axes of ``eps``."""
    _, mu, bulk = lame(E, v, dim, planeStress)
    Rp, Rm = Rp_Rm(eps, dim)
    II = IxI(dim)
    n = II.shape[0]
    mu = np.asarray(mu)[..., None, None] if np.ndim(mu) else mu
    bulk = np.asarray(bulk)[..., None, None] if np.ndim(bulk) else bulk
    cP = bulk * (Rp[..., None, None] * II) + 2 * mu * (np.eye(n) - 1 / dim * II)
    cM = bulk * (Rm[..., None, None] * II)
    return cP, cM


def vec_to_mat(vec):
    """``Project_vector_to_matrix`` (``Models/_utils.py:155-188``)."""
    n = vec.shape[-1]
    dim = 2 if n == 3 else 3
    M = np.zeros(vec.shape[:-1] + (dim, dim))
    for d in range(dim):
        M[..., d, d] = vec[..., d]
    if dim == 2:
        M[..., 0, 1] = M[..., 1, 0] = vec[..., 2] / SQRT2
    else:
        M[..., 1, 2] = M[..., 2, 1] = vec[..., 3] / SQRT2
        M[..., 0, 2] = M[..., 2, 0] = vec[..., 4] / SQRT2
        M[..., 0, 1] = M[..., 1, 0] = vec[..., 5] / SQRT2
    return M


def mat_to_vec(M):
    """``Project_matrix_to_vector`` (``Models/_utils.py:191-222``)."""
    dim = M.shape[-1]
    if dim == 2:
        return np.stack([M[..., 0, 0], M[..., 1, 1], M[..., 0, 1] * SQRT2], -1)
    return np.stack([M[..., 0, 0], M[..., 1, 1], M[..., 2, 2], M[..., 1, 2] * SQRT2,
                     M[..., 0, 2] * SQRT2, M[..., 0, 1] * SQRT2], -1)


def eigen_2d(vec):
    """2D branch of ``_Eigen_values_vectors_projectors`` (:751-802)."""
    M = vec_to_mat(vec)
    det = M[..., 0, 0] * M[..., 1, 1] - M[..., 1, 0] * M[..., 0, 1]
    tr = M[..., 0, 0] + M[..., 1, 1]
    delta = tr**2 - 4 * det
    with np.errstate(invalid="ignore"):
        s = np.sqrt(delta)
    eigs = np.stack([(tr - s) / 2, (tr + s) / 2], -1)
    v1mv2 = eigs[..., 0] - eigs[..., 1]
    neq = eigs[..., 0] != eigs[..., 1]
    M1 = np.zeros_like(M)
    M1[..., 0, 0] = 1
    if neq.any():
        d = np.where(v1mv2 == 0, 1.0, v1mv2)
        I = np.eye(2)
        m1 = (M - eigs[..., 1, None, None] * I) / d[..., None, None]
        M1[neq] = m1[neq]
    M2 = np.eye(2) - M1
    return eigs, [mat_to_vec(M1), mat_to_vec(M2)], [M1, M2]


def eigen_3d(vec):
    """3D branch of ``_Eigen_values_vectors_projectors`` (:804-936): He's
    closed form with the reference's element-level case masks (a case-2/3
    hit at any Gauss point applies that case to the element's points)."""
    M = vec_to_mat(vec)
    Ne, nPg = M.shape[:2]
    I = np.broadcast_to(np.eye(3), M.shape)
    I1 = np.trace(M, axis1=-2, axis2=-1)
    I2 = 0.5 * (I1**2 - np.trace(M @ M, axis1=-2, axis2=-1))
    I3 = _det(M)
    g = I1**2 - 3 * I2
    sg = np.sqrt(g)
    gn0 = g != 0
    arg = 0.5 * (2 * I1**3 - 9 * I1 * I2 + 27 * I3)
    np.divide(arg, g ** (3 / 2), out=arg, where=gn0)
    theta = np.arccos(arg) / 3
    v1, v2, v3 = I1 / 3, I1 / 3, I1 / 3
    v1, v2, v3 = v1.copy(), v2.copy(), v3.copy()
    M1 = np.zeros_like(M)
    M1[..., 0, 0] = 1
    M3 = np.zeros_like(M)
    M3[..., 2, 2] = 1
    Irg = (1 / 3) * ((I1 - sg)[..., None, None] * I)
    case2 = np.unique(np.nonzero(gn0 & (theta == np.pi / 3))[0])
    if case2.size:
        v1[case2] += -2 / 3 * sg[case2]
        v2[case2] += 1 / 3 * sg[case2]
        v3[case2] += 1 / 3 * sg[case2]
        M1[case2] = ((g ** (-1 / 2))[..., None, None] * (Irg - M))[case2]
        M3[case2] = 0.5 * (I - M1)[case2]
    case3 = np.unique(np.nonzero(gn0 & (theta == 0))[0])
    if case3.size:
        v1[case3] += -1 / 3 * sg[case3]
        v2[case3] += -1 / 3 * sg[case3]
        v3[case3] += 2 / 3 * sg[case3]
        M3[case3] = ((g ** (-1 / 2))[..., None, None] * (M - Irg))[case3]
        M1[case3] = 0.5 * (I - M3)[case3]
    case1 = np.setdiff1d(np.unique(np.nonzero(gn0 & (theta != 0) & (theta != np.pi / 3))[0]),
                         np.union1d(case2, case3))
    if case1.size:
        v1[case1] += 2 / 3 * (sg * np.cos(2 * np.pi / 3 + theta))[case1]
        v2[case1] += 2 / 3 * (sg * np.cos(2 * np.pi / 3 - theta))[case1]
        v3[case1] += 2 / 3 * (sg * np.cos(theta))[case1]
        a, b, c, m = v1[case1][..., None, None], v2[case1][..., None, None], v3[case1][..., None, None], M[case1]
        E3 = np.eye(3)
        M1[case1] = (m - b * E3) @ (m - c * E3) / ((a - b) * (a - c))
        M3[case1] = (m - a * E3) @ (m - b * E3) / ((c - a) * (c - b))
    eigs = np.stack([v1, v2, v3], -1)
    M1 = M1 / np.linalg.norm(M1, axis=(-2, -1))[..., None, None]
    M3 = M3 / np.linalg.norm(M3, axis=(-2, -1))[..., None, None]
    M2 = I - (M1 + M3)
    return eigs, [mat_to_vec(M1), mat_to_vec(M2), mat_to_vec(M3)], [M1, M2, M3]


def spectral_projectors(vec, dim):
    """``__Spectral_Decomposition`` (:1043-1243): ``projP``, ``projM``."""
    if dim == 2:
        val, (m1, m2), _ = eigen_2d(vec)
        valp = (val + np.abs(val)) / 2
        dvalp = np.heaviside(val, 0.5)
        v = val[..., 0] - val[..., 1]
        v = np.where(v == 0, 1.0, v)
        BetaP = (valp[..., 0] - valp[..., 1]) / v
        gammap = dvalp - BetaP[..., None]
        projP = (BetaP[..., None, None] * np.eye(3)
                 + gammap[..., 0, None, None] * np.einsum("...i,...j->...ij", m1, m1)
                 + gammap[..., 1, None, None] * np.einsum("...i,...j->...ij", m2, m2))
        return projP, np.eye(3) - projP
    val, ms, Ms = eigen_3d(vec)
    valp = (val + np.abs(val)) / 2
    dvalp = np.heaviside(val, 0.5)
    d12 = np.where(val[..., 0] - val[..., 1] == 0, 1.0, val[..., 0] - val[..., 1])
    d13 = np.where(val[..., 0] - val[..., 2] == 0, 1.0, val[..., 0] - val[..., 2])
    d23 = np.where(val[..., 1] - val[..., 2] == 0, 1.0, val[..., 1] - val[..., 2])
    th = np.stack([(valp[..., 0] - valp[..., 1]) / (2 * d12), (valp[..., 0] - valp[..., 2]) / (2 * d13),
                   (valp[..., 1] - valp[..., 2]) / (2 * d23)], -1)
    P = sum(dvalp[..., a, None, None] * np.einsum("...i,...j->...ij", ms[a], ms[a]) for a in range(3))
    rI = np.array([0, 1, 2, 1, 0, 0])
    rJ = np.array([0, 1, 2, 2, 2, 1])
    scale = np.ones((6, 6))
    scale[3:, :3] = scale[:3, 3:] = SQRT2
    scale[3:, 3:] = 2.0
    for t, (A, Bm) in enumerate([(Ms[0], Ms[1]), (Ms[0], Ms[2]), (Ms[1], Ms[2])]):
        Aik = A[..., rI[:, None], rI[None, :]]
        Ail = A[..., rI[:, None], rJ[None, :]]
        Ajl = A[..., rJ[:, None], rJ[None, :]]
        Bik = Bm[..., rI[:, None], rI[None, :]]
        Bil = Bm[..., rI[:, None], rJ[None, :]]
        Bjl = Bm[..., rJ[:, None], rJ[None, :]]
        G = (Aik * Bjl + Ail * np.swapaxes(Bil, -2, -1) + Bik * Ajl + Bil * np.swapaxes(Ail, -2, -1)) * scale
        P = P + th[..., t, None, None] * G
    return P, np.eye(6) - P


def split_miehe(eps, E, v, dim, planeStress):
    """Miehe branch of ``__Split_Strain`` (:503-537)."""
    projP, projM = spectral_projectors(eps, dim)
    lmbda, mu, _ = lame(E, v, dim, planeStress)
    Rp, Rm = Rp_Rm(eps, dim)
    II = IxI(dim)
    lm = np.asarray(lmbda)[..., None, None] if np.ndim(lmbda) else lmbda
    mu = np.asarray(mu)[..., None, None] if np.ndim(mu) else mu
    cP = lm * (Rp[..., None, None] * II) + 2 * mu * projP
    cM = lm * (Rm[..., None, None] * II) + 2 * mu * projM
    return cP, cM


def _mT(A):
    return np.swapaxes(A, -2, -1)


def _C_iso(E, v, dim, planeStress):
    """Isotropic Kelvin-Mandel stiffness ``lam IxI + 2 mu I`` broadcast over
    the leading axes of ``E`` (``Isotropic._Behavior``, ``Models/Elastic/_laws.py:407-482``)."""
    lmbda, mu, _ = lame(E, v, dim, planeStress)
    n = 3 if dim == 2 else 6
    lm = np.asarray(lmbda)[..., None, None]
    mu = np.asarray(mu)[..., None, None]
    return lm * IxI(dim) + 2 * mu * np.eye(n)


def _matvec(A, x):
    """``A x`` summed in index order with unfused products.  The reference
    forms ``C eps`` with numpy's matmul, whose BLAS may fuse multiply-adds:
    equal principal stresses (a hydrostatic strain) then come out an ulp
    apart, the 2D discriminant ``tr^2 - 4 det`` rounds negative and its
    unguarded square root (``Models/_phasefield.py:783-787``) returns NaN.
    Summed this way the equal components are equal and the discriminant is 0,
    the value the decomposition is written for."""
    A = np.broadcast_to(A, x.shape + (x.shape[-1],))
    out = A[..., :, 0] * x[..., 0, None]
    for j in range(1, x.shape[-1]):
        out = out + A[..., :, j] * x[..., j, None]
    return out


def _cross_variants(split, Cpp, Cpm, Cmm, Cmp):
    """The four ways of assigning the cross terms (:548-566, :660-676)."""
    if split.endswith("_PM"):
        return Cpp + Cpm, Cmm + Cmp
    if split.endswith("_MP"):
        return Cpp + Cmp, Cmm + Cpm
    if split.endswith("_NoCross"):
        return Cpp, Cmm + Cpm + Cmp
    return Cpp + Cpm + Cmp, Cmm


def split_bourdin(eps, E, v, dim, planeStress):
    """``__Split_Bourdin`` (:433-449): ``c+ = C``, ``c- = 0``."""
    C = np.broadcast_to(_C_iso(E, v, dim, planeStress), eps.shape + (eps.shape[-1],))
    return C.copy(), np.zeros_like(C)


def split_anisot_strain(split, eps, E, v, dim, planeStress):
    """"Strain" branch of ``__Split_Strain`` (:539-566)."""
    projP, projM = spectral_projectors(eps, dim)
    C = _C_iso(E, v, dim, planeStress)
    projPTC = _mT(projP) @ C
    projMTc = _mT(projM) @ C
    return _cross_variants(split, projPTC @ projP, projPTC @ projM, projMTc @ projM, projMTc @ projP)


def split_stress(split, eps, E, v, dim, planeStress):
    """``__Split_Stress`` (:573-678): the Stress split, Zhang and the
    AnisotStress variants, from the projectors of ``sigma = C eps``."""
    C = _C_iso(E, v, dim, planeStress)
    Sigma = _matvec(C, eps)
    projP, projM = spectral_projectors(Sigma, dim)
    if split == "Stress":
        E_ = np.asarray(E, float)[..., None, None] if np.ndim(E) else E
        mu = E_ / (2 * (1 + v))
        Rp, Rm = Rp_Rm(Sigma, dim)
        Rp, Rm = Rp[..., None, None], Rm[..., None, None]
        II = IxI(dim)
        if dim == 2:
            cv = v / E_ if planeStress else v * (1 + v) / E_
            sP = ((1 + v) / E_ * projP) - (cv * Rp * II)
            sM = ((1 + v) / E_ * projM) - (cv * Rm * II)
        else:
            sP = (1 / (2 * mu) * projP) - (v / E_ * Rp * II)
            sM = (1 / (2 * mu) * projM) - (v / E_ * Rm * II)
        return _mT(C) @ sP @ C, _mT(C) @ sM @ C
    Cp, Cm = projP @ C, projM @ C
    if split == "Zhang":
        return Cp, Cm
    S = np.linalg.inv(C)
    ps, ms = _mT(Cp) @ S, _mT(Cm) @ S
    return _cross_variants(split, ps @ Cp, ps @ Cm, ms @ Cm, ms @ Cp)


def split_he(eps, E, v, dim, planeStress):
    """``__Split_He`` (:680-749) with ``Get_sqrt_C_S``
    (``Models/Elastic/_laws.py:223-247``: eigh of C)."""
    C = _C_iso(E, v, dim, planeStress)
    lam, Q = np.linalg.eigh(C)
    sqrt_lam = np.sqrt(lam)[..., None, :]
    sqrtC = (Q * sqrt_lam) @ _mT(Q)
    inv_sqrtC = (Q / sqrt_lam) @ _mT(Q)
    epst = _matvec(sqrtC, eps)
    projPt, projMt = spectral_projectors(epst, dim)
    projP = inv_sqrtC @ projPt @ sqrtC
    projM = inv_sqrtC @ projMt @ sqrtC
    return C @ projP, C @ projM


SPLITS = ("Bourdin", "Amor", "Miehe", "He", "Stress", "Zhang", "AnisotStrain", "AnisotStrain_PM",
          "AnisotStrain_MP", "AnisotStrain_NoCross", "AnisotStress", "AnisotStress_PM", "AnisotStress_MP",
          "AnisotStress_NoCross")


def calc_C(split, eps, E, v, dim, planeStress):
    """``PhaseField.Calc_C`` (:396-431) for an isotropic material."""
    if split == "Bourdin":
        return split_bourdin(eps, E, v, dim, planeStress)
    if split == "Amor":
        return split_amor(eps, E, v, dim, planeStress)
    if split == "Miehe":
        return split_miehe(eps, E, v, dim, planeStress)
    if "Strain" in split:
        return split_anisot_strain(split, eps, E, v, dim, planeStress)
    if split == "Zhang" or "Stress" in split:
        return split_stress(split, eps, E, v, dim, planeStress)
    if split == "He":
        return split_he(eps, E, v, dim, planeStress)
    raise NotImplementedError(split)


def calc_psi(split, eps, E, v, dim, planeStress):
    """``Calc_psi_e_pg`` (:335-358) via ``Calc_Sigma_e_pg`` (:360-394)."""
    cP, cM = calc_C(split, eps, E, v, dim, planeStress)
    sP = (cP @ eps[..., None])[..., 0]
    sM = (cM @ eps[..., None])[..., 0]
    return np.sum(0.5 * eps * sP, -1), np.sum(0.5 * eps * sM, -1)


def pf_k(Gc, l0, regu):
    """``PhaseField.k`` (:236-251)."""
    return 3 / 4 * Gc * l0 if regu == "AT1" else Gc * l0


def pf_r(psiP, Gc, l0, regu):
    """``Get_r_e_pg`` (:253-271); ``Gc`` scalar or ``(Ne,)``."""
    Gc = np.asarray(Gc, float)
    if Gc.ndim:
        Gc = Gc[:, None]
    return 2 * psiP if regu == "AT1" else 2 * psiP + Gc / l0


def pf_f(psiP, Gc, l0, regu):
    """``Get_f_e_pg`` (:273-293)."""
    Gc = np.asarray(Gc, float)
    if Gc.ndim:
        Gc = Gc[:, None]
    if regu == "AT1":
        f = 2 * psiP - (3 * Gc) / (8 * l0)
        return (f + np.abs(f)) / 2
    return 2 * psiP


def pf_g(d, connect, N_pg, k_res=1e-12):
    """``Get_g_e_pg`` (:295-317)."""
    d_pg = np.einsum("pn,en->ep", N_pg, d[connect])
    return (1 - d_pg) ** 2 + k_res


# ---------------------------------------------------------------------------
# Post-processing of linear-elastic results (section 8(f), rank 1)
# ---------------------------------------------------------------------------


def strain_e_pg(geo: Geometry, u: np.ndarray, connect: np.ndarray) -> np.ndarray:
    """``Calc_Epsilon_e_pg`` (``Models/Elastic/_laws.py:127-157``):
    ``B_e_pg @ u_e`` (Kelvin-Mandel) at the rule's points, ``(Ne, nPg, ns)``."""
    ue = u[assembly_e(connect, geo.dim)]
    return np.einsum("epij,ej->epi", geo.B(), ue)


def stress_e_pg(eps: np.ndarray, C: np.ndarray) -> np.ndarray:
    """``Calc_Sigma_e_pg`` (``_laws.py:159-185``): ``C @ eps`` with ``C``
    ``(ns, ns)`` or ``(Ne, ns, ns)``."""
    C = np.asarray(C, dtype=float)
    if C.ndim == 2:
        return np.einsum("ij,epj->epi", C, eps)
    return np.einsum("eij,epj->epi", C, eps)


def result_in_field(field_e_pg: np.ndarray, result: str, coef=SQRT2) -> np.ndarray:
    """``__Result_in_Strain_or_Stress_field`` (``Models/_utils.py:302-392``):
    the shear components divided by ``coef``, then the component, the von
    Mises measure or the whole field at every point."""
    f = np.array(field_e_pg, dtype=float, copy=True)
    ns = f.shape[-1]
    if ns == 3:
        f[:, :, 2] *= 1 / coef
        xx, yy, xy = (f[:, :, i] for i in range(3))
        comps = {"xx": xx, "yy": yy, "xy": xy}
        vm = np.sqrt(xx**2 + yy**2 - xx * yy + 3 * xy**2)
    else:
        f[:, :, 3:] *= 1 / coef
        xx, yy, zz, yz, xz, xy = (f[:, :, i] for i in range(6))
        comps = {"xx": xx, "yy": yy, "zz": zz, "yz": yz, "xz": xz, "xy": xy}
        vm = np.sqrt(0.5 * ((xx - yy) ** 2 + (yy - zz) ** 2 + (zz - xx) ** 2 + 6 * (xy**2 + yz**2 + xz**2)))
    for k, v in comps.items():
        if k in result:
            return v
    if "vm" in result:
        return vm
    if result in ("Strain", "Stress"):
        return f
    raise ValueError(result)


def strain_or_stress_result_e(field_e_pg: np.ndarray, result: str, coef=SQRT2) -> np.ndarray:
    """``Result_strain_or_stress_field_e`` (``Models/_utils.py:395-434``) for
    one group: the point values averaged over the points, ``(Ne,)`` or
    ``(Ne, ns)``."""
    return np.asarray(result_in_field(field_e_pg, result, coef)).mean(1)


def psi_elas_e(geo: Geometry, eps: np.ndarray, sig: np.ndarray, thickness: float = 1.0) -> np.ndarray:
    """``Elastic._Calc_Psi_Elas(returnScalar=False)`` (``Simulations/_elastic.py:323-380``):
    ``1/2 sum_p wJ sig . eps`` per element, times the thickness in 2D."""
    w = 0.5 * np.einsum("ep,epi,epi->e", geo.wJ, sig, eps)
    return w * thickness if geo.dim == 2 else w


def pf_stress_e_pg(geo: Geometry, u: np.ndarray, d: np.ndarray, connect: np.ndarray, split: str, E, v: float,
                   planeStress: bool) -> np.ndarray:
    """``PhaseField._Calc_Sigma_e_pg`` (``Simulations/_phasefield.py:868-916``):
    ``g(d) * (c+ eps) + c- eps`` at the rule's points, with the split of
    ``calc_C`` and ``g = (1 - N.d)^2 + 1e-12``."""
    eps = strain_e_pg(geo, u, connect)
    E = np.asarray(E, float)
    cP, cM = calc_C(split, eps, E[:, None] if E.ndim else E, v, geo.dim, planeStress)
    sP = np.einsum("epij,epj->epi", cP, eps)
    sM = np.einsum("epij,epj->epi", cM, eps)
    g = pf_g(d, connect, geo.N_pg)
    return g[..., None] * sP + sM


def zz1(geo: Geometry, u: np.ndarray, connect: np.ndarray, C: np.ndarray, Nn: int, thickness: float = 1.0):
    """``Elastic._Calc_ZZ1`` (``Simulations/_elastic.py:398-418``) with the
    smoothed energy of ``_Calc_Psi_Elas(smoothedStress=True)`` (:357-378):
    the element-mean stresses (Kelvin-Mandel) averaged at the nodes,
    interpolated back to the points with ``N`` and dotted with the strain.
    Returns ``(error, error_e)``."""
    eps = strain_e_pg(geo, u, connect)
    sig = stress_e_pg(eps, C)
    W_e = psi_elas_e(geo, eps, sig, thickness)
    sig_n = node_values(connect, Nn, sig.mean(axis=1))
    sig_s = np.einsum("eni,pn->epi", sig_n[connect], geo.N_pg)
    Ws_e = psi_elas_e(geo, eps, sig_s, thickness)
    W, Ws = W_e.sum(), Ws_e.sum()
    return abs(W - Ws) / W, np.abs(Ws_e - W_e) / W


def node_values(connect: np.ndarray, Nn: int, values_e: np.ndarray) -> np.ndarray:
    """``Mesh.Get_Node_Values`` (``FEM/_mesh.py:822-873``) with the incidence of
    ``_GroupElem.Get_connect_n_e`` (``FEM/_group_elem.py:322-344``): per node,
    the sum of the surrounding elements' values divided by their count; nodes
    of no element stay 0."""
    Ne, nPe = connect.shape
    v = np.asarray(values_e, dtype=float)
    one_d = v.ndim == 1
    v = v.reshape(Ne, -1)
    G = sp.csr_matrix((np.ones(nPe * Ne), (connect.ravel(), np.repeat(np.arange(Ne), nPe))), shape=(Nn, Ne))
    nodes = np.unique(connect)
    Gn = G[nodes]
    count = np.asarray(Gn.sum(axis=1)).reshape(-1, 1)
    out = np.zeros((Nn, v.shape[1]))
    for c in range(v.shape[1]):
        out[nodes, c] = ((Gn @ v[:, c].reshape(Ne, 1)) * 1 / count).ravel()
    return out.ravel() if one_d else out


# ---------------------------------------------------------------------------
# Simulations: the driver-level semantics
# ---------------------------------------------------------------------------


class _SimuBase:
    """Boundary-condition bookkeeping shared by the oracle simulations
    (``_simu.py:2190-2469, 2572-2630``)."""

    def __init__(self, connect, coord, elemType, dim, boundary):
        self.connect = np.asarray(connect, dtype=np.int64)
        self.coord = np.asarray(coord, dtype=float)
        self.elemType = elemType
        self.dim = dim
        self.Nn = self.coord.shape[0]
        self.boundary = boundary  # {elemType: connect} of lower-dimensional groups
        used = np.zeros(self.Nn, dtype=bool)
        used[self.connect.ravel()] = True
        self.orphan = np.nonzero(~used)[0]
        self.Bc_Init()

    def Bc_Init(self):
        self.dir = {}
        self.neu = {}

    def _add(self, store, pt, dofs, vals):
        d, v = store.get(pt, (np.array([], int), np.array([])))
        store[pt] = (np.concatenate([d, dofs]), np.concatenate([v, vals]))

    def add_dirichlet(self, nodes, values, unknowns, problemType="elastic"):
        nodes = np.asarray(nodes)
        avail = self.unknowns(problemType)
        vals = np.zeros((nodes.size, len(unknowns)))
        c = self.coord[nodes]
        for d, val in enumerate(values):
            vals[:, d] = val(c[:, 0], c[:, 1], c[:, 2]) if callable(val) else val
        self._add(self.dir, problemType, dofs_nodes(nodes, unknowns, avail), vals.ravel())

    def add_surfLoad(self, nodes, values, unknowns, problemType="elastic", thickness=1.0):
        avail = self.unknowns(problemType)
        inDim = 2 if self.dim == 2 else 3
        bdim = self.dim - 1
        for et, cb in self.boundary.items():
            if ELEM_DIM[et] != bdim:
                continue
            vals, dofs = integrate_load(et, cb, self.coord, inDim, nodes, values, unknowns, avail)
            if self.dim == 2:
                vals = vals * thickness
            self._add(self.neu, problemType, dofs, vals)

    def add_volumeLoad(self, nodes, values, unknowns, problemType="elastic", thickness=1.0):
        """``add_volumeLoad`` (``_simu.py:2683-2741``): integral over the
        main elements, times the thickness in 2D."""
        avail = self.unknowns(problemType)
        vals, dofs = integrate_load(self.elemType, self.connect, self.coord, self.dim, nodes, values, unknowns,
                                    avail)
        if self.dim == 2:
            vals = vals * thickness
        self._add(self.neu, problemType, dofs, vals)

    def add_pressureLoad(self, nodes, magnitude, problemType="elastic", thickness=1.0):
        """``add_pressureLoad`` (``_simu.py:2632-2681, 2910-2944``): nodal
        normals times ``magnitude`` (times the thickness in 2D), integrated
        as nodal values over the boundary elements."""
        avail = self.unknowns(problemType)
        inDim = 2 if self.dim == 2 else 3
        if self.dim == 2:
            magnitude = magnitude * thickness
        for et, cb in self.boundary.items():
            if ELEM_DIM[et] != self.dim - 1:
                continue
            normals, used = node_normals(et, cb, self.coord, np.asarray(nodes))
            if used.size == 0:
                continue
            values = [normals[:, d] * magnitude for d in range(inDim)]
            vals, dofs = integrate_load(et, cb, self.coord, inDim, used, values, avail[:inDim], avail)
            self._add(self.neu, problemType, dofs, vals)

    def unknowns(self, problemType):
        return ["x", "y"] if self.dim == 2 else ["x", "y", "z"]

    def neumann_vector(self, problemType, Ndof):
        dofs, vals = self.neu.get(problemType, (np.array([], int), np.array([])))
        b = sp.csr_matrix((vals, (dofs, np.zeros(len(dofs)))), shape=(Ndof, 1))
        return b.toarray().ravel()


class ElasticOracle(_SimuBase):
    """``Simulations.Elastic`` static solve (``Simulations/_elastic.py:123-152``
    and the static pipeline: assembly, Dirichlet reduction, direct solve)."""

    def __init__(self, connect, coord, elemType, E, v, planeStress=True, thickness=1.0, boundary=None, C=None):
        dim = ELEM_DIM[elemType]
        super().__init__(connect, coord, elemType, dim, boundary or {})
        self.E, self.v, self.planeStress, self.thickness = E, v, planeStress, thickness
        self.C = C  # a Kelvin-Mandel stiffness replacing the isotropic one (anisotropic laws)
        self.Ndof = self.Nn * dim
        self._cmap = None
        self.u = np.zeros(self.Ndof)

    def element_K(self):
        geo = Geometry(self.elemType, self.connect, self.coord, "rigi", self.dim)
        C = C_isotropic(self.E, self.v, self.dim, self.planeStress) if self.C is None else self.C
        Ke = K_elastic(geo, C)
        if self.dim == 2:
            Ke = Ke * self.thickness
        return Ke

    def Get_K(self):
        if self._cmap is None:
            self._cmap = csr_map(self.connect, self.dim, self.Ndof)
        return assemble_csr(self.element_K(), self._cmap, self.Ndof)

    def add_surfLoad(self, nodes, values, unknowns, problemType="elastic"):
        super().add_surfLoad(nodes, values, unknowns, problemType, self.thickness)

    def Solve(self):
        K = self.Get_K()
        b = self.neumann_vector("elastic", self.Ndof)
        dofs, vals = self.dir.get("elastic", (np.array([], int), np.array([])))
        orphan = dofs_nodes(self.orphan, self.unknowns("elastic"), self.unknowns("elastic"))
        self.u = solve_r1(K, b, dofs, vals, self.Ndof, orphan)
        return self.u


class PhaseFieldOracle(_SimuBase):
    """``Simulations.PhaseField`` staggered solve with the ``History``,
    ``HistoryDamage`` or ``BoundConstrain`` irreversibility
    (``Simulations/_phasefield.py:248-657``), every split, AT1/AT2.
    Reproduces the assembly caching order of the reference (``Get_K_C_M_F``
    :248-269): ``Ku`` is built from the pre-solve displacement and the new
    damage, ``Kd`` from the displacement current when it is next needed (the
    energies of options 1 and 2 assemble it from the post-solve one), the
    history ``H`` advances only in ``Save_Iter`` (to the last assembled psi+),
    and ``Calc_Reaction`` uses the cached ``Ku``."""

    def __init__(self, connect, coord, elemType, E, v, planeStress, thickness, split, regu, Gc, l0,
                 boundary=None, solver="History", lsq="exact"):
        dim = ELEM_DIM[elemType]
        super().__init__(connect, coord, elemType, dim, boundary or {})
        assert solver in ("History", "HistoryDamage", "BoundConstrain") and lsq in ("exact", "reference")
        self.solver, self.lsq = solver, lsq
        self.E, self.v, self.planeStress, self.thickness = E, v, planeStress, thickness
        self.split, self.regu, self.Gc, self.l0 = split, regu, Gc, l0
        self.u = np.zeros(self.Nn * dim)
        self.d = np.zeros(self.Nn)
        self.geo_r = Geometry(elemType, self.connect, self.coord, "rigi", dim)
        self.geo_m = Geometry(elemType, self.connect, self.coord, "mass", dim)
        self.Bm = self.geo_m.B()
        self.Br = self.geo_r.B()
        self.cmap_u = csr_map(self.connect, dim, self.Nn * dim)
        self.cmap_d = csr_map(self.connect, 1, self.Nn)
        self.cmap_f = csr_map(self.connect, 1, self.Nn, isMatrix=False)
        self.old_psiP = np.zeros((self.connect.shape[0], self.geo_m.nPg))
        self.psiP = None
        self.updU = False
        self.updD = False
        self.Ku = self.Kd = self.Fd = None
        self.Niter = 0

    def unknowns(self, problemType):
        return ["d"] if problemType == "damage" else super().unknowns(problemType)

    def _E_pg(self):
        E = np.asarray(self.E, float)
        return E[:, None] if E.ndim else E

    def _eps(self, B):
        ue = self.u[assembly_e(self.connect, self.dim)]
        return np.einsum("epij,ej->epi", B, ue)

    def calc_psiP(self):
        """``__Calc_psiPlus_e_pg`` (:494-538)."""
        eps = self._eps(self.Bm)
        psiP, _ = calc_psi(self.split, eps, self._E_pg(), self.v, self.dim, self.planeStress)
        if self.solver == "History":  # only History clamps psi+ (_phasefield.py:512-529)
            inc = psiP - self.old_psiP
            psiP = np.where(inc < 0, self.old_psiP, psiP)
        self.psiP = psiP
        return psiP

    def assemble_damage(self):
        """``__Construct_Damage_Matrix`` (:540-573)."""
        psiP = self.calc_psiP()
        r = pf_r(psiP, self.Gc, self.l0, self.regu)
        Ke = UV(self.geo_m, r) + GradU_A_GradV(self.geo_r, np.eye(self.dim), pf_k(self.Gc, self.l0, self.regu))
        Fe = V(self.geo_m, pf_f(psiP, self.Gc, self.l0, self.regu))
        if self.dim == 2:
            Ke = Ke * self.thickness
            Fe = Fe * self.thickness
        self.Kd = assemble_csr(Ke, self.cmap_d, self.Nn)
        self.Fd = assemble_csr(Fe, self.cmap_f, self.Nn, isMatrix=False).toarray().ravel()
        self.updD = True

    def element_Ku(self):
        """``__Construct_Elastic_Matrix`` (:444-482)."""
        eps = self._eps(self.Br)
        cP, cM = calc_C(self.split, eps, self._E_pg(), self.v, self.dim, self.planeStress)
        g = pf_g(self.d, self.connect, self.geo_r.N_pg)
        c = g[..., None, None] * cP + cM
        Ke = K_elastic(self.geo_r, c)
        if self.dim == 2:
            Ke = Ke * self.thickness
        return Ke

    def assemble_elastic(self):
        self.Ku = assemble_csr(self.element_Ku(), self.cmap_u, self.Nn * self.dim)
        self.updU = True

    def Get_K(self, problemType):
        """``Get_K_C_M_F`` (:248-269) caching."""
        if problemType == "elastic":
            if not self.updU:
                self.assemble_elastic()
            return self.Ku
        if not self.updD:
            self.assemble_damage()
        return self.Kd

    def psi_crack(self):
        """``_Calc_Psi_Crack`` (:798-816): ``1/2 d.Kd d`` with the cached
        (or freshly assembled) ``Kd``; 0 for a zero damage field."""
        Kd = self.Get_K("damage")
        if np.linalg.norm(self.d) == 0:
            return 0.0
        return 0.5 * self.d @ (Kd @ self.d)

    def psi_elas(self):
        """``_Calc_Psi_Elas`` (:775-796): ``1/2 u.Ku u`` with the cached (or
        freshly assembled) ``Ku``; 0 for a zero displacement."""
        Ku = self.Get_K("elastic")
        if np.linalg.norm(self.u) == 0:
            return 0.0
        return 0.5 * self.u @ (Ku @ self.u)

    def psi_ext(self, fu):
        """``_Calc_Psi_Ext`` (:818-836): ``u.f`` (no 1/2); 0 when either is zero."""
        if np.linalg.norm(self.u) == 0 or np.linalg.norm(fu) == 0:
            return 0.0
        return self.u @ fu

    def _energy(self, convOption, fu):
        """The convergence energy of ``Solve`` (:354-361, 376-381): crack
        energy (option 1) or total energy crack + elastic - external (option 2,
        Ambati 2015 eq. 39), each term in the reference's order of calls."""
        E = self.psi_crack()
        if convOption == 2:
            E += self.psi_elas() - self.psi_ext(fu)
        return E

    def _solve(self, problemType):
        if problemType == "damage":
            K = self.Get_K("damage")
            b = self.Fd.copy()
            dofs, vals = self.dir.get("damage", (np.array([], int), np.array([])))
            if self.solver == "BoundConstrain":
                assert len(dofs) == 0, "BoundConstrain: the reference's bounds cover every damage DOF"
                lb = self.d.copy()  # Get_lb_ub (_phasefield.py:148-163)
                lb[lb >= 1] = 1 - np.finfo(float).eps
                solve = bounded_lsq_exact if self.lsq == "exact" else bounded_lsq
                self.d = solve(K, b, lb, np.ones_like(lb))
            else:
                self.d = solve_r1(K, b, dofs, vals, self.Nn)
            return self.d
        K = self.Get_K("elastic")
        b = self.neumann_vector("elastic", self.Nn * self.dim)
        dofs, vals = self.dir.get("elastic", (np.array([], int), np.array([])))
        self.u = solve_r1(K, b, dofs, vals, self.Nn * self.dim)
        return self.u

    def Solve(self, tolConv=1.0, maxIter=500, convOption=2):
        """``PhaseField.Solve`` (:303-432), every convergence option:
        0 max |d_np1 - d_n|, 1 relative change of the crack energy, 2 of the
        total energy, 3 the summed relative changes of u and d (converged when
        convD <= tolConv and convU <= 0.999 tolConv).  A pass solves the
        damage with the cached ``Kd``, invalidates ``Ku``, solves the
        displacement, invalidates ``Kd``; ``tolConv == 1`` or an all-zero
        damage stops after one pass.  ``self.trace`` records, per pass,
        ``(u_n, d_n, u_np1, d_np1, convIter)``; ``HistoryDamage`` returns
        the node-wise maximum of the damage before and after the call."""
        assert 0 < tolConv <= 1, "tolConv must be between 0 and 1."
        assert maxIter > 1, "Must be > 1."
        assert convOption in (0, 1, 2, 3)
        Niter = 0
        converged = False
        old_damage = self.d.copy()
        fu = self.neumann_vector("elastic", self.Nn * self.dim) if convOption == 2 else None
        self.trace = []
        while not converged and Niter < maxIter:
            Niter += 1
            d_n = self.d.copy()
            u_n = self.u.copy()
            if convOption in (1, 2):
                E_n = self._energy(convOption, fu)
            d_np1 = self._solve("damage").copy()
            self.updU = False
            u_np1 = self._solve("elastic").copy()
            self.updD = False
            if convOption == 0:
                conv = np.max(np.abs(d_np1 - d_n))
            elif convOption in (1, 2):
                E_np1 = self._energy(convOption, fu)
                conv = np.abs(E_n - E_np1) if E_np1 == 0 else np.abs((E_n - E_np1) / E_np1)
            else:
                diffU = np.abs(u_np1 - u_n)
                nz = u_np1 != 0
                diffU[nz] *= 1 / np.abs(u_np1[nz])
                diffD = np.abs(d_np1 - d_n)
                nz = d_np1 != 0
                diffD[nz] *= 1 / np.abs(d_np1[nz])
                convU, convD = np.sum(diffU), np.sum(diffD)
                conv = max(convD, convU)
            self.trace.append((u_n, d_n, u_np1, d_np1, float(conv)))
            if tolConv == 1 or d_np1.max() == 0:
                converged = True
            elif convOption == 3:
                converged = (convD <= tolConv) and (convU <= tolConv * 0.999)
            else:
                converged = conv <= tolConv
        self.Niter = Niter
        self.convIter = float(conv)
        if self.solver == "HistoryDamage":
            d_np1 = np.maximum(old_damage, d_np1)
        return u_np1, d_np1, converged

    def Save_Iter(self):
        """History update of ``Save_Iter`` (:613-630)."""
        self.old_psiP = self.psiP

    def Calc_Reaction(self, dofs):
        """``_Simu.Calc_Reaction`` (``_simu.py:212-260``), elliptic."""
        K = self.Get_K("elastic")
        return K[dofs] @ self.u


# ---------------------------------------------------------------------------
# a24-a28: time schemes, dynamic elastic and thermal simulations
# ---------------------------------------------------------------------------

HYPERBOLIC = ("newmark", "midpoint", "hht", "hht_newmark", "euler_implicit", "euler_explicit")


def hyperbolic_params(dt, algo="newmark", beta=0.25, gamma=0.5, alpha=0.5):
    """``Solver_Set_Hyperbolic_Algorithm`` (``Simulations/_simu.py:1214-1281``):
    ``(dt, beta, gamma, alpha)``, hht_newmark deriving beta and gamma."""
    assert algo in HYPERBOLIC and dt > 0
    if algo == "hht_newmark":
        assert 0 <= alpha <= 1 / 3
        beta = 1 / 4 * (1 + alpha) ** 2
        gamma = 1 / 2 + alpha
    else:
        assert 0 <= alpha < 1
    return (dt, beta, gamma, alpha)


def time_coefs(algo, params):
    """``_Solver_Get_K_C_M_coefs_for_time_scheme`` (``_simu.py:1384-1436``)."""
    if algo == "parabolic":
        dt, alpha = params
        return 1, 1 / (alpha * dt), 0
    dt, beta, gamma, alpha = params
    return {
        "newmark": (1, gamma / (beta * dt), 1 / (beta * dt ** 2)),
        "hht": (1 - alpha, (1 - alpha) * gamma / (beta * dt), (1 - alpha) / (beta * dt ** 2)),
        "midpoint": (0.5, 1 / dt, 2 / dt ** 2),
        "hht_newmark": (1 - alpha, gamma / (beta * dt), 1 / (beta * dt ** 2)),
        "euler_implicit": (1, 1 / dt, 1 / dt ** 2),
        "euler_explicit": (0, 0, 1),
    }[algo]


def time_rhs(algo, params, K, C, M, F, u_n, v_n, a_n):
    """``_Solver_Apply_Neumann`` (``_simu.py:1735-1860``): ``b`` of the step
    with the history terms applied as the reference writes them (matrix sums
    times predictors).  ``F``: the Neumann vector (dense)."""
    b = np.array(F, dtype=float)
    if algo == "parabolic":
        dt, alpha = params
        ut = u_n + (1 - alpha) * dt * v_n
        return b + 1 / (alpha * dt) * (C @ ut)
    dt, beta, gamma, alpha = params
    if algo in ("newmark", "hht_newmark"):
        ut = u_n + dt * v_n + dt ** 2 / 2 * (1 - 2 * beta) * a_n
        vt = v_n + dt * (1 - gamma) * a_n
        cC, cM = gamma / (beta * dt), 1 / (beta * dt ** 2)
        b = b + (cC * C + cM * M) @ ut
        b = b - C @ vt
        if algo == "hht_newmark":
            b = b - alpha * (K @ u_n)
        return b
    if algo == "midpoint":
        b = b + (2 / dt ** 2 * M + 1 / dt * C - 1 / 2 * K) @ u_n
        return b + (2 / dt * M) @ v_n
    if algo == "hht":
        cM, cC = 1 / (beta * dt ** 2), gamma / (beta * dt)
        b = b - ((alpha - 1) * (cM * M + cC * C) + alpha * K) @ u_n
        cM, cC = (alpha - 1) / (beta * dt), (alpha - 1) * (gamma / beta) + 1
        b = b - (cM * M + cC * C) @ v_n
        cM, cC = (alpha - 1) / (2 * beta) + 1, dt * (alpha - 1) * (gamma / (2 * beta) - 1)
        return b - (cM * M + cC * C) @ a_n
    if algo == "euler_implicit":
        b = b + (1 / dt ** 2 * M + 1 / dt * C) @ u_n
        return b + (1 / dt * M) @ v_n
    if algo == "euler_explicit":
        return b - K @ u_n - C @ v_n
    raise ValueError(algo)


def time_update(algo, params, u_n, v_n, a_n, u_np1):
    """``_Solver_Update_solutions`` (``_simu.py:1552-1640``): ``(u, v, a)``
    after the step (``None`` = unchanged), in the reference's operation order."""
    if algo == "parabolic":
        dt, alpha = params
        vt_np1 = u_n + ((1 - alpha) * dt * v_n)
        return u_np1, (u_np1 - vt_np1) / (alpha * dt), None
    dt, beta, gamma, alpha = params
    if algo in ("newmark", "hht_newmark"):
        ut_np1 = u_n + dt * v_n + dt ** 2 / 2 * (1 - 2 * beta) * a_n
        vt_np1 = v_n + dt * (1 - gamma) * a_n
        a_np1 = (u_np1 - ut_np1) / (beta * dt ** 2)
        return u_np1, vt_np1 + gamma * dt * a_np1, a_np1
    if algo == "midpoint":
        v_np1 = 2 / dt * (u_np1 - u_n) - v_n
        return u_np1, v_np1, 2 / dt * (v_np1 - v_n) - a_n
    if algo == "hht":
        a_np1 = 1 / (beta * dt) * ((u_np1 - u_n) / dt - v_n) + (1 - 1 / (2 * beta)) * a_n
        return u_np1, dt * ((1 - gamma) * a_n + gamma * a_np1) + v_n, a_np1
    if algo == "euler_implicit":
        v_np1 = (u_np1 - u_n) / dt
        return u_np1, v_np1, (v_np1 - v_n) / dt
    if algo == "euler_explicit":
        a_np1 = u_np1
        return u_n + dt * v_n, v_n + dt * a_np1, a_np1
    raise ValueError(algo)


def time_evaluate(algo, params, u_n, v_n, a_n, u_np1):
    """``_Solver_Evaluate_u_v_a_for_time_scheme`` (``_simu.py:1292-1382``):
    the states ``(u_t, v_t, a_t)`` at which the scheme enforces
    ``K u_t + C v_t + M a_t = F`` (``a_t`` None: no mass term)."""
    if algo == "parabolic":
        dt, alpha = params
        ut_np1 = u_n + (1 - alpha) * dt * v_n
        return u_np1, (u_np1 - ut_np1) / (alpha * dt), None
    dt, beta, gamma, alpha = params
    if algo == "newmark":
        ut_np1 = u_n + dt * v_n + dt ** 2 / 2 * (1 - 2 * beta) * a_n
        vt_np1 = v_n + dt * (1 - gamma) * a_n
        a_t = (u_np1 - ut_np1) / (beta * dt ** 2)
        return u_np1, vt_np1 + gamma * dt * a_t, a_t
    if algo == "hht":
        a_np1 = 1 / (beta * dt ** 2) * (u_np1 - u_n - dt * v_n) + (1 - 1 / (2 * beta)) * a_n
        v_np1 = dt * ((1 - gamma) * a_n + gamma * a_np1) + v_n
        return ((1 - alpha) * u_np1 + alpha * u_n, (1 - alpha) * v_np1 + alpha * v_n,
                (1 - alpha) * a_np1 + alpha * a_n)
    if algo == "midpoint":
        v_np1 = 2 / dt * (u_np1 - u_n) - v_n
        a_np1 = 2 / dt * (v_np1 - v_n) - a_n
        return (u_np1 + u_n) / 2, (v_np1 + v_n) / 2, (a_np1 + a_n) / 2
    if algo == "hht_newmark":
        a_np1 = 1 / (beta * dt ** 2) * (u_np1 - u_n - dt * v_n) + (1 - 1 / (2 * beta)) * a_n
        v_np1 = dt * ((1 - gamma) * a_n + gamma * a_np1) + v_n
        return (1 - alpha) * u_np1 + alpha * u_n, v_np1, a_np1
    if algo == "euler_implicit":
        v_np1 = (u_np1 - u_n) / dt
        return u_np1, v_np1, (v_np1 - v_n) / dt
    if algo == "euler_explicit":
        return u_n, v_n, None
    raise ValueError(algo)


def time_step(algo, params, K, C, M, F, u_n, v_n, a_n, dir_dofs=(), dir_vals=(), orphan_dofs=()):
    """One step of ``_Solver_Solve_problemType`` (``_simu.py:1529-1550``):
    ``A = coefK K + coefC C + coefM M`` (:1893-1899), ``b`` from
    ``time_rhs``, the r1 solve with the Dirichlet values (zeros for the
    explicit scheme's unknown ``a^n``, :1905-1907), then ``time_update``.
    Returns ``(u, v, a)`` with unchanged entries carried over."""
    Ndof = len(u_n)
    cK, cC, cM = time_coefs(algo, params)
    A = (cK * K + cC * C + cM * M).tocsr()
    b = time_rhs(algo, params, K, C, M, F, u_n, v_n, a_n)
    vals = np.zeros(len(dir_vals)) if algo == "euler_explicit" else np.asarray(dir_vals, float)
    x = solve_r1(A, b, np.asarray(dir_dofs, int), vals, Ndof, orphan_dofs)
    u, v, a = time_update(algo, params, u_n, v_n, a_n, x)
    return u, (v_n if v is None else v), (a_n if a is None else a)


class DynamicElasticOracle(ElasticOracle):
    """``Simulations.Elastic`` with a time scheme: mass ``M_e = UV(rho,
    dof_n = dim)`` over the mass rule and Rayleigh ``C = coefK K + coefM M``
    (``Simulations/_elastic.py:123-156``), element matrices times the
    thickness in 2D."""

    def __init__(self, *args, rho=1.0, rayleigh=(0.0, 0.0), **kw):
        super().__init__(*args, **kw)
        self.rho = rho
        self.rayleigh = rayleigh  # (coefM, coefK)
        self.speed = np.zeros(self.Ndof)
        self.accel = np.zeros(self.Ndof)

    def Get_M(self):
        geo = Geometry(self.elemType, self.connect, self.coord, "mass", self.dim)
        Me = UV(geo, self.rho, self.dim)
        if self.dim == 2:
            Me = Me * self.thickness
        if self._cmap is None:
            self._cmap = csr_map(self.connect, self.dim, self.Ndof)
        return assemble_csr(Me, self._cmap, self.Ndof)

    def Get_K_C_M(self):
        K, M = self.Get_K(), self.Get_M()
        cM, cK = self.rayleigh
        return K, (cK * K + cM * M).tocsr(), M

    def Step(self, algo, params):
        K, C, M = self.Get_K_C_M()
        F = self.neumann_vector("elastic", self.Ndof)
        dofs, vals = self.dir.get("elastic", (np.array([], int), np.array([])))
        orphan = dofs_nodes(self.orphan, self.unknowns("elastic"), self.unknowns("elastic"))
        self.u, self.speed, self.accel = time_step(algo, params, K, C, M, F, self.u, self.speed, self.accel, dofs,
                                                   vals, orphan)
        return self.u


class ThermalOracle(_SimuBase):
    """``Simulations.Thermal`` (``Simulations/_thermal.py``): ``K_e =
    GradUGradV(k)`` over the rigi rule and ``C_e = UV(rho c)`` over the mass
    rule, neither scaled by the thickness (the simulation's dimension is the
    model's, 1; ``_simu.py:621``, ``_thermal.py:130-134``); loads on 2D meshes
    are (``add_surfLoad``, ``_simu.py:2572-2630``)."""

    def __init__(self, connect, coord, elemType, k, c, rho=1.0, thickness=1.0, boundary=None):
        dim = ELEM_DIM[elemType]
        super().__init__(connect, coord, elemType, dim, boundary or {})
        self.k, self.c, self.rho, self.thickness = k, c, rho, thickness
        self.cmap = csr_map(self.connect, 1, self.Nn)
        self.u = np.zeros(self.Nn)
        self.speed = np.zeros(self.Nn)
        self.accel = np.zeros(self.Nn)

    def unknowns(self, problemType):
        return ["t"]

    def add_surfLoad(self, nodes, values, unknowns, problemType="thermal"):
        super().add_surfLoad(nodes, values, unknowns, problemType, self.thickness)

    def add_volumeLoad(self, nodes, values, unknowns, problemType="thermal"):
        super().add_volumeLoad(nodes, values, unknowns, problemType, self.thickness)

    def Get_K_C(self):
        geo_r = Geometry(self.elemType, self.connect, self.coord, "rigi", self.dim)
        geo_m = Geometry(self.elemType, self.connect, self.coord, "mass", self.dim)
        Ke = GradU_A_GradV(geo_r, np.eye(self.dim), self.k)
        Ce = UV(geo_m, np.asarray(self.rho) * np.asarray(self.c), 1)
        return assemble_csr(Ke, self.cmap, self.Nn), assemble_csr(Ce, self.cmap, self.Nn)

    def Step(self, algo="elliptic", params=None):
        K, C = self.Get_K_C()
        F = self.neumann_vector("thermal", self.Nn)
        dofs, vals = self.dir.get("thermal", (np.array([], int), np.array([])))
        orphan = dofs_nodes(self.orphan, ["t"], ["t"])
        if algo == "elliptic":
            self.u = solve_r1(K, F, dofs, vals, self.Nn, orphan)
        else:
            M = sp.csr_matrix(K.shape)
            self.u, self.speed, self.accel = time_step(algo, params, K, C, M, F, self.u, self.speed, self.accel,
                                                       dofs, vals, orphan)
        return self.u
