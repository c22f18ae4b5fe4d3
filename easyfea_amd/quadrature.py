"""Integration points, weights and shape-function tables.

The device kernels receive these tables as a small constant block, so this
module is the single host-side source of the reference-element data.

Point orderings and weights follow ``EasyFEA/FEM/_gauss.py``:
``_Triangle`` (:55-108), ``_Quadrangle`` (:111-139), ``_Tetrahedron``
(:142-198), ``_Hexahedron`` (:201-265) and the per-type selection in
``Gauss_factory`` (:364-516).  Shape functions follow
``EasyFEA/FEM/Elems/_seg.py`` (SEG2), ``_tri.py`` (TRI3), ``_quad.py``
(QUAD4), ``_tetra.py`` (TETRA4/TETRA10) and ``_hexa.py`` (HEXA8).
"""

from functools import lru_cache

import numpy as np

from ._utils import ElemType, MatrixType


def _npg(elemType: ElemType, matrixType: MatrixType) -> int:
    rigi = MatrixType(matrixType) == MatrixType.rigi
    table = {
        ElemType.SEG2: (1, 2),
        ElemType.TRI3: (1, 3),
        ElemType.QUAD4: (4, 4),
        ElemType.TETRA4: (1, 4),
        ElemType.TETRA10: (4, 4),
        ElemType.HEXA8: (8, 8),
    }
    r, m = table[ElemType(elemType)]
    return r if rigi else m


@lru_cache(maxsize=None)
def gauss_points(elemType: ElemType, matrixType: MatrixType):
    """Returns ``(coord (nPg, dim), weights (nPg,))`` for an element type."""
    elemType = ElemType(elemType)
    nPg = _npg(elemType, matrixType)
    if elemType == ElemType.SEG2:
        x, w = np.polynomial.legendre.leggauss(nPg)
        coord = x.reshape(nPg, 1)
    elif elemType == ElemType.TRI3:
        if nPg == 1:
            coord = np.array([[1 / 3, 1 / 3]])
            w = np.array([1 / 2])
        else:
            coord = np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]])
            w = np.full(3, 1 / 6)
    elif elemType == ElemType.QUAD4:
        a = 1 / np.sqrt(3)
        coord = np.array([[-a, -a], [a, -a], [a, a], [-a, a]])
        w = np.ones(4)
    elif elemType in (ElemType.TETRA4, ElemType.TETRA10):
        if nPg == 1:
            coord = np.array([[1 / 4, 1 / 4, 1 / 4]])
            w = np.array([1 / 6])
        else:
            a = (5 - np.sqrt(5)) / 20
            b = (5 + 3 * np.sqrt(5)) / 20
            coord = np.array([[a, a, a], [a, a, b], [a, b, a], [b, a, a]])
            w = np.full(4, 1 / 24)
    elif elemType == ElemType.HEXA8:
        a = 1 / np.sqrt(3)
        x = [-a, -a, -a, -a, a, a, a, a]
        y = [-a, -a, a, a, -a, -a, a, a]
        z = [-a, a, -a, a, -a, a, -a, a]
        coord = np.array([x, y, z]).T
        w = np.ones(8)
    else:  # pragma: no cover
        raise NotImplementedError(elemType)
    coord = np.ascontiguousarray(coord, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    coord.setflags(write=False)
    w.setflags(write=False)
    return coord, w


def _N(elemType: ElemType, p: np.ndarray) -> np.ndarray:
    """Shape functions at one reference point, ``(nPe,)``."""
    if elemType == ElemType.SEG2:
        (r,) = p
        return np.array([-(r - 1) / 2, (r + 1) / 2])
    if elemType == ElemType.TRI3:
        r, s = p
        return np.array([-r - s + 1, r, s])
    if elemType == ElemType.QUAD4:
        r, s = p
        return np.array(
            [
                (r - 1) * (s - 1) / 4,
                -(r + 1) * (s - 1) / 4,
                (r + 1) * (s + 1) / 4,
                -(r - 1) * (s + 1) / 4,
            ]
        )
    if elemType == ElemType.TETRA4:
        x, y, z = p
        return np.array([1 - x - y - z, x, y, z])
    if elemType == ElemType.TETRA10:
        x, y, z = p
        L0 = 1 - x - y - z
        return np.array(
            [
                L0 * (2 * L0 - 1),
                x * (2 * x - 1),
                y * (2 * y - 1),
                z * (2 * z - 1),
                4 * x * L0,
                4 * x * y,
                4 * y * L0,
                4 * z * L0,
                4 * y * z,
                4 * x * z,
            ]
        )
    if elemType == ElemType.HEXA8:
        r, s, t = p
        return np.array(
            [
                -(r - 1) * (s - 1) * (t - 1) / 8,
                (r + 1) * (s - 1) * (t - 1) / 8,
                -(r + 1) * (s + 1) * (t - 1) / 8,
                (r - 1) * (s + 1) * (t - 1) / 8,
                (r - 1) * (s - 1) * (t + 1) / 8,
                -(r + 1) * (s - 1) * (t + 1) / 8,
                (r + 1) * (s + 1) * (t + 1) / 8,
                -(r - 1) * (s + 1) * (t + 1) / 8,
            ]
        )
    raise NotImplementedError(elemType)  # pragma: no cover


def _dN(elemType: ElemType, p: np.ndarray) -> np.ndarray:
    """Reference derivatives at one point, ``(dim, nPe)``."""
    if elemType == ElemType.SEG2:
        return np.array([[-1 / 2, 1 / 2]])
    if elemType == ElemType.TRI3:
        return np.array([[-1.0, 1.0, 0.0], [-1.0, 0.0, 1.0]])
    if elemType == ElemType.QUAD4:
        r, s = p
        return np.array(
            [
                [s / 4 - 1 / 4, 1 / 4 - s / 4, s / 4 + 1 / 4, -s / 4 - 1 / 4],
                [r / 4 - 1 / 4, -r / 4 - 1 / 4, r / 4 + 1 / 4, 1 / 4 - r / 4],
            ]
        )
    if elemType == ElemType.TETRA4:
        return np.array(
            [[-1.0, 1.0, 0.0, 0.0], [-1.0, 0.0, 1.0, 0.0], [-1.0, 0.0, 0.0, 1.0]]
        )
    if elemType == ElemType.TETRA10:
        x, y, z = p
        L0 = 1 - x - y - z
        d0 = -(4 * L0 - 1)
        return np.array(
            [
                [d0, 4 * x - 1, 0, 0, 4 * (L0 - x), 4 * y, -4 * y, -4 * z, 0, 4 * z],
                [d0, 0, 4 * y - 1, 0, -4 * x, 4 * x, 4 * (L0 - y), -4 * z, 4 * z, 0],
                [d0, 0, 0, 4 * z - 1, -4 * x, 0, -4 * y, 4 * (L0 - z), 4 * y, 4 * x],
            ],
            dtype=float,
        )
    if elemType == ElemType.HEXA8:
        r, s, t = p
        return np.array(
            [
                [
                    -(s - 1) * (t - 1) / 8,
                    (s - 1) * (t - 1) / 8,
                    -(s + 1) * (t - 1) / 8,
                    (s + 1) * (t - 1) / 8,
                    (s - 1) * (t + 1) / 8,
                    -(s - 1) * (t + 1) / 8,
                    (s + 1) * (t + 1) / 8,
                    -(s + 1) * (t + 1) / 8,
                ],
                [
                    -(r - 1) * (t - 1) / 8,
                    (r + 1) * (t - 1) / 8,
                    -(r + 1) * (t - 1) / 8,
                    (r - 1) * (t - 1) / 8,
                    (r - 1) * (t + 1) / 8,
                    -(r + 1) * (t + 1) / 8,
                    (r + 1) * (t + 1) / 8,
                    -(r - 1) * (t + 1) / 8,
                ],
                [
                    -(r - 1) * (s - 1) / 8,
                    (r + 1) * (s - 1) / 8,
                    -(r + 1) * (s + 1) / 8,
                    (r - 1) * (s + 1) / 8,
                    (r - 1) * (s - 1) / 8,
                    -(r + 1) * (s - 1) / 8,
                    (r + 1) * (s + 1) / 8,
                    -(r - 1) * (s + 1) / 8,
                ],
            ]
        )
    raise NotImplementedError(elemType)  # pragma: no cover


@lru_cache(maxsize=None)
def N_pg(elemType: ElemType, matrixType: MatrixType) -> np.ndarray:
    """Shape functions at the integration points, ``(nPg, nPe)``."""
    coord, _ = gauss_points(elemType, matrixType)
    out = np.array([_N(ElemType(elemType), p) for p in coord], dtype=np.float64)
    out.setflags(write=False)
    return out


@lru_cache(maxsize=None)
def dN_pg(elemType: ElemType, matrixType: MatrixType) -> np.ndarray:
    """Reference derivatives at the integration points, ``(nPg, dim, nPe)``."""
    coord, _ = gauss_points(elemType, matrixType)
    out = np.array([_dN(ElemType(elemType), p) for p in coord], dtype=np.float64)
    out.setflags(write=False)
    return out
