"""Element and matrix type enumerations.

Mirrors the names of ``EasyFEA/FEM/_utils.py`` (``ElemType``, ``MatrixType``)
for the element families the device path supports: TRI3, QUAD4, HEXA8 and
TETRA10 volume/surface groups, plus SEG2 and QUAD4 boundary groups used by
Neumann loads.
"""

from enum import Enum


class ElemType(str, Enum):
    """Element types (string values follow the reference's names)."""

    SEG2 = "SEG2"
    TRI3 = "TRI3"
    QUAD4 = "QUAD4"
    TETRA4 = "TETRA4"
    TETRA10 = "TETRA10"
    HEXA8 = "HEXA8"

    def __str__(self) -> str:
        return self.name

    @property
    def dim(self) -> int:
        return _DIM[self]

    @property
    def nPe(self) -> int:
        return _NPE[self]


_DIM = {
    ElemType.SEG2: 1,
    ElemType.TRI3: 2,
    ElemType.QUAD4: 2,
    ElemType.TETRA4: 3,
    ElemType.TETRA10: 3,
    ElemType.HEXA8: 3,
}
_NPE = {
    ElemType.SEG2: 2,
    ElemType.TRI3: 3,
    ElemType.QUAD4: 4,
    ElemType.TETRA4: 4,
    ElemType.TETRA10: 10,
    ElemType.HEXA8: 8,
}


class MatrixType(str, Enum):
    """Integration rule selector (``EasyFEA/FEM/_utils.py`` ``MatrixType``)."""

    rigi = "rigi"
    mass = "mass"

    def __str__(self) -> str:
        return self.name
