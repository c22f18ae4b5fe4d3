"""Material models of the solve path.

``Elastic.Isotropic`` mirrors ``EasyFEA/Models/Elastic/_laws.py:333-505``
(``E`` scalar or per element, Lame constants, Kelvin-Mandel ``C``);
``PhaseField`` mirrors ``EasyFEA/Models/_phasefield.py:38-317`` for the
Amor and Miehe splits, AT1/AT2 regularisations and the ``History``
irreversibility.  Parameter changes notify observing simulations, which
re-assemble lazily (``Utilities/_observers.py``, ``_simu.py:972-983``).
"""

from __future__ import annotations

from enum import Enum
from typing import Union

import numpy as np


class _Observable:
    def __init__(self):
        self._observers = []

    def _Add_observer(self, obs) -> None:
        if obs not in self._observers:
            self._observers.append(obs)

    def Need_Update(self, value: bool = True) -> None:
        if value:
            for o in self._observers:
                o.Need_Update()


class Isotropic(_Observable):
    """Isotropic linear elastic material (Kelvin-Mandel notation)."""

    def __init__(self, dim: int, E=210000.0, v=0.3, planeStress=True, thickness=1.0):
        super().__init__()
        assert dim in (2, 3), "dim must be 2 or 3"
        self.dim = dim
        self.E = E
        self.v = v
        self.planeStress = planeStress
        self.thickness = thickness

    def __setattr__(self, name, value):
        if name == "E":
            arr = np.asarray(value, dtype=float)
            assert np.all(arr > 0), "E must be > 0"
            value = float(arr) if arr.ndim == 0 else arr
        elif name == "v":
            assert -1 < value < 0.5, "v must be in ]-1, 0.5["
            value = float(value)
        elif name == "thickness":
            assert value > 0, "thickness must be > 0"
            value = float(value)
        elif name == "planeStress":
            value = bool(value)
        super().__setattr__(name, value)
        if name in ("E", "v", "thickness", "planeStress") and "_observers" in self.__dict__:
            self.Need_Update()

    @property
    def isHeterogeneous(self) -> bool:
        return isinstance(self.E, np.ndarray)

    @property
    def coef(self) -> float:
        return np.sqrt(2)

    def get_lambda(self):
        E, v = self.E, self.v
        lmbda = E * v / ((1 + v) * (1 - 2 * v))
        if self.dim == 2 and self.planeStress:
            lmbda = E * v / (1 - v**2)
        return lmbda

    def get_mu(self):
        return self.E / (2 * (1 + self.v))

    def get_bulk(self):
        return self.get_lambda() + 2 * self.get_mu() / self.dim

    @property
    def C(self) -> np.ndarray:
        """Kelvin-Mandel stiffness, ``(ns, ns)`` or ``(Ne, ns, ns)``."""
        lam, mu = np.asarray(self.get_lambda()), np.asarray(self.get_mu())
        ns = 3 if self.dim == 2 else 6
        C = np.zeros(lam.shape + (ns, ns))
        d = self.dim
        for i in range(d):
            for j in range(d):
                C[..., i, j] = lam
            C[..., i, i] = lam + 2 * mu
        for k in range(d, ns):
            C[..., k, k] = 2 * mu
        return C

    def _E_array(self) -> np.ndarray:
        return np.atleast_1d(np.asarray(self.E, dtype=np.float64))


def _mandel_rotation(axis_1: np.ndarray, axis_2: np.ndarray) -> np.ndarray:
    """The 6 x 6 map P of Kelvin-Mandel strain vectors from the material
    basis (axis_1, axis_2, axis_1 x axis_2) to the global one: column k is
    the rotated Kelvin-Mandel basis tensor k, rotated as p T p^T with the
    axes as the columns of p.  ``C_global = P C_material P^T``."""
    p = np.stack([axis_1, axis_2, np.cross(axis_1, axis_2)], axis=1)
    r = np.sqrt(0.5)
    pairs = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    P = np.empty((6, 6))
    for k, (i, j) in enumerate(pairs):
        T = np.zeros((3, 3))
        if i == j:
            T[i, i] = 1.0
        else:
            T[i, j] = T[j, i] = r
        R = p @ T @ p.T
        P[:, k] = [R[0, 0], R[1, 1], R[2, 2], R[1, 2] / r, R[0, 2] / r, R[0, 1] / r]
    return P


class Anisotropic(_Observable):
    """Anisotropic linear elastic material (``Models/Elastic/_laws.py:
    1111-1259``): ``C`` (6 x 6, or one per element) given in the anisotropy
    basis (axis1, axis2), kept in Kelvin-Mandel notation in the global basis,
    ``C = P C_material P^T``.  Like the reference, a 6 x 6 ``C`` is taken as
    given whatever ``useVoigtNotation`` says (its Kelvin-Mandel scaling only
    applies to 3 x 3 inputs, ``_laws.py:1230-1246``).  3D; the device builds
    its element matrices on the matrix cores (HEXA8, TETRA10)."""

    def __init__(self, dim: int, C, useVoigtNotation: bool, axis1=(1, 0, 0), axis2=(0, 1, 0), thickness=1.0):
        super().__init__()
        assert dim == 3, "anisotropic laws run in 3D on the device"
        a1 = np.asarray(axis1, dtype=float).reshape(-1)
        a2 = np.asarray(axis2, dtype=float).reshape(-1)
        assert a1.size == 3 and a2.size == 3, "axis1 and axis2 must be 3D vectors"
        assert a1 @ a2 <= 1e-12, "axis1 and axis2 must be perpendicular"
        self.dim = dim
        self.planeStress = False
        self.thickness = float(thickness)
        self._axis1 = a1 / np.linalg.norm(a1)
        self._axis2 = a2 / np.linalg.norm(a2)
        self.Set_C(C, useVoigtNotation)

    def Set_C(self, C, useVoigtNotation: bool = True, update_S: bool = True) -> None:
        """``Anisotropic.Set_C`` (``_laws.py:1172-1188``)."""
        C = np.asarray(C, dtype=float)
        assert C.shape[-2:] == (6, 6) and C.ndim in (2, 3), "C must be (6, 6) or (Ne, 6, 6)"
        sym = np.linalg.norm(np.swapaxes(C, -1, -2) - C, axis=(-2, -1)) / np.linalg.norm(C, axis=(-2, -1))
        assert np.max(sym) <= 1e-12, "The matrix is not symmetrical."
        P = _mandel_rotation(self._axis1, self._axis2)
        self.C = np.ascontiguousarray(P @ C @ P.T)
        if update_S:
            self.S = np.linalg.inv(self.C)
        if "_observers" in self.__dict__:
            self.Need_Update()

    @property
    def isHeterogeneous(self) -> bool:
        return self.C.ndim == 3

    @property
    def coef(self) -> float:
        return np.sqrt(2)


class Elastic:
    """Namespace mirroring ``EasyFEA.Models.Elastic``."""

    Isotropic = Isotropic
    Anisotropic = Anisotropic


class PhaseField(_Observable):
    """Phase-field fracture model."""

    class ReguType(str, Enum):
        AT1 = "AT1"
        AT2 = "AT2"

        def __str__(self):
            return self.name

    class SplitType(str, Enum):
        """``PhaseField.SplitType`` (``Models/_phasefield.py:55-78``)."""

        Bourdin = "Bourdin"
        Amor = "Amor"
        Miehe = "Miehe"
        He = "He"
        Stress = "Stress"
        Zhang = "Zhang"
        AnisotStrain = "AnisotStrain"
        AnisotStrain_PM = "AnisotStrain_PM"
        AnisotStrain_MP = "AnisotStrain_MP"
        AnisotStrain_NoCross = "AnisotStrain_NoCross"
        AnisotStress = "AnisotStress"
        AnisotStress_PM = "AnisotStress_PM"
        AnisotStress_MP = "AnisotStress_MP"
        AnisotStress_NoCross = "AnisotStress_NoCross"

        def __str__(self):
            return self.name

    # the splits of the device kernels (EFEA_SPLIT_* of include/efea.h); the
    # others make g c+ + c- non-symmetric, which the PCG cannot solve
    SPLIT_CODE = {"Amor": 0, "Miehe": 1, "Bourdin": 2, "Stress": 3, "He": 4, "AnisotStrain": 5,
                  "AnisotStrain_NoCross": 6, "AnisotStress": 7, "AnisotStress_NoCross": 8, "Zhang": 9,
                  "AnisotStrain_PM": 10, "AnisotStrain_MP": 11, "AnisotStress_PM": 12, "AnisotStress_MP": 13}
    # splits whose c+ / c- (hence the displacement stiffness) are not symmetric
    NONSYMMETRIC = ("Zhang", "AnisotStrain_PM", "AnisotStrain_MP", "AnisotStress_PM", "AnisotStress_MP")

    @staticmethod
    def Get_splits() -> list:
        return list(PhaseField.SplitType)

    class SolverType(str, Enum):
        """Irreversibility (``Models/_phasefield.py:103-112``)."""

        History = "History"
        HistoryDamage = "HistoryDamage"
        BoundConstrain = "BoundConstrain"

        def __str__(self):
            return self.name

    def __init__(self, material: Isotropic, split, regularization, Gc: Union[float, np.ndarray], l0: float,
                 solver="History"):
        super().__init__()
        assert isinstance(material, Isotropic), "Must be an Isotropic elastic material"
        self.material = material
        self.split = PhaseField.SplitType(split)
        self.regularization = PhaseField.ReguType(regularization)
        self.Gc = Gc
        self.l0 = l0
        self.solver = PhaseField.SolverType(solver)
        material._Add_observer(self)

    def __setattr__(self, name, value):
        if name == "Gc":
            arr = np.asarray(value, dtype=float)
            assert np.all(arr > 0), "Gc must be > 0"
            value = float(arr) if arr.ndim == 0 else arr
        elif name == "l0":
            assert value > 0, "l0 must be > 0"
            value = float(value)
        super().__setattr__(name, value)
        if name in ("Gc", "l0", "split", "regularization") and "_observers" in self.__dict__:
            self.Need_Update()

    @property
    def dim(self) -> int:
        return self.material.dim

    @property
    def thickness(self) -> float:
        return self.material.thickness

    @property
    def isHeterogeneous(self) -> bool:
        return isinstance(self.Gc, np.ndarray)

    @property
    def k(self):
        return 3 / 4 * self.Gc * self.l0 if self.regularization == PhaseField.ReguType.AT1 else self.Gc * self.l0

    @property
    def c_w(self) -> float:
        return 8 / 3 if self.regularization == PhaseField.ReguType.AT1 else 2.0

    def _Gc_array(self) -> np.ndarray:
        return np.atleast_1d(np.asarray(self.Gc, dtype=np.float64))


class Thermal(_Observable):
    """Thermal model (``EasyFEA/Models/_thermal.py:17-56``): conductivity
    ``k`` and mass heat capacity ``c``, each a scalar or one value per
    element (>= 0), and a thickness.  Its ``dim`` is 1 (one unknown)."""

    def __init__(self, k: Union[float, np.ndarray], c: Union[float, np.ndarray] = 0.0, thickness: float = 1.0):
        super().__init__()
        self.dim = 1
        self.k = k
        self.c = c
        self.thickness = thickness

    def __setattr__(self, name, value):
        if name in ("k", "c"):
            arr = np.asarray(value, dtype=float)
            assert np.all(arr >= 0.0), f"{name} must be >= 0"
            value = float(arr) if arr.ndim == 0 else arr
        elif name == "thickness":
            assert np.ndim(value) == 0 and value >= 0.0, "thickness must be a scalar >= 0"
            value = float(value)
        super().__setattr__(name, value)
        if name in ("k", "c", "thickness") and "_observers" in self.__dict__:
            self.Need_Update()

    @property
    def isHeterogeneous(self) -> bool:
        return isinstance(self.k, np.ndarray) or isinstance(self.c, np.ndarray)
