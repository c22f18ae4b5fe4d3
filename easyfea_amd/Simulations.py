"""Simulations: the reference's ``Simulations.Elastic`` / ``Simulations.PhaseField``
API over device-resident systems.

Mirrors (EasyFEA v3.5.1):
* ``_Simu`` boundary conditions ``Bc_Init`` / ``add_dirichlet`` / ``add_neumann`` /
  ``add_lineLoad`` / ``add_surfLoad`` / ``Bc_dofs_nodes`` / ``Bc_dofs_known_unknown``
  (``Simulations/_simu.py:2190-2630``), ``Get_K_C_M_F`` returning scipy CSR copies
  (``:161-175``), ``Solve`` (``:1504-1515``), ``Calc_Reaction`` (``:212-260``),
  ``Calc_Energy`` (``:177-210``), ``Save_Iter`` / ``Set_Iter`` (``:288-471``);
* ``Elastic`` (``Simulations/_elastic.py:52-176``), static or dynamic;
* ``Thermal`` (``Simulations/_thermal.py``), steady or transient;
* ``PhaseField`` with its staggered ``Solve`` (``Simulations/_phasefield.py:86-657``);
* the time schemes ``Solver_Set_Elliptic/Parabolic/Hyperbolic_Algorithm``
  (``Simulations/_simu.py:1161-1290``) and one step per ``Solve``
  (``:1517-1640, 1735-1899``).

Assembly, scatter, solve and the phase-field updates run in ``libefea_hip.so``;
boundary-condition bookkeeping (which DOFs, which values) stays on the host
as in the reference.
"""

from __future__ import annotations

import time
from typing import Callable, Optional

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from ._utils import ElemType, MatrixType
from .device import (DeviceCsr, DeviceMesh, PendingSolve, SolveJob, Workspace, drive, lincomb, ptr, stream_ptr,
                     to_dev, to_host)
from .mesh import Mesh
from .Models import Anisotropic, Isotropic
from .Models import PhaseField as PhaseFieldModel
from .Models import Thermal as ThermalModel
from .quadrature import N_pg, dN_pg, gauss_points
from .Solvers import REFERENCE_ONLY, AlgoType, SolverType
from .time_schemes import expand, history_coefs, matrix_coefs


def _print_error(msg: str) -> None:
    print(f"\033[91m{msg}\033[0m")


def _frame_coords(coord_e: np.ndarray, dim: int) -> np.ndarray:
    """Boundary-element coordinates in the element's own orthonormal frame
    (``_group_elem.py:404-479``), ``(Ne, nPe, dim)``."""
    if dim == 1:
        t = coord_e[:, 1] - coord_e[:, 0]
        t = t / np.linalg.norm(t, axis=1, keepdims=True)
        return np.einsum("end,ed->en", coord_e, t)[..., None]
    i = coord_e[:, 1] - coord_e[:, 0]
    i = i / np.linalg.norm(i, axis=1, keepdims=True)
    k = np.cross(i, coord_e[:, -1] - coord_e[:, 0])
    k = k / np.linalg.norm(k, axis=1, keepdims=True)
    j = np.cross(k, i)
    return np.stack([np.einsum("end,ed->en", coord_e, i), np.einsum("end,ed->en", coord_e, j)], -1)


class _Bc:
    __slots__ = ("problemType", "nodes", "dofs", "values", "unknowns", "description")

    def __init__(self, problemType, nodes, dofs, values, unknowns, description):
        self.problemType = problemType
        self.nodes = np.asarray(nodes, dtype=np.int64)
        self.dofs = np.asarray(dofs, dtype=np.int64)
        self.values = np.asarray(values, dtype=np.float64)
        self.unknowns = list(unknowns)
        self.description = description


class _Simu:
    """Shared machinery of the device-backed simulations."""

    def __init__(self, mesh: Mesh, model, folder: str = "", verbosity: bool = False):
        self.mesh = mesh
        self.model = model
        self.folder = folder
        self._verbosity = verbosity
        self.dev = DeviceMesh(mesh)
        self.__solver = SolverType.hip_mixed
        self._ws = Workspace()
        self._results: list = []
        self.Niter = 0
        self.tol_residual = 1e-12
        self.tol_anorm = 1e-9
        self.replace_ratio = 0.0  # mixed mode: FP64 residual replacement at every poll
        self.maxit = 200000
        self.check_every = 32
        self.preconditioner = "auto"  # "bj" | "two-level" | "auto"
        # hip_mixed: keep FP64 Krylov products for matrices under
        # device.MIXED_MIN_MATRIX_BYTES (latency-bound); False forces FP32 values
        self.mixed_adaptive = True
        # hip_mixed two-level solves: "auto" | "values" | "refine" (device.py)
        self.mixed_strategy = "auto"
        # pipelined PCG "auto" (latency-bound systems; block Jacobi: one launch per iteration,
        # two-level: two) | True | False
        self.pipelined = "auto"
        self.max_blocks = 0  # grid cap of the solver kernels (0: the library's default)
        # keep the two-level coarse inverse over successive solves of one matrix
        # until their iteration count grows (device.COARSE_REUSE_GROWTH)
        self.coarse_reuse = True
        # read the solver state back first near the previous solve's iteration count
        self.predict_polls = True
        # largest aggregate of the two-level coarse space (None: 256 nodes in 2D, 1024 in 3D)
        self.coarse_target = None
        # time steps start the solve from the state extrapolated over the step
        self.extrapolate_guess = True
        self.last_solve = {}
        self._algo = AlgoType.elliptic
        self._parabolicParams = None
        self._hyperbolicParams = None
        self.Bc_Init()
        model._Add_observer(self)

    # -------------------------------------------------------------- solver
    @property
    def solver(self) -> SolverType:
        return self.__solver

    @solver.setter
    def solver(self, value):
        try:
            self.__solver = SolverType(value)
        except ValueError:
            extra = " (a host backend of the reference)" if str(value) in REFERENCE_ONLY else ""
            _print_error(f"The solver {value}{extra} cannot be used. The solver must be in {list(SolverType)}")

    def _mode(self) -> str:
        return "fp64" if self.__solver == SolverType.hip_fp64 else "mixed"

    # -------------------------------------------------------- time schemes
    @property
    def algo(self) -> AlgoType:
        return self._algo

    def Solver_Set_Elliptic_Algorithm(self) -> None:
        """``K u = F`` (``_simu.py:1161-1171``)."""
        self._algo = AlgoType.elliptic

    def Solver_Set_Parabolic_Algorithm(self, dt: float, alpha=1 / 2) -> None:
        """``K u^{n+1} + C v^{n+1} = F^{n+1}`` with ``u^{n+1} = u^n + dt
        v^{n+alpha}`` (``_simu.py:1173-1212``)."""
        assert dt > 0, "Time increment must be > 0"
        self._algo = AlgoType.parabolic
        self._parabolicParams = (float(dt), float(alpha))

    def Solver_Set_Hyperbolic_Algorithm(self, dt: float, algo=AlgoType.newmark, beta=0.25, gamma=0.5,
                                        alpha=0.5) -> None:
        """``K u + C v + M a = F`` with ``algo`` in ``AlgoType.Get_Hyperbolic_Types()``
        (``_simu.py:1214-1290``); ``hht_newmark`` derives beta and gamma from alpha."""
        algo = AlgoType(algo)
        types = AlgoType.Get_Hyperbolic_Types()
        assert algo in types, f"algo must be in {types}"
        assert dt > 0, "Time increment must be > 0"
        if algo == AlgoType.hht_newmark:
            assert 0 <= alpha <= 1 / 3, "hht_newmark requires alpha in [0, 1/3]"
            beta = 1 / 4 * (1 + alpha) ** 2
            gamma = 1 / 2 + alpha
        else:
            assert 0 <= alpha < 1
        self._algo = algo
        self._hyperbolicParams = (float(dt), float(beta), float(gamma), float(alpha))

    def _time_params(self) -> tuple:
        if self._algo == AlgoType.parabolic:
            return self._parabolicParams
        if self._algo in AlgoType.Get_Hyperbolic_Types():
            return self._hyperbolicParams
        return ()

    def _time_bases(self, problemType) -> dict:
        """``{"K"|"C"|"M": DeviceCsr}``: the assembled base matrices of
        ``problemType`` on one pattern (absent ones are zero)."""
        raise NotImplementedError

    def _time_damping(self, problemType) -> dict:
        """``C`` as weights of the base matrices (``{"C": 1.0}`` if C is a base)."""
        return {"C": 1.0}

    def _scheme_matrix(self, problemType, bases: dict, coefs: dict) -> DeviceCsr:
        """``A = coefK K + coefC C + coefM M`` on the device
        (``_simu.py:1893-1899``), kept while the coefficients and the base
        matrices' values stay the same.  A changed ``A`` drops its lagged
        coarse inverse."""
        if len(coefs) == 1:
            (name, c), = coefs.items()
            if c == 1.0:
                return bases[name]
        names = sorted(coefs)
        K = bases["K"]
        A = self.dev.matrix("A:" + problemType, K.dof_n)
        key = tuple((n, coefs[n], id(bases[n]), bases[n].version) for n in names)
        if getattr(A, "_scheme_key", None) != key:
            A.combine([(coefs[n], bases[n]) for n in names])
            A._scheme_key = key
            caches = self.__dict__.get("_coarse_caches")
            if caches:
                for k in [k for k in caches if k[0] == id(A)]:
                    del caches[k]
        return A

    def _time_step_gen(self, problemType, u: torch.Tensor, v: torch.Tensor, a: torch.Tensor):
        """One step of the current scheme for ``problemType`` from the state
        ``(u, v, a)``: builds ``A`` and ``b`` on the device, yields the solve,
        then returns the new ``(u, v, a)`` (``_Solver_Solve_problemType``,
        ``_simu.py:1529-1550``, with ``_Solver_Apply_Neumann`` :1735-1860,
        ``_Solver_Apply_Dirichlet`` :1862-1911 and
        ``_Solver_Update_solutions`` :1552-1640).  Not for the elliptic
        scheme, whose solve is ``K u = F`` alone."""
        algo = self._algo
        assert algo != AlgoType.elliptic
        bases = self._time_bases(problemType)
        b = self._neumann_device(problemType)
        params = self._time_params()
        damping = self._time_damping(problemType)
        acoefs = {n: c for n, c in expand(dict(zip("KCM", matrix_coefs(algo, params))), damping).items()
                  if n in bases}
        A = self._scheme_matrix(problemType, bases, acoefs)
        b = b.clone()
        w = torch.empty_like(b)
        for name, (cu, cv, ca) in expand(history_coefs(algo, params), damping).items():
            if name not in bases:
                continue
            lincomb(w, [(c, s) for c, s in ((cu, u), (cv, v), (ca, a)) if c != 0])
            lincomb(b, [(1.0, b), (1.0, bases[name].spmv(w))])
        explicit = algo == AlgoType.euler_explicit
        if explicit or not self.extrapolate_guess:
            x = (a if explicit else u).clone()
        else:
            # initial guess u + dt v + dt^2/2 a (u + dt v for the parabolic scheme):
            # the solution within the solver tolerance does not depend on it, the
            # iteration count does; the Dirichlet values replace it at the known DOFs
            dt = params[0]
            terms = [(1.0, u), (dt, v)] + ([] if algo == AlgoType.parabolic else [(dt * dt / 2, a)])
            x = lincomb(torch.empty_like(u), terms)
        # the explicit scheme's unknown is a^n: constrained DOFs have zero acceleration (_simu.py:1905-1907)
        yield from self._solve_system_gen(A, b, x, problemType, zero_dirichlet=explicit)
        u1, v1, a1 = torch.empty_like(u), torch.empty_like(v), torch.empty_like(a)
        dt, *rest = params
        beta, gamma, alpha = (0.0, 0.0, rest[0]) if algo == AlgoType.parabolic else tuple(rest)
        _lib.call("efea_time_update", _lib.SCHEME_CODE[algo.value], u.numel(), dt, beta, gamma, alpha, ptr(u),
                  ptr(v), ptr(a), ptr(x), ptr(u1), ptr(v1), ptr(a1), stream_ptr())
        return u1, v1, a1

    # ------------------------------------------------------------ problems
    def Get_problemTypes(self) -> list:
        raise NotImplementedError

    @property
    def problemType(self) -> str:
        return self.Get_problemTypes()[0]

    def Get_unknowns(self, problemType=None) -> list:
        raise NotImplementedError

    def Get_dof_n(self, problemType=None) -> int:
        return len(self.Get_unknowns(problemType))

    def Get_dofs(self, problemType=None) -> np.ndarray:
        return np.arange(self.mesh.Nn * self.Get_dof_n(problemType))

    # ---------------------------------------------------------------- BCs
    def Bc_Init(self) -> None:
        self._Bc_Dirichlet: list = []
        self._Bc_Neumann: list = []

    @property
    def Bc_Dirichlet(self) -> list:
        return list(self._Bc_Dirichlet)

    @property
    def Bc_Neuman(self) -> list:
        return list(self._Bc_Neumann)

    def _check_problemType(self, problemType):
        assert problemType in self.Get_problemTypes(), f"problemType must be in {self.Get_problemTypes()}"

    def Bc_dofs_nodes(self, nodes, unknowns, problemType=None) -> np.ndarray:
        """``BoundaryCondition.Get_dofs_nodes`` (``FEM/_boundary_conditions.py:160-201``)."""
        nodes = np.asarray(nodes, dtype=np.int64).ravel()
        if nodes.size == 0:
            return np.array([], dtype=np.int64)
        problemType = problemType or self.problemType
        avail = self.Get_unknowns(problemType)
        dofs = np.zeros((nodes.size, len(unknowns)), dtype=np.int64)
        for d, u in enumerate(unknowns):
            assert u in avail, f"{u} is not in {avail}"
            dofs[:, d] = nodes * len(avail) + avail.index(u)
        return dofs.ravel()

    def _eval_nodes(self, coord_n, value):
        if callable(value):
            return np.asarray(value(coord_n[:, 0], coord_n[:, 1], coord_n[:, 2]), dtype=float) * np.ones(len(coord_n))
        return np.broadcast_to(np.asarray(value, dtype=float), (len(coord_n),)).copy()

    def add_dirichlet(self, nodes, values: list, unknowns: list, problemType=None, description="") -> None:
        if len(nodes) == 0 or len(values) == 0:
            return
        assert len(values) == len(unknowns), f"len(values)={len(values)} must equal len(unknowns)={len(unknowns)}."
        problemType = problemType or self.problemType
        self._check_problemType(problemType)
        nodes = np.asarray(nodes, dtype=np.int64)
        coord_n = self.mesh.coord[nodes]
        vals = np.zeros((nodes.size, len(unknowns)))
        for d in range(len(unknowns)):
            vals[:, d] = self._eval_nodes(coord_n, values[d])
        dofs = self.Bc_dofs_nodes(nodes, unknowns, problemType)
        self._Bc_Dirichlet.append(_Bc(problemType, nodes, dofs, vals.ravel(), unknowns, f"Dirichlet {description}"))

    def add_neumann(self, nodes, values: list, unknowns: list, problemType=None, description="") -> None:
        """Point loads shared equally by the nodes (``_simu.py:2471-2517, 2743-2768``)."""
        if len(nodes) == 0 or len(values) == 0:
            return
        assert len(values) == len(unknowns)
        problemType = problemType or self.problemType
        nodes = np.asarray(nodes, dtype=np.int64)
        coord_n = self.mesh.coord[nodes]
        vals = np.zeros((nodes.size, len(unknowns)))
        for d in range(len(unknowns)):
            vals[:, d] = self._eval_nodes(coord_n, values[d]) / len(nodes)
        dofs = self.Bc_dofs_nodes(nodes, unknowns, problemType)
        self._Bc_Neumann.append(_Bc(problemType, nodes, dofs, vals.ravel(), unknowns, f"Neumann {description}"))

    def _integrate_boundary(self, bdim: int, nodes, values, unknowns, problemType):
        """Load integral over the elements of dimension ``bdim`` whose nodes all
        belong to ``nodes`` (``_simu.py:2770-2842``), integrated on the device."""
        nodes = np.asarray(nodes, dtype=np.int64)
        all_vals, all_dofs, used = [], [], []
        for g in self.mesh.Get_list_groupElem(bdim):
            elements = g.Get_Elements_Nodes(nodes, exclusively=True)
            if elements.size == 0:
                continue
            connect = g.connect[elements]
            used.append(connect.ravel())
            et = g.elemType
            mt = MatrixType.mass
            coord_e = self.mesh.coord[connect]
            inDim = 2 if self.mesh.dim == 2 else 3
            if bdim == inDim:
                Xloc = coord_e[..., :bdim]
            else:
                Xloc = _frame_coords(coord_e, bdim)
            Npg = N_pg(et, mt)
            nPg = Npg.shape[0]
            coord_gp = np.einsum("pn,end->epd", Npg, coord_e)
            tab = np.concatenate([gauss_points(et, mt)[1], Npg.ravel(), dN_pg(et, mt).ravel()])
            d_X = to_dev(np.ascontiguousarray(Xloc, dtype=np.float64))
            d_tab = to_dev(tab)
            vals = np.zeros((connect.shape[0] * connect.shape[1], len(unknowns)))
            dofs = np.zeros_like(vals, dtype=np.int64)
            for u, unk in enumerate(unknowns):
                val = values[u]
                if callable(val):
                    f = np.asarray(val(coord_gp[..., 0], coord_gp[..., 1], coord_gp[..., 2]), dtype=float)
                    f = np.broadcast_to(f, (connect.shape[0], nPg))
                elif np.ndim(val) == 0:
                    f = np.full((connect.shape[0], nPg), float(val))
                else:  # nodal values
                    ev = np.zeros(self.mesh.Nn)
                    ev[nodes] = val
                    f = np.einsum("pn,en->ep", Npg, ev[connect])
                d_f = to_dev(np.ascontiguousarray(f, dtype=np.float64))
                out = torch.empty(connect.size, dtype=torch.float64, device="cuda")
                _lib.call("efea_boundary_load", _lib.ELEM_CODE[et.value], connect.shape[0], ptr(d_X), ptr(d_tab),
                          nPg, ptr(d_f), ptr(out), stream_ptr())
                vals[:, u] = out.cpu().numpy()
                dofs[:, u] = self.Bc_dofs_nodes(connect.ravel(), [unk], problemType)
            all_vals.append(vals.ravel())
            all_dofs.append(dofs.ravel())
        if not all_vals:
            return np.array([]), np.array([], dtype=np.int64), np.array([], dtype=np.int64)
        return np.concatenate(all_vals), np.concatenate(all_dofs), np.concatenate(used)

    def add_lineLoad(self, nodes, values: list, unknowns: list, problemType=None, description="") -> None:
        if len(nodes) == 0 or len(values) == 0:
            return
        assert len(values) == len(unknowns)
        problemType = problemType or self.problemType
        vals, dofs, used = self._integrate_boundary(1, nodes, values, unknowns, problemType)
        self._Bc_Neumann.append(_Bc(problemType, used, dofs, vals, unknowns, f"Neumann {description}"))

    def add_surfLoad(self, nodes, values: list, unknowns: list, problemType=None, description="") -> None:
        """2D: line load times thickness; 3D: surface load (``_simu.py:2572-2630``)."""
        if len(nodes) == 0 or len(values) == 0:
            return
        assert len(values) == len(unknowns)
        problemType = problemType or self.problemType
        if self.mesh.dim == 2:
            vals, dofs, used = self._integrate_boundary(1, nodes, values, unknowns, problemType)
            vals = vals * self.model.thickness
        else:
            vals, dofs, used = self._integrate_boundary(2, nodes, values, unknowns, problemType)
        self._Bc_Neumann.append(_Bc(problemType, used, dofs, vals, unknowns, f"Neumann {description}"))

    def add_volumeLoad(self, nodes, values: list, unknowns: list, problemType=None, description="") -> None:
        """Body load integrated over the main elements whose every node is in
        ``nodes``, times the thickness in 2D (``_simu.py:2683-2741``).
        Values are constants, callables of ``(x, y, z)`` at the Gauss points
        or arrays of nodal values."""
        if len(nodes) == 0 or len(values) == 0:
            return
        assert len(values) == len(unknowns), f"len(values)={len(values)} must equal len(unknowns)={len(unknowns)}."
        problemType = problemType or self.problemType
        self._check_problemType(problemType)
        if self.mesh.dim not in (2, 3):
            raise NotImplementedError("Unknown configuration.")
        vals, dofs, used = self._integrate_boundary(self.mesh.dim, nodes, values, unknowns, problemType)
        if self.mesh.dim == 2:
            vals = vals * self.model.thickness
        self._Bc_Neumann.append(_Bc(problemType, used, dofs, vals, unknowns, f"Neumann {description}"))

    def add_pressureLoad(self, nodes, magnitude: float, problemType=None, description="") -> None:
        """Pressure ``magnitude`` along the nodal normals of the boundary
        elements whose every node is in ``nodes`` (``_simu.py:2632-2681,
        2910-2944``): the normals of ``Mesh.Get_normals`` times the magnitude
        (times the thickness in 2D) integrated as nodal values.  The sign
        follows the normals, hence the node order of the boundary elements.
        Refused on 1D meshes and scalar problems, as the reference does."""
        if len(nodes) == 0 or magnitude == 0:
            return
        problemType = problemType or self.problemType
        self._check_problemType(problemType)
        if self.mesh.dim == 1:
            print("Cant apply pressure on 1D mesh.")
            return
        if self.Get_dof_n(problemType) == 1:
            print("Cant apply pressure on scalar problems.")
            return
        if self.mesh.dim == 2:
            magnitude = magnitude * self.model.thickness
        normals, used = self.mesh.Get_normals(np.asarray(nodes, dtype=np.int64))
        if used.size == 0:
            return
        inDim = self.mesh.inDim
        values = [normals[:, d] * magnitude for d in range(inDim)]
        unknowns = self.Get_unknowns(problemType)[:inDim]
        vals, dofs, used = self._integrate_boundary(self.mesh.dim - 1, used, values, unknowns, problemType)
        self._Bc_Neumann.append(_Bc(problemType, used, dofs, vals, unknowns, f"Neumann {description}"))

    def _bc_arrays(self, store, problemType):
        bcs = [bc for bc in store if bc.problemType == problemType]
        if not bcs:
            return np.array([], dtype=np.int64), np.array([])
        return np.concatenate([bc.dofs for bc in bcs]), np.concatenate([bc.values for bc in bcs])

    def Bc_dofs_Dirichlet(self, problemType=None) -> np.ndarray:
        return self._bc_arrays(self._Bc_Dirichlet, problemType or self.problemType)[0]

    def Bc_values_Dirichlet(self, problemType=None) -> np.ndarray:
        return self._bc_arrays(self._Bc_Dirichlet, problemType or self.problemType)[1]

    def Bc_dofs_Neumann(self, problemType=None) -> np.ndarray:
        return self._bc_arrays(self._Bc_Neumann, problemType or self.problemType)[0]

    def Bc_values_Neumann(self, problemType=None) -> np.ndarray:
        return self._bc_arrays(self._Bc_Neumann, problemType or self.problemType)[1]

    def Bc_dofs_known_unknown(self, problemType) -> tuple:
        """Known = unique sorted Dirichlet DOFs (``_simu.py:2295-2316``)."""
        nDof = self.mesh.Nn * self.Get_dof_n(problemType)
        mask = np.ones(nDof, dtype=bool)
        mask[self.Bc_dofs_Dirichlet(problemType)] = False
        return np.nonzero(~mask)[0], np.nonzero(mask)[0]

    def _neumann_device(self, problemType) -> torch.Tensor:
        """``_neumann_vector`` on the device, kept while the Neumann conditions
        stay the same (a zero vector when there are none)."""
        dofs, vals = self._bc_arrays(self._Bc_Neumann, problemType)
        key = (dofs.tobytes(), vals.tobytes())
        cache = self.__dict__.setdefault("_neumann_cache", {})
        c = cache.get(problemType)
        if c is None or c[0] != key:
            Ndof = self.mesh.Nn * self.Get_dof_n(problemType)
            v = torch.zeros(Ndof, dtype=torch.float64, device="cuda") if dofs.size == 0 else \
                to_dev(self._neumann_vector(problemType))
            c = cache[problemType] = (key, v)
        return c[1]

    def _neumann_vector(self, problemType) -> np.ndarray:
        dofs, vals = self._bc_arrays(self._Bc_Neumann, problemType)
        Ndof = self.mesh.Nn * self.Get_dof_n(problemType)
        # duplicates add, as the (Ndof, 1) CSR of _Solver_Apply_Neumann (_simu.py:1750-1756)
        return np.bincount(dofs, weights=vals, minlength=Ndof) if dofs.size else np.zeros(Ndof)

    def _dirichlet_state(self, problemType):
        """(known mask uint8 on device, x values at known DOFs on host)."""
        Ndof = self.mesh.Nn * self.Get_dof_n(problemType)
        dofs, vals = self._bc_arrays(self._Bc_Dirichlet, problemType)
        xk = np.bincount(dofs, weights=vals, minlength=Ndof) if dofs.size else np.zeros(Ndof)
        known = np.zeros(Ndof, dtype=np.uint8)
        known[dofs] = 1
        orphan = self.mesh.orphanNodes
        if orphan.size:
            od = self.Bc_dofs_nodes(orphan, self.Get_unknowns(problemType), problemType)
            known[od] = 1  # orphan rows hold only the unit diagonal (_simu.py:1949-1958): x = 0
            xk[od] = np.where(np.isin(od, dofs), xk[od], 0.0)
        return known, xk

    def _make_job(self, A, b_dev: torch.Tensor, x_dev: torch.Tensor, problemType,
                  zero_dirichlet: bool = False) -> SolveJob:
        """The solve of ``problemType``: Dirichlet values (zeros with
        ``zero_dirichlet``) written into ``x_dev`` at the known DOFs, the
        known mask on the device."""
        # One device mask per problem type, kept while the Dirichlet DOFs stay the
        # same (the tensor's identity then also keys the lagged coarse inverse):
        # known = the Dirichlet DOFs and the orphan DOFs, sorted; per solve only
        # the values at the known DOFs are formed (duplicates add, orphans 0
        # unless prescribed), as _dirichlet_state does over the full vector.
        dofs, vals = self._bc_arrays(self._Bc_Dirichlet, problemType)
        key = dofs.tobytes()
        cache = self.__dict__.setdefault("_mask_cache", {})
        inv_cache = self.__dict__.setdefault("_mask_inverse", {})
        if cache.get(problemType, (None,))[0] != key:
            Ndof = self.mesh.Nn * self.Get_dof_n(problemType)
            orphan = self.mesh.orphanNodes
            od = self.Bc_dofs_nodes(orphan, self.Get_unknowns(problemType), problemType) if orphan.size else \
                np.array([], dtype=np.int64)
            idx = np.unique(np.concatenate([dofs.astype(np.int64), np.asarray(od, dtype=np.int64)]))
            known_np = np.zeros(Ndof, dtype=np.uint8)
            known_np[idx] = 1
            cache[problemType] = (key, to_dev(known_np), to_dev(idx), idx)
            inv_cache[problemType] = np.searchsorted(idx, dofs)
        _, known, idx_dev, idx = cache[problemType]
        if idx.size and zero_dirichlet:
            x_dev.index_fill_(0, idx_dev, 0.0)
        elif idx.size:
            xk_idx = np.bincount(inv_cache[problemType], weights=vals, minlength=idx.size) if dofs.size else \
                np.zeros(idx.size)
            # pinned staging, copied without a host wait; the buffer is rewritten only
            # at this problem's next solve, after this solve (which follows the copy
            # on the device) has finished
            pins = self.__dict__.setdefault("_dirichlet_pinned", {})
            pin = pins.get(problemType)
            if pin is None or pin.numel() != idx.size:
                pin = pins[problemType] = torch.empty(idx.size, dtype=torch.float64, pin_memory=True)
            pin.numpy()[:] = xk_idx
            x_dev.index_copy_(0, idx_dev, pin.to("cuda", non_blocking=True))
        prev = self.last_solve.get(problemType)
        expected = prev.iterations if self.predict_polls and prev is not None and prev.status == 1 else None
        job = SolveJob(A, b_dev, x_dev, known, self._mode(), self.tol_residual, self.tol_anorm,
                       self.replace_ratio, self.maxit, self.check_every, adaptive=self.mixed_adaptive,
                       owner=self, precond=self.preconditioner, max_blocks=self.max_blocks,
                       expected_iterations=expected, strategy=self.mixed_strategy, pipelined=self.pipelined,
                       coarse_target=self.coarse_target)
        job.isolate = self._isolation(problemType)
        job.method = self._solve_method(problemType)
        job.defer = self._defer(problemType)
        job.before_solve = self._before_solve(problemType)
        return job

    def _before_solve(self, problemType):
        """``SolveJob.before_solve`` of ``problemType``'s solve (none outside
        the phase field)."""
        return None

    def _defer(self, problemType) -> bool:
        """Whether the single-system driver may queue ``problemType``'s solve
        and hand it back unfinished (``SolveJob.defer``); never outside the
        phase field."""
        return False

    def _solve_method(self, problemType) -> str:
        """The Krylov method of ``problemType``'s system: "pcg" (symmetric)."""
        return "pcg"

    def _isolation(self, problemType):
        """Nodes a stalled solve of ``problemType`` may give aggregates of
        their own (``SolveJob.isolate``); none outside the phase field."""
        return None

    def _solve_system_gen(self, A, b_dev: torch.Tensor, x_dev: torch.Tensor, problemType,
                          zero_dirichlet: bool = False):
        """Generator form of a solve: yields the ``SolveJob`` and receives its
        ``SolveInfo`` (from the single-system path or a batch executor)."""
        job = self._make_job(A, b_dev, x_dev, problemType, zero_dirichlet)
        t = time.perf_counter()
        info = yield job
        if isinstance(info, PendingSolve):
            # queued without a wait: _finish_pending ends it
            self._pending = (info, problemType, t)
            return
        self._after_solve(job, info, problemType, t)

    def _finish_pending(self) -> bool:
        """Ends the solve the single-system driver left queued, if any; True
        when its solution changed after the queued iterations, so that work
        queued from it since must be redone."""
        p = self.__dict__.pop("_pending", None)
        if p is None:
            return False
        pending, problemType, t = p
        info = pending.finish()
        self._after_solve(pending.job, info, problemType, t)
        return pending.resumed

    def _after_solve(self, job: SolveJob, info, problemType, t: float) -> None:
        job.record(info)
        self.last_job = job
        self.last_solve[problemType] = info
        if info.status != 1:
            raise _lib.EfeaError("solve", _lib.ERR_NOCONV,
                                 f"{problemType}: PCG status {info.status} after {info.iterations} iterations")
        if self._verbosity:
            print(f"{problemType}: {info.iterations} it, rel res {info.rel_residual:.2e} "
                  f"({time.perf_counter() - t:.3f} s)")

    def _reshape_values(self, values, nodeValues: bool) -> np.ndarray:
        """``Results_Reshape_values`` (``Simulations/_simu.py:3095-3153``):
        element values go to the nodes by ``Mesh.Get_Node_Values`` (on the
        device), node values to the elements as the mean over their nodes."""
        Nn, Ne = self.mesh.Nn, self.dev.Ne
        is_t = isinstance(values, torch.Tensor)
        size = values.numel() if is_t else np.asarray(values).size
        is1d = values.dim() == 1 if is_t else np.asarray(values).ndim == 1
        if nodeValues:
            if size % Nn == 0:
                v = values.cpu().numpy() if is_t else np.asarray(values)
                return v.ravel() if is1d else v.reshape(Nn, -1)
            if size % Ne == 0:
                vt = values if is_t else to_dev(np.asarray(values, dtype=np.float64))
                vn = self.dev.node_values(vt.reshape(Ne, -1)).cpu().numpy()
                return vn.ravel() if is1d else vn
        else:
            if size % Ne == 0:
                v = values.cpu().numpy() if is_t else np.asarray(values)
                return v.ravel() if is1d else v.reshape(Ne, -1)
            if size % Nn == 0:
                v = (values.cpu().numpy() if is_t else np.asarray(values)).reshape(Nn, -1)
                ve = v[self.dev.connect_np].mean(axis=1)
                return ve.ravel() if is1d else ve
        raise Exception("Unexpected conditions occurred during the calculation.")

    def _solve_system(self, A, b_dev: torch.Tensor, x_dev: torch.Tensor, problemType) -> None:
        drive(self._solve_system_gen(A, b_dev, x_dev, problemType))
        self._finish_pending()

    # ------------------------------------------------------------ results
    def Save_Iter(self, iter: Optional[dict] = None) -> None:
        it = {} if iter is None else dict(iter)
        self._results.append(it)
        self.Niter += 1

    @property
    def results(self) -> list:
        return self._results

    def Set_Iter(self, iter: int = -1) -> Optional[dict]:
        if not self._results:
            return None
        return self._results[iter]

    def Save(self, folder: str, filename: str = "simulation", additionalInfos: str = "") -> str:
        """Writes the mesh, the simulation (model, solver settings, saved
        iterations, current fields) and a summary to ``folder``
        (``_simu.py:3157-3214``) as ``.npz`` data; ``Load_Simu`` reads it."""
        from .sim_io import save_simu

        self.folder = folder
        return save_simu(self, folder, filename, additionalInfos)

    def Need_Update(self, value: bool = True) -> None:
        raise NotImplementedError

    def Calc_Energy(self, A: sp.csr_matrix, x: np.ndarray, dofs=None) -> float:
        if dofs is None:
            dofs = self.Get_dofs()
        return 0.5 * x[dofs] @ (A[dofs] @ x)


class Elastic(_Simu):
    """Linear elastic simulation (``Simulations/_elastic.py``)."""

    def __init__(self, mesh: Mesh, model: Isotropic, folder: str = "", verbosity=False):
        assert isinstance(model, (Isotropic, Anisotropic)), "model must be a elastic model"
        self._needUpdate = True
        super().__init__(mesh, model, folder, verbosity)
        assert mesh.dim == model.dim, "mesh and material dimensions differ"
        self._rho = 1.0
        self._rayleigh = (0.0, 0.0)  # (coefM, coefK) of C = coefK K + coefM M
        self._K = None
        self._M = None
        n = mesh.Nn * model.dim
        self._u = torch.zeros(n, dtype=torch.float64, device="cuda")
        self._v = torch.zeros(n, dtype=torch.float64, device="cuda")
        self._a = torch.zeros(n, dtype=torch.float64, device="cuda")

    @property
    def rho(self):
        """Density, a scalar or one value per element (``_simu.py:706``)."""
        return self._rho

    @rho.setter
    def rho(self, value):
        arr = np.asarray(value, dtype=float)
        assert np.all(arr >= 0.0), "rho must be >= 0"
        self._rho = float(arr) if arr.ndim == 0 else arr
        self.Need_Update()

    def Set_Rayleigh_Damping_Coefs(self, coefM=0.0, coefK=0.0) -> None:
        """Damping ``C = coefK K + coefM M`` (``Simulations/_elastic.py:152-156``)."""
        self._rayleigh = (float(coefM), float(coefK))
        self.Need_Update()

    def Get_problemTypes(self) -> list:
        return ["elastic"]

    def Get_unknowns(self, problemType=None) -> list:
        return ["x", "y"] if self.model.dim == 2 else ["x", "y", "z"]

    @property
    def material(self) -> Isotropic:
        return self.model

    @property
    def needUpdate(self) -> bool:
        return self._needUpdate

    def Need_Update(self, value: bool = True) -> None:
        self._needUpdate = value
        if value:
            self._M = None

    def Assembly(self):
        """Element stiffness on the device and deterministic scatter."""
        dev = self.dev
        A = dev.csr_map(self.model.dim)
        Ke = A.new_entries()
        tab, nPg = dev.tables(MatrixType.rigi)
        if isinstance(self.model, Anisotropic):
            # a general constitutive matrix: B^T C B on the matrix cores
            if dev.code not in (_lib.ELEM_CODE["HEXA8"], _lib.ELEM_CODE["TETRA10"]):
                raise NotImplementedError("anisotropic laws: HEXA8 and TETRA10 meshes on the device")
            C = to_dev(np.ascontiguousarray(self.model.C).reshape(-1))
            _lib.call("efea_elastic_ke_dense", dev.code, dev.Ne, ptr(dev.connect), ptr(dev.coord), ptr(tab), nPg,
                      ptr(C), 1 if self.model.isHeterogeneous else 0, None, 0, 0.0, ptr(Ke), stream_ptr())
        else:
            E = to_dev(self.model._E_array())
            _lib.call("efea_elastic_ke", dev.code, 64, dev.Ne, ptr(dev.connect), ptr(dev.coord), ptr(tab), nPg,
                      ptr(E), 1 if self.model.isHeterogeneous else 0, self.model.v, int(self.model.planeStress),
                      self.model.thickness, ptr(Ke), stream_ptr())
        A.scatter(Ke)
        self._K = A
        self._needUpdate = False
        return A

    def _device_K(self):
        if self._needUpdate or self._K is None:
            self.Assembly()
        return self._K

    def _device_M(self) -> DeviceCsr:
        """Mass matrix ``sum rho wJ N^T N`` (x thickness in 2D) over the mass
        rule (``Simulations/_elastic.py:135-142``), assembled on first use
        after each update."""
        self._device_K()  # brings the simulation up to date (clears a stale M)
        if self._M is None:
            dev = self.dev
            M = dev.matrix("M", self.model.dim)
            Me = M.new_entries()
            tab, nPg = dev.tables(MatrixType.mass)
            rho = self._rho
            het = isinstance(rho, np.ndarray)
            rho_dev = to_dev(np.ascontiguousarray(rho, dtype=np.float64)) if het else None
            scale = self.model.thickness if self.model.dim == 2 else 1.0
            _lib.call("efea_uv_ke", dev.code, 64, dev.Ne, ptr(dev.connect), ptr(dev.coord), ptr(tab), nPg,
                      ptr(rho_dev), 1 if het else 0, 0.0 if het else float(rho), self.model.dim, scale, ptr(Me),
                      stream_ptr())
            M.scatter(Me)
            self._M = M
        return self._M

    def _time_bases(self, problemType) -> dict:
        return {"K": self._device_K(), "M": self._device_M()}

    def _time_damping(self, problemType) -> dict:
        cM, cK = self._rayleigh
        return {n: c for n, c in (("K", cK), ("M", cM)) if c != 0.0}

    def Get_K_C_M_F(self, problemType=None):
        """Fresh scipy copies ``(K, C, M, F)`` with the mass matrix ``M``, the
        Rayleigh damping ``C = coefK K + coefM M`` and ``F`` the assembled
        load (none for this model) (``_simu.py:161-175``)."""
        K = self._device_K().to_scipy()
        M = self._device_M().to_scipy()
        cM, cK = self._rayleigh
        C = (cK * K + cM * M).tocsr()
        return K, C, M, sp.csr_matrix((K.shape[0], 1))

    @property
    def displacement(self) -> np.ndarray:
        return self._u.cpu().numpy()

    def Get_x0(self, problemType=None) -> np.ndarray:
        return self.displacement

    @property
    def speed(self) -> np.ndarray:
        return self._v.cpu().numpy()

    @property
    def accel(self) -> np.ndarray:
        return self._a.cpu().numpy()

    def _solve_gen(self):
        if self.algo == AlgoType.elliptic:
            A = self._device_K()
            b = self._neumann_device("elastic")
            x = self._u.clone()
            yield from self._solve_system_gen(A, b, x, "elastic")
            self._u = x
        else:
            self._u, self._v, self._a = yield from self._time_step_gen("elastic", self._u, self._v, self._a)

    def Solve(self) -> np.ndarray:
        drive(self._solve_gen())
        return self.displacement

    def Calc_Reaction(self, dofs=None, problemType=None) -> np.ndarray:
        A = self._device_K()
        y = A.spmv(self._u)
        dofs = self.Get_dofs() if dofs is None else np.asarray(dofs)
        return y.cpu().numpy()[dofs]

    def _Calc_Psi_Elas(self) -> float:
        A = self._device_K()
        y = A.spmv(self._u)
        return 0.5 * float(torch.dot(self._u, y))

    def Save_Iter(self, iter: Optional[dict] = None) -> None:
        """Displacement, plus speed and acceleration under a hyperbolic scheme
        (``Simulations/_elastic.py:164-176``)."""
        it = {} if iter is None else dict(iter)
        it["displacement"] = self.displacement
        if self.algo in AlgoType.Get_Hyperbolic_Types():
            it["speed"] = self.speed
            it["accel"] = self.accel
        super().Save_Iter(it)

    def Set_Iter(self, iter: int = -1) -> Optional[dict]:
        r = super().Set_Iter(iter)
        if r is not None:
            self._u = to_dev(r["displacement"])
            if self.algo in AlgoType.Get_Hyperbolic_Types() and "speed" in r and "accel" in r:
                self._v, self._a = to_dev(r["speed"]), to_dev(r["accel"])
            else:
                self._v, self._a = torch.zeros_like(self._u), torch.zeros_like(self._u)
        return r

    # ------------------------------------------------------------ results
    def Results_Available(self) -> list:
        """``Elastic.Results_Available`` (``Simulations/_elastic.py:203-229``;
        ZZ1 is not provided).  Speed and acceleration are zero until a
        hyperbolic scheme has stepped."""
        dim = self.model.dim
        out = ["displacement", "displacement_norm", "displacement_matrix", "speed", "speed_norm", "accel",
               "accel_norm"]
        comps = ["xx", "yy", "xy"] if dim == 2 else ["xx", "yy", "zz", "yz", "xz", "xy"]
        axes = ["x", "y"] if dim == 2 else ["x", "y", "z"]
        out += ["u" + a for a in axes] + ["v" + a for a in axes] + ["a" + a for a in axes]
        out += ["S" + c for c in comps] + ["E" + c for c in comps]
        out += ["Svm", "Stress", "Evm", "Strain", "Wdef", "Wdef_e", "ZZ1", "ZZ1_e"]
        return out

    def _Calc_ZZ1(self):
        """ZZ1 error estimate (``Simulations/_elastic.py:398-418``): the
        deformation energy with the stresses smoothed at the nodes against
        the element energy, ``(|W - Ws| / W, |Ws_e - W_e| / W)``, on the
        device (``efea_elastic_strain_stress``, ``efea_node_average``,
        ``efea_zz1_energy``)."""
        dev = self.dev
        ns = 3 if self.model.dim == 2 else 6
        e = self._strain_stress_e()
        W_e = e[:, -1]
        sig_n = self.dev.node_values(e[:, ns + 1:2 * ns + 1].contiguous()).contiguous()
        Ws_e = torch.empty(dev.Ne, dtype=torch.float64, device="cuda")
        tab, nPg = dev.tables(MatrixType.rigi)
        _lib.call("efea_zz1_energy", dev.code, dev.Ne, ptr(dev.connect), ptr(dev.coord), ptr(tab), nPg, ptr(self._u),
                  ptr(sig_n), self.model.thickness, ptr(Ws_e), stream_ptr())
        W = W_e.sum()
        return float(torch.abs(W - Ws_e.sum()) / W), torch.abs(Ws_e - W_e) / W

    def _strain_stress_e(self) -> torch.Tensor:
        """(Ne, 2 ns + 3) per element on the device: strain components, Evm,
        stress components, Svm, W_e (``efea_elastic_strain_stress``)."""
        dev = self.dev
        if isinstance(self.model, Anisotropic):
            raise NotImplementedError("strain / stress results of anisotropic laws are not on the device")
        ns = 3 if self.model.dim == 2 else 6
        out = torch.empty((dev.Ne, 2 * ns + 3), dtype=torch.float64, device="cuda")
        tab, nPg = dev.tables(MatrixType.rigi)
        E = to_dev(self.model._E_array())
        _lib.call("efea_elastic_strain_stress", dev.code, dev.Ne, ptr(dev.connect), ptr(dev.coord), ptr(tab), nPg,
                  ptr(self._u), ptr(E), 1 if self.model.isHeterogeneous else 0, self.model.v,
                  int(self.model.planeStress), self.model.thickness, ptr(out), stream_ptr())
        return out

    def Result(self, result: str, nodeValues: bool = True, iter: Optional[int] = None):
        """``Elastic.Result`` (``Simulations/_elastic.py:231-321``): node (or,
        with ``nodeValues=False``, element) values of ``result``; ``Wdef`` is
        a float.  Strain and stress results come from the device kernel."""
        if iter is not None:
            self.Set_Iter(iter)
        if result not in self.Results_Available():
            print(f"\n The result '{result}' is not available.\n Available results: {self.Results_Available()}")
            return None
        dim, Nn = self.model.dim, self.mesh.Nn
        axes = {"x": 0, "y": 1, "z": 2}
        if result in ("ux", "uy", "uz"):
            values = self.displacement.reshape(Nn, -1)[:, axes[result[1]]]
        elif result == "displacement":
            values = self.displacement
        elif result == "displacement_norm":
            values = np.linalg.norm(self.displacement.reshape(Nn, -1), axis=1)
        elif result == "displacement_matrix":
            values = np.zeros((Nn, 3))
            values[:, :dim] = self.displacement.reshape(Nn, -1)
        elif result[0] in "va" and result[1:] in axes:
            values = (self.speed if result[0] == "v" else self.accel).reshape(Nn, -1)[:, axes[result[1]]]
        elif result in ("speed", "accel"):
            values = self.speed if result == "speed" else self.accel
        elif result in ("speed_norm", "accel_norm"):
            values = np.linalg.norm((self.speed if result == "speed_norm" else self.accel).reshape(Nn, -1), axis=1)
        elif result == "Wdef":
            return float(self._strain_stress_e()[:, -1].sum())
        elif result == "ZZ1":
            return self._Calc_ZZ1()[0]
        elif result == "ZZ1_e":
            values = self._Calc_ZZ1()[1]
        else:
            ns = 3 if dim == 2 else 6
            e = self._strain_stress_e()
            comps = ["xx", "yy", "xy"] if dim == 2 else ["xx", "yy", "zz", "yz", "xz", "xy"]
            if result == "Wdef_e":
                values = e[:, -1]
            elif result == "Strain":
                values = e[:, :ns]
            elif result == "Stress":
                values = e[:, ns + 1:2 * ns + 1]
            elif result == "Evm":
                values = e[:, ns]
            elif result == "Svm":
                values = e[:, 2 * ns + 1]
            elif result[0] == "E":
                values = e[:, comps.index(result[1:])]
            else:
                values = e[:, ns + 1 + comps.index(result[1:])]
            values = values.contiguous()
        return self._reshape_values(values, nodeValues)


class Thermal(_Simu):
    """Heat transfer (``Simulations/_thermal.py``): the temperature ``t``, one
    DOF per node, steady (elliptic scheme) or transient (parabolic scheme).

    ``K_e = sum_rigi k wJ dN^T dN`` and ``C_e = sum_mass rho c wJ N^T N``
    (``_thermal.py:114-137``).  The reference scales both by the thickness
    only when the simulation's dimension -- the model's, 1 for a thermal
    model (``_simu.py:621``) -- is 2, so its element matrices carry no
    thickness, while line and surface loads on 2D meshes do
    (``_simu.py:2572-2630``); both are mirrored."""

    def __init__(self, mesh: Mesh, model: ThermalModel, folder: str = "", verbosity=False):
        assert isinstance(model, ThermalModel), "model must be a thermal model"
        self._needUpdate = True
        super().__init__(mesh, model, folder, verbosity)
        self._rho = 1.0
        self._K = None
        self._C = None
        self._u = torch.zeros(mesh.Nn, dtype=torch.float64, device="cuda")
        self._v = torch.zeros(mesh.Nn, dtype=torch.float64, device="cuda")
        self._a = torch.zeros(mesh.Nn, dtype=torch.float64, device="cuda")

    def Get_problemTypes(self) -> list:
        return ["thermal"]

    def Get_unknowns(self, problemType=None) -> list:
        return ["t"]

    def Get_dof_n(self, problemType=None) -> int:
        return 1

    @property
    def thermalModel(self) -> ThermalModel:
        return self.model

    @property
    def rho(self):
        """Density, a scalar or one value per element (``_simu.py:706``)."""
        return self._rho

    @rho.setter
    def rho(self, value):
        arr = np.asarray(value, dtype=float)
        assert np.all(arr >= 0.0), "rho must be >= 0"
        self._rho = float(arr) if arr.ndim == 0 else arr
        self.Need_Update()

    @property
    def needUpdate(self) -> bool:
        return self._needUpdate

    def Need_Update(self, value: bool = True) -> None:
        self._needUpdate = value

    def Assembly(self):
        """Conduction ``K`` (``efea_reaction_diffusion_ke`` without reaction)
        and capacity ``C`` (``efea_uv_ke``) on the device, scattered."""
        dev, m = self.dev, self.thermalModel
        K = dev.matrix("K", 1)
        C = dev.matrix("C", 1)
        tab_m, nPgm = dev.tables(MatrixType.mass)
        tab_r, nPgr = dev.tables(MatrixType.rigi)
        k_het = isinstance(m.k, np.ndarray)
        k_dev = to_dev(np.ascontiguousarray(np.atleast_1d(m.k), dtype=np.float64))
        Ke = K.new_entries()
        _lib.call("efea_reaction_diffusion_ke", dev.code, 64, dev.Ne, ptr(dev.connect), ptr(dev.coord), ptr(tab_m),
                  nPgm, ptr(tab_r), nPgr, None, 0.0, ptr(k_dev), 1 if k_het else 0, None, 1.0, ptr(Ke), None,
                  stream_ptr())
        K.scatter(Ke)
        coef = self._rho * m.c
        c_het = isinstance(coef, np.ndarray)
        c_dev = to_dev(np.ascontiguousarray(coef, dtype=np.float64)) if c_het else None
        Ce = C.new_entries()
        _lib.call("efea_uv_ke", dev.code, 64, dev.Ne, ptr(dev.connect), ptr(dev.coord), ptr(tab_m), nPgm, ptr(c_dev),
                  1 if c_het else 0, 0.0 if c_het else float(coef), 1, 1.0, ptr(Ce), stream_ptr())
        C.scatter(Ce)
        self._K, self._C = K, C
        self._needUpdate = False
        return K

    def _device_KC(self):
        if self._needUpdate or self._K is None:
            self.Assembly()
        return self._K, self._C

    def _time_bases(self, problemType) -> dict:
        K, C = self._device_KC()
        return {"K": K, "C": C}

    def Get_K_C_M_F(self, problemType=None):
        """Fresh scipy copies ``(K, C, M, F)``: conduction, capacity, ``M = 0``
        and no assembled load (``_simu.py:161-175``)."""
        K, C = self._device_KC()
        K, C = K.to_scipy(), C.to_scipy()
        return K, C, sp.csr_matrix(K.shape), sp.csr_matrix((K.shape[0], 1))

    @property
    def thermal(self) -> np.ndarray:
        return self._u.cpu().numpy()

    @property
    def thermalDot(self) -> np.ndarray:
        return self._v.cpu().numpy()

    def Get_x0(self, problemType=None) -> np.ndarray:
        return self.thermal

    def _solve_gen(self):
        if self.algo == AlgoType.elliptic:
            K, _ = self._device_KC()
            x = self._u.clone()
            yield from self._solve_system_gen(K, self._neumann_device("thermal"), x, "thermal")
            self._u = x
        else:
            self._u, self._v, self._a = yield from self._time_step_gen("thermal", self._u, self._v, self._a)

    def Solve(self) -> np.ndarray:
        drive(self._solve_gen())
        return self.thermal

    def Calc_Reaction(self, dofs=None, problemType=None) -> np.ndarray:
        """``K t`` at ``dofs`` (all by default)."""
        K, _ = self._device_KC()
        y = K.spmv(self._u)
        dofs = self.Get_dofs() if dofs is None else np.asarray(dofs)
        return y.cpu().numpy()[dofs]

    def Save_Iter(self, iter: Optional[dict] = None) -> None:
        """Temperature, plus its rate under the parabolic scheme (``_thermal.py:139-150``)."""
        it = {} if iter is None else dict(iter)
        it["thermal"] = self.thermal
        if self.algo == AlgoType.parabolic:
            it["thermalDot"] = self.thermalDot
        super().Save_Iter(it)

    def Set_Iter(self, iter: int = -1) -> Optional[dict]:
        r = super().Set_Iter(iter)
        if r is not None:
            self._u = to_dev(r["thermal"])
            if self.algo == AlgoType.parabolic and "thermalDot" in r:
                self._v = to_dev(r["thermalDot"])
            else:
                self._v = torch.zeros_like(self._u)
        return r

    def Results_Available(self) -> list:
        return ["thermal", "thermalDot", "displacement_matrix"]

    def Result(self, result: str, nodeValues: bool = True, iter: Optional[int] = None):
        """``Thermal.Result`` (``_thermal.py:171-200``)."""
        if iter is not None:
            self.Set_Iter(iter)
        if result not in self.Results_Available():
            print(f"\n The result '{result}' is not available.\n Available results: {self.Results_Available()}")
            return None
        if result == "thermal":
            values = self.thermal
        elif result == "thermalDot":
            values = self.thermalDot
        else:
            values = np.zeros((self.mesh.Nn, 3))
        return self._reshape_values(values, nodeValues)


# damage from which a node's elements count as broken for a stalled displacement solve
ISOLATE_DAMAGE = 0.99


class PhaseField(_Simu):
    """Staggered phase-field fracture simulation (``Simulations/_phasefield.py``)."""

    class ProblemTypes:
        damage = "damage"
        elastic = "elastic"

    def __init__(self, mesh: Mesh, model: PhaseFieldModel, folder: str = "", verbosity=False):
        assert isinstance(model, PhaseFieldModel), "model must be a phase field model"
        assert mesh.elemType in (ElemType.TRI3, ElemType.QUAD4, ElemType.TETRA4, ElemType.TETRA10, ElemType.HEXA8), \
            "the device phase-field path supports TRI3 / QUAD4 / TETRA4 / TETRA10 / HEXA8 meshes"
        assert mesh.dim == model.dim, "mesh and material dimensions differ"
        self._updU = False
        self._updD = False
        super().__init__(mesh, model, folder, verbosity)
        model.material._Add_observer(self)
        dim = model.dim
        Nn, Ne = mesh.Nn, mesh.Ne
        self._u = torch.zeros(Nn * dim, dtype=torch.float64, device="cuda")
        self._d = torch.zeros(Nn, dtype=torch.float64, device="cuda")
        nPgm = gauss_points(mesh.elemType, MatrixType.mass)[1].size
        self._psiP = torch.zeros(Ne * nPgm, dtype=torch.float64, device="cuda")
        self._H = None  # history (old psi+), None until the first Save_Iter
        self._H0 = None  # zeros: the history before the first Save_Iter
        self._Ku = self._Kd = None
        self._Fd = torch.zeros(Nn, dtype=torch.float64, device="cuda")
        self.Niter_staggered = 0
        self.convIter = 0.0
        # The single-system path queues a damage solve with an expected
        # iteration count (the pipelined block-Jacobi PCG) without waiting for
        # it, queues the displacement assembly and its solve's set-up behind
        # it, and reads the damage solve's end just before the displacement
        # solve; a damage solve that needed more iterations than it queued is
        # continued, and the displacement system assembled and set up again.
        # The displacement solve likewise, with the damage assembly and the
        # energy of the convergence measure (convOption 1 / 2) queued behind
        # it.  Same iterates either way.
        self.defer_damage = True
        self.defer_displacement = True
        # BoundConstrain solve: CGLS reduction per pass, first-order tolerance
        # (damage units), CGLS iterations per pass, active-set updates
        self.bounded_tol = 1e-10
        self.bounded_kkt_tol = 1e-9
        self.bounded_maxit = 100000
        # active-set updates: None = max(200, number of damage DOFs) (a primal-dual
        # update frees or binds a few variables, so the count grows with the width
        # of the damage band in nodes)
        self.bounded_max_outer = None

    def Get_problemTypes(self) -> list:
        return ["damage", "elastic"]

    @property
    def problemType(self) -> str:
        return "elastic"

    def Get_unknowns(self, problemType=None) -> list:
        if problemType == "damage":
            return ["d"]
        return ["x", "y"] if self.model.dim == 2 else ["x", "y", "z"]

    def Get_dof_n(self, problemType=None) -> int:
        return 1 if problemType == "damage" else self.model.dim

    @property
    def phaseFieldModel(self) -> PhaseFieldModel:
        return self.model

    @property
    def needUpdate(self) -> bool:
        return not self._updD or not self._updU

    def Need_Update(self, value: bool = True) -> None:
        self._updD = not value
        self._updU = not value
        if value:
            self.__dict__.pop("_dev_params", None)

    @property
    def displacement(self) -> np.ndarray:
        return self._u.cpu().numpy()

    @property
    def damage(self) -> np.ndarray:
        return self._d.cpu().numpy()

    def _material_args(self):
        mat = self.model.material
        c = self.__dict__.setdefault("_dev_params", {})
        if "E" not in c:
            c["E"] = to_dev(mat._E_array())
        return c["E"], (1 if mat.isHeterogeneous else 0), mat.v, int(mat.planeStress)

    def _history_arg(self) -> Optional[torch.Tensor]:
        """The old psi+ the damage kernel clamps against: with the History
        solver the saved field, zeros before the first Save_Iter (the
        reference starts from zeros, ``_phasefield.py:512-529``; a split whose
        psi+ can be negative, e.g. AnisotStrain, sees the difference); no
        clamp (None) for the other solvers."""
        if self.model.solver != PhaseFieldModel.SolverType.History:
            return None
        if self._H is not None:
            return self._H
        if self._H0 is None or self._H0.numel() != self._psiP.numel():
            self._H0 = torch.zeros_like(self._psiP)
        return self._H0

    def _split_code(self) -> int:
        """``EFEA_SPLIT_*`` of the model's split (all fourteen in 2D)."""
        split = str(self.model.split)
        code = PhaseFieldModel.SPLIT_CODE.get(split)
        if code is None:
            raise NotImplementedError(f"split {split}. Device splits: {list(PhaseFieldModel.SPLIT_CODE)}")
        if self.model.dim == 3 and split not in ("Amor", "Bourdin"):
            raise NotImplementedError(f"split {split} in 3D: the device kernels provide Amor and Bourdin in 3D")
        return code

    def _Gc_dev(self):
        c = self.__dict__.setdefault("_dev_params", {})
        if "Gc" not in c:
            c["Gc"] = to_dev(self.model._Gc_array())
        return c["Gc"]

    # --------------------------------------------------------- assembly
    def _assemble_damage(self) -> None:
        dev, pfm = self.dev, self.model
        Kd = dev.csr_map(1)
        Ke = Kd.new_entries()
        Fe = torch.empty(dev.Ne * dev.nPe, dtype=torch.float64, device="cuda")
        tm, nm = dev.tables(MatrixType.mass)
        tr, nr = dev.tables(MatrixType.rigi)
        E, Es, nu, ps = self._material_args()
        Gc = self._Gc_dev()
        _lib.call("efea_pf_damage_ke", dev.code, dev.Ne, ptr(dev.connect), ptr(dev.coord), ptr(tm), nm, ptr(tr), nr,
                  ptr(self._u), ptr(self._history_arg()), ptr(E), Es, nu, ps, ptr(Gc), 1 if pfm.isHeterogeneous else 0,
                  pfm.l0, self._split_code(), 0 if pfm.regularization == "AT1" else 1,
                  pfm.thickness, ptr(self._psiP), ptr(Ke), ptr(Fe), stream_ptr())
        Kd.scatter(Ke)
        dev.vec_map(1).assemble(Fe, self._Fd)
        self._Kd = Kd
        self._updD = True

    def _assemble_elastic(self) -> None:
        dev, pfm = self.dev, self.model
        Ku = dev.csr_map(pfm.dim)
        Ke = Ku.new_entries()
        tr, nr = dev.tables(MatrixType.rigi)
        E, Es, nu, ps = self._material_args()
        _lib.call("efea_pf_elastic_ke", dev.code, dev.Ne, ptr(dev.connect), ptr(dev.coord), ptr(tr), nr,
                  ptr(self._u), ptr(self._d), ptr(E), Es, nu, ps, self._split_code(), pfm.thickness,
                  ptr(Ke), stream_ptr())
        Ku.scatter(Ke)
        self._Ku = Ku
        self._updU = True

    def _device_K(self, problemType):
        """``Get_K_C_M_F`` caching (``_phasefield.py:248-269``)."""
        if problemType == "elastic":
            if not self._updU:
                self._assemble_elastic()
            return self._Ku
        if not self._updD:
            self._assemble_damage()
        return self._Kd

    def Get_K_C_M_F(self, problemType=None):
        problemType = problemType or "elastic"
        A = self._device_K(problemType)
        K = A.to_scipy()
        z = sp.csr_matrix(K.shape)
        if problemType == "elastic":
            return K, z, z.copy(), sp.csr_matrix((K.shape[0], 1))
        return K, z, z.copy(), self.dev.vec_map(1).to_scipy(self._Fd)

    # --------------------------------------------------------- energies
    def _quad(self, A, x) -> float:
        y = A.spmv(x)
        return 0.5 * float(torch.dot(x, y))

    def _Calc_Psi_Crack(self) -> float:
        Kd = self._device_K("damage")
        if float(torch.linalg.vector_norm(self._d)) == 0:
            return 0.0
        return self._quad(Kd, self._d)

    def _Calc_Psi_Elas(self) -> float:
        Ku = self._device_K("elastic")
        if float(torch.linalg.vector_norm(self._u)) == 0:
            return 0.0
        return self._quad(Ku, self._u)

    def _Calc_Psi_Ext(self, fu: torch.Tensor) -> float:
        if float(torch.linalg.vector_norm(self._u)) == 0 or float(torch.linalg.vector_norm(fu)) == 0:
            return 0.0
        return float(torch.dot(self._u, fu))

    # device scalars of the same energies (no host wait): with zero d or u the
    # products are exactly 0, the value the early returns above give
    def _psi_crack_dev(self, d: Optional[torch.Tensor] = None) -> torch.Tensor:
        Kd = self._device_K("damage")
        d = self._d if d is None else d
        return 0.5 * torch.dot(d, Kd.spmv(d))

    def _psi_elas_dev(self, u: Optional[torch.Tensor] = None) -> torch.Tensor:
        Ku = self._device_K("elastic")
        u = self._u if u is None else u
        return 0.5 * torch.dot(u, Ku.spmv(u))

    def _energy_dev(self, convOption: int, fu: Optional[torch.Tensor], d: Optional[torch.Tensor] = None,
                    u: Optional[torch.Tensor] = None) -> torch.Tensor:
        """The convergence energy of the fields ``d`` and ``u`` (the current
        ones by default) with the current matrices."""
        u = self._u if u is None else u
        E = self._psi_crack_dev(d)
        if convOption == 2:
            E = E + self._psi_elas_dev(u) - torch.dot(u, fu)
        return E

    # ------------------------------------------------------------ solves
    def _start(self, field: str) -> torch.Tensor:
        """Start of the iterative solve of ``field`` ("u" or "d").  The
        damage starts, on the first pass of an increment that follows two
        saved ones (the field is still the last saved one), from the secant
        extrapolation ``2 d_n - d_(n-1)`` of those saves (SENT 200 x 200
        history: 72.2 instead of 76.6 iterations per damage solve); the
        displacement, and the damage otherwise, from the current field.  (A
        secant displacement start measured 197.6 instead of 182.7 iterations:
        the stopping test is relative to max(sum alpha r.z, x0.A x0), and the
        current field with the new Dirichlet values carries the energy of the
        top-edge jump, which the secant start does not.)  Only the Krylov
        iteration's start changes: the solution meets the same stopping test
        (Dirichlet values are written over it)."""
        cur = self._u if field == "u" else self._d
        saved = self.__dict__.get("_saved_states")
        if field == "u" or not self.extrapolate_guess or saved is None:
            return cur.clone()
        prev, last = saved[field]
        if prev is None or last is not cur or prev.shape != cur.shape:
            return cur.clone()
        return 2.0 * cur - prev

    def _solve_damage_gen(self):
        Kd = self._device_K("damage")
        b = self._Fd + self._neumann_device("damage")
        if self.model.solver == PhaseFieldModel.SolverType.BoundConstrain:
            x = self._d.clone()
            self._solve_bounded(Kd, b, x)
        else:
            x = self._start("d")
            yield from self._solve_system_gen(Kd, b, x, "damage")
        self._d = x

    def _solve_bounded(self, Kd, b: torch.Tensor, x: torch.Tensor) -> None:
        """BoundConstrain damage solve: ``min |Kd x - b|^2`` with the current
        damage as lower bound (values >= 1 lowered by one ulp) and 1 as
        upper bound (``_phasefield.py:148-163``; the reference calls
        ``scipy.optimize.lsq_linear``, ``Solvers.py:376-383``), solved to
        its minimiser on the device by ``efea_solve_bounded_lsq``."""
        assert self.Bc_dofs_Dirichlet("damage").size == 0, \
            "BoundConstrain bounds every damage DOF: Dirichlet conditions on the damage are not supported"
        assert Kd.dof_n == 1 and Kd.node_ptr.dtype == torch.int32
        lb = self._d.clone()
        lb[lb >= 1] = 1 - np.finfo(float).eps
        ub = torch.ones_like(lb)
        N = Kd.Nn
        nbytes = int(_lib.load().efea_bounded_lsq_workspace_bytes(N))
        ws = self._ws.get(nbytes)
        info = np.zeros(4)
        rc = _lib.call("efea_solve_bounded_lsq", N, ptr(Kd.node_ptr), ptr(Kd.node_cols), ptr(Kd.values), ptr(b),
                       ptr(lb), ptr(ub), ptr(x), ptr(ws), nbytes, self.bounded_tol, self.bounded_kkt_tol,
                       self.bounded_maxit,
                       self.bounded_max_outer if self.bounded_max_outer is not None else max(200, N),
                       info.ctypes.data, stream_ptr())
        self.last_bounded = {"updates": int(info[0]), "iterations": int(info[1]), "violations": int(info[2])}
        if rc == _lib.ERR_NOCONV:
            raise RuntimeError(f"BoundConstrain damage solve did not converge: {_lib.last_error()}")

    def _solve_elastic_gen(self):
        Ku = self._device_K("elastic")
        b = self._neumann_device("elastic")
        x = self._start("u")
        yield from self._solve_system_gen(Ku, b, x, "elastic")
        self._u = x

    def _settle_damage(self) -> bool:
        """``before_solve`` of the displacement solve: ends a deferred damage
        solve, queued before the displacement assembly and set-up.  If it went
        on past its queued iterations, the assembly is redone from the final
        damage (True: the solve's set-up is redone too)."""
        if self._finish_pending():
            self._assemble_elastic()
            return True
        return False

    def _staggered(self, tolConv=1.0, maxIter=500, convOption=2):
        """The staggered loop as a generator of solve jobs; returns
        ``converged``.  ``Solve`` runs it on the single-system path, the
        serving loop on a batch executor."""
        assert 0 < tolConv <= 1, "tolConv must be between 0 and 1."
        assert maxIter > 1, "Must be > 1."
        Niter, converged, convIter = 0, False, 0.0
        t0 = time.perf_counter()
        self._finish_pending()  # one a failed earlier call left queued
        fu = self._neumann_device("elastic") if convOption == 2 else None
        # the convergence measure stays a device scalar unless the loop's
        # control needs it (tolConv < 1): no host wait between the solves
        while not converged and Niter < maxIter:
            Niter += 1
            d_n, u_n = self._d, self._u
            yield from self._solve_damage_gen()
            if convOption in (1, 2):
                # the energy of the pass's start, queued behind a deferred damage
                # solve: from d_n and u_n with the matrices before the pass (the
                # damage solve wrote a new field; the displacement matrix is
                # assembled again only below, the damage one already was)
                E_n = self._energy_dev(convOption, fu, d_n, u_n)
            self._updU = False
            yield from self._solve_elastic_gen()
            self._updD = False
            if convOption == 3:
                self._finish_pending()  # its measure is read on the host
            d_np1, u_np1 = self._d, self._u
            if convOption == 0:
                convIter = torch.max(torch.abs(d_np1 - d_n))
            elif convOption in (1, 2):
                def measure():
                    E_np1 = self._energy_dev(convOption, fu)
                    dE = torch.abs(E_n - E_np1)
                    return torch.where(E_np1 == 0, dE, dE / torch.abs(E_np1))

                convIter = measure()
                if self._finish_pending():
                    # a deferred displacement solve went on past its queued
                    # iterations: the damage system and the measure again
                    self._updD = False
                    convIter = measure()
            elif convOption == 3:
                diffU = torch.abs(u_np1 - u_n)
                nz = u_np1 != 0
                diffU[nz] = diffU[nz] / torch.abs(u_np1[nz])
                diffD = torch.abs(d_np1 - d_n)
                nz = d_np1 != 0
                diffD[nz] = diffD[nz] / torch.abs(d_np1[nz])
                convU, convD = float(diffU.sum()), float(diffD.sum())
                convIter = max(convD, convU)
            self._finish_pending()  # (convOption 0: its measure reads no displacement)
            if tolConv == 1 or float(d_np1.max()) == 0:
                converged = True
            elif convOption == 3:
                converged = (convD <= tolConv) and (convU <= tolConv * 0.999)
            else:
                converged = float(convIter) <= tolConv
        self.Niter_staggered = Niter
        self._convIter = convIter
        self.timeIter = time.perf_counter() - t0
        return converged

    @property
    def convIter(self) -> float:
        c = self.__dict__.get("_convIter", 0.0)
        return float(c)

    @convIter.setter
    def convIter(self, value: float) -> None:
        self._convIter = value

    def Solve(self, tolConv=1.0, maxIter=500, convOption=2):
        """Staggered scheme (``_phasefield.py:303-432``): returns ``(u, d, converged)``.
        With the ``HistoryDamage`` solver the returned ``d`` is the larger of
        the damage before the call and after it, node by node; the
        simulation's own damage field is the solved one (``:338, 412-416``)."""
        history_damage = self.model.solver == PhaseFieldModel.SolverType.HistoryDamage
        d_old = self.damage if history_damage else None
        converged = drive(self._staggered(tolConv, maxIter, convOption))
        # u, d and a device convergence measure in one wait
        c = self.__dict__.get("_convIter", 0.0)
        if isinstance(c, torch.Tensor):
            u, d, c = to_host(self._u, self._d, c)
            self._convIter = float(c)
        else:
            u, d = to_host(self._u, self._d)
        if history_damage:
            d = np.maximum(d_old, d)
        return u, d, converged

    def Save_Iter(self, iter: Optional[dict] = None) -> None:
        """Advances the history field (``_phasefield.py:613-630``)."""
        it = {} if iter is None else dict(iter)
        it["Niter"] = self.Niter_staggered
        it["convIter"] = self.convIter
        it["timeIter"] = getattr(self, "timeIter", 0.0)
        self._H = self._psiP.clone()
        it["displacement"] = self._u.clone()
        it["damage"] = self._d.clone()
        prev = self.__dict__.get("_saved_states")
        # the last two saved fields (tensor identity marks an unchanged field)
        self._saved_states = {"u": (prev["u"][1] if prev else None, self._u),
                              "d": (prev["d"][1] if prev else None, self._d)}
        super().Save_Iter(it)

    def Set_Iter(self, iter: int = -1, resetAll=False) -> Optional[dict]:
        r = super().Set_Iter(iter)
        if r is None:
            return None
        self._u = r["displacement"].clone()
        self._d = r["damage"].clone()
        self._updD = self._updU = False
        self._saved_states = None
        if resetAll:
            self._H = None
            self._assemble_damage()
            self._H = self._psiP.clone()
        return r

    def Calc_Reaction(self, dofs=None, problemType="elastic") -> np.ndarray:
        A = self._device_K(problemType)
        x = self._u if problemType == "elastic" else self._d
        y = A.spmv(x)
        if dofs is None:
            return y.cpu().numpy()
        return y[self._dofs_dev(dofs)].cpu().numpy()

    def _dofs_dev(self, dofs) -> torch.Tensor:
        """``dofs`` as a device index, kept while the same DOFs are asked for
        (a load history reads the reaction at the same DOFs every increment):
        no blocking host-to-device copy per call."""
        dofs = np.asarray(dofs, dtype=np.int64)
        key = dofs.tobytes()
        c = self.__dict__.get("_dofs_index")
        if c is None or c[0] != key:
            c = self._dofs_index = (key, to_dev(dofs))
        return c[1]

    def _solve_method(self, problemType) -> str:
        """Zhang and the ``_PM`` / ``_MP`` splits make the displacement
        stiffness ``g c+ + c-`` non-symmetric (``_phasefield.py:554, 634-676``;
        EasyFEA solves it directly): BiCGStab with block Jacobi."""
        if problemType == "elastic" and str(self.model.split) in PhaseFieldModel.NONSYMMETRIC:
            return "bicgstab"
        return "pcg"

    def _defer(self, problemType) -> bool:
        """The damage solve, with ``defer_damage``: the displacement
        assembly and its solve's set-up are queued while it runs; the
        displacement solve, with ``defer_displacement``: the damage assembly
        and energy of the convergence measure are queued while it runs."""
        if problemType == "damage":
            return self.defer_damage
        return problemType == "elastic" and self.defer_displacement

    def _before_solve(self, problemType):
        return self._settle_damage if problemType == "elastic" else None

    def _isolation(self, problemType):
        """Fully damaged nodes (d >= ISOLATE_DAMAGE), most damaged first: a
        displacement solve that stalls on a broken band is redone with them as
        aggregates of their own (``SolveJob._isolated_retry``)."""
        if problemType != "elastic":
            return None

        def nodes():
            d = self._d
            sel = torch.nonzero(d >= ISOLATE_DAMAGE).flatten()
            return sel[torch.argsort(d[sel], descending=True, stable=True)].cpu().numpy()

        return nodes

    def Detect_Damage(self, nodes, threshold: float = 1.0) -> bool:
        nodes = np.asarray(nodes, dtype=np.int64)
        if nodes.size == 0:
            return False
        return bool(torch.any(self._d[to_dev(nodes)] >= threshold))

    def _comps(self) -> list:
        return ["xx", "yy", "xy"] if self.model.dim == 2 else ["xx", "yy", "zz", "yz", "xz", "xy"]

    def Results_Available(self) -> list:
        """``PhaseField.Results_Available`` (``Simulations/_phasefield.py:659-680``)."""
        axes = ["x", "y"] if self.model.dim == 2 else ["x", "y", "z"]
        return (["displacement", "displacement_norm", "displacement_matrix"] + ["u" + a for a in axes]
                + ["S" + c for c in self._comps()] + ["E" + c for c in self._comps()]
                + ["Svm", "Stress", "Evm", "Strain", "damage", "psiP", "Psi_Crack", "Wdef"])

    def _strain_stress_e(self) -> torch.Tensor:
        """(Ne, 2 ns + 3) per element on the device: strain components, Evm,
        damaged stress components, Svm, W_e (``efea_pf_strain_stress``)."""
        dev, pfm = self.dev, self.model
        ns = len(self._comps())
        out = torch.empty((dev.Ne, 2 * ns + 3), dtype=torch.float64, device="cuda")
        tr, nr = dev.tables(MatrixType.rigi)
        E, Es, nu, ps = self._material_args()
        _lib.call("efea_pf_strain_stress", dev.code, dev.Ne, ptr(dev.connect), ptr(dev.coord), ptr(tr), nr,
                  ptr(self._u), ptr(self._d), ptr(E), Es, nu, ps, self._split_code(), pfm.thickness,
                  ptr(out), stream_ptr())
        return out

    def _psiP_now(self) -> torch.Tensor:
        """psi+ at the mass points from the current u with the history
        (``__Calc_psiPlus_e_pg``, ``_phasefield.py:494-538``), (Ne, nPg)."""
        dev, pfm = self.dev, self.model
        tm, nm = dev.tables(MatrixType.mass)
        tr, nr = dev.tables(MatrixType.rigi)
        E, Es, nu, ps = self._material_args()
        psiP = torch.empty(dev.Ne * nm, dtype=torch.float64, device="cuda")
        Ke = torch.empty(dev.Ne * dev.nPe * dev.nPe, dtype=torch.float64, device="cuda")
        Fe = torch.empty(dev.Ne * dev.nPe, dtype=torch.float64, device="cuda")
        _lib.call("efea_pf_damage_ke", dev.code, dev.Ne, ptr(dev.connect), ptr(dev.coord), ptr(tm), nm, ptr(tr), nr,
                  ptr(self._u), ptr(self._history_arg()), ptr(E), Es, nu, ps, ptr(self._Gc_dev()),
                  1 if pfm.isHeterogeneous else 0, pfm.l0, self._split_code(),
                  0 if pfm.regularization == "AT1" else 1, pfm.thickness, ptr(psiP), ptr(Ke), ptr(Fe), stream_ptr())
        return psiP.reshape(dev.Ne, nm)

    def Result(self, result: str, nodeValues: bool = True, iter: Optional[int] = None):
        """``PhaseField.Result`` (``Simulations/_phasefield.py:682-760``)."""
        if iter is not None:
            self.Set_Iter(iter)
        if result not in self.Results_Available():
            print(f"\n The result '{result}' is not available.\n Available results: {self.Results_Available()}")
            return None
        Nn = self.mesh.Nn
        if result == "Wdef":
            return float(self._psi_elas_dev())
        if result == "Psi_Crack":
            return float(self._psi_crack_dev())
        if result == "psiP":
            values = self._psiP_now().mean(1).contiguous()
        elif result == "damage":
            values = self.damage
        elif result in ("ux", "uy", "uz"):
            values = self.displacement.reshape(Nn, -1)[:, "xyz".index(result[1])]
        elif result == "displacement":
            values = self.displacement
        elif result == "displacement_norm":
            values = np.linalg.norm(self.displacement.reshape(Nn, -1), axis=1)
        elif result == "displacement_matrix":
            values = np.zeros((Nn, 3))
            values[:, :self.model.dim] = self.displacement.reshape(Nn, -1)
        else:
            e = self._strain_stress_e()
            comps = self._comps()
            ns = len(comps)
            if result == "Strain":
                values = e[:, :ns]
            elif result == "Stress":
                values = e[:, ns + 1:2 * ns + 1]
            elif result == "Evm":
                values = e[:, ns]
            elif result == "Svm":
                values = e[:, 2 * ns + 1]
            elif result[0] == "E":
                values = e[:, comps.index(result[1:])]
            else:
                values = e[:, ns + 1 + comps.index(result[1:])]
            values = values.contiguous()
        return self._reshape_values(values, nodeValues)


def Load_Simu(folder: str, filename: str = "simulation") -> _Simu:
    """The simulation written by ``_Simu.Save`` (``_simu.py:3333-3366``)."""
    from .sim_io import load_simu

    return load_simu(folder, filename)
