"""Device-resident mesh data, assembled systems and solves.

PyTorch provides device memory and the stream; every computation is a call
into ``libefea_hip.so``.  Nothing here falls back to the CPU: on a machine
without a GPU the constructors raise.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from ._utils import ElemType, MatrixType
from .assembly_map import CsrMap, VecMap, build_csr_map, build_vec_map
from .quadrature import N_pg, dN_pg, gauss_points


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("easyfea_amd needs a ROCm GPU (torch.cuda.is_available() is False); no CPU fallback")
    _lib.load()
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream


def ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


def to_dev(a: np.ndarray, dtype=None) -> torch.Tensor:
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda", non_blocking=False)


def table_block(elemType: ElemType, matrixType: MatrixType) -> np.ndarray:
    """[w (nPg)] [N (nPg*nPe)] [dN (nPg*dim*nPe)] as one FP64 block."""
    _, w = gauss_points(elemType, matrixType)
    return np.concatenate([w, N_pg(elemType, matrixType).ravel(), dN_pg(elemType, matrixType).ravel()])


class DeviceMesh:
    """Connectivity, coordinates and integration tables of the main group."""

    def __init__(self, mesh):
        require_gpu()
        g = mesh.groupElem
        self.mesh = mesh
        self.elemType = g.elemType
        self.code = _lib.ELEM_CODE[g.elemType.value]
        self.Ne = g.Ne
        self.nPe = g.nPe
        self.dim = g.dim
        self.Nn = mesh.Nn
        self.connect_np = g.connect
        self.connect = to_dev(g.connect.astype(np.int32))
        self.coord = to_dev(mesh.coord.astype(np.float64))
        self._tables = {}
        self._maps = {}

    def tables(self, matrixType: MatrixType):
        mt = MatrixType(matrixType)
        if mt not in self._tables:
            self._tables[mt] = (to_dev(table_block(self.elemType, mt)), gauss_points(self.elemType, mt)[1].size)
        return self._tables[mt]

    def csr_map(self, dof_n: int) -> "DeviceCsr":
        key = ("K", dof_n)
        if key not in self._maps:
            self._maps[key] = DeviceCsr(build_csr_map(self.connect_np, self.Nn, dof_n), self.Ne)
        return self._maps[key]

    def vec_map(self, dof_n: int) -> "DeviceVecMap":
        key = ("F", dof_n)
        if key not in self._maps:
            self._maps[key] = DeviceVecMap(build_vec_map(self.connect_np, self.Nn, dof_n))
        return self._maps[key]


class DeviceVecMap:
    def __init__(self, vm: VecMap):
        self.host = vm
        self.seg = to_dev(vm.seg_ptr)
        self.perm = to_dev(vm.perm)
        self.Ndof = vm.Ndof

    def assemble(self, Fe: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        _lib.call("efea_csr_scatter", 64, self.Ndof, ptr(self.seg), ptr(self.perm), ptr(Fe), ptr(out), stream_ptr())
        return out

    def to_scipy(self, F: torch.Tensor) -> sp.csr_matrix:
        """(Ndof, 1) CSR with the entries of every DOF an element touches,
        the pattern ``__Assemble_csr(isMatrix=False)`` produces."""
        vals = F.detach().cpu().numpy()
        used = self.host.used
        indptr = np.zeros(self.Ndof + 1, dtype=np.int32)
        indptr[used + 1] = 1
        indptr = np.cumsum(indptr).astype(np.int32)
        m = sp.csr_matrix((vals[used], np.zeros(used.size, dtype=np.int32), indptr), shape=(self.Ndof, 1))
        m.has_canonical_format = True
        return m


class DeviceCsr:
    """A global matrix on the device: canonical CSR values (FP64, optional
    FP32 copy), the node-adjacency indices the kernels use and the scatter map."""

    def __init__(self, cmap: CsrMap, Ne: int):
        self.host = cmap
        self.Nn = cmap.Nn
        self.dof_n = cmap.dof_n
        self.nnz = cmap.nnz
        self.Ne = Ne
        self.node_ptr = to_dev(cmap.node_ptr)
        self.node_cols = to_dev(cmap.node_cols)
        self.seg = to_dev(cmap.seg_ptr)
        self.perm = to_dev(cmap.perm)
        self.group = _lib.load().efea_default_group(cmap.dof_n, cmap.max_neighbours)
        self.values = torch.empty(self.nnz, dtype=torch.float64, device="cuda")
        self.values32: Optional[torch.Tensor] = None
        self.spmv_scratch = torch.empty(_lib.load().efea_spmv_scratch_bytes(self.Nn, self.dof_n), dtype=torch.uint8,
                                        device="cuda")
        nd = cmap.nPe * cmap.dof_n
        self.n_entries = Ne * nd * nd

    @property
    def Ndof(self) -> int:
        return self.Nn * self.dof_n

    def new_entries(self, dtype=torch.float64) -> torch.Tensor:
        return torch.empty(self.n_entries, dtype=dtype, device="cuda")

    def scatter(self, Ke: torch.Tensor) -> None:
        assert Ke.numel() == self.n_entries and Ke.dtype == torch.float64
        _lib.call("efea_csr_scatter", 64, self.nnz, ptr(self.seg), ptr(self.perm), ptr(Ke), ptr(self.values),
                  stream_ptr())
        self.values32 = None

    def make_fp32(self) -> torch.Tensor:
        if self.values32 is None:
            self.values32 = torch.empty(self.nnz, dtype=torch.float32, device="cuda")
            _lib.call("efea_cast_f64_f32", self.nnz, ptr(self.values), ptr(self.values32), stream_ptr())
        return self.values32

    def spmv(self, x: torch.Tensor, y: Optional[torch.Tensor] = None, known: Optional[torch.Tensor] = None,
             precision: int = 64) -> torch.Tensor:
        if y is None:
            y = torch.empty(self.Ndof, dtype=torch.float64, device="cuda")
        vals = self.values if precision == 64 else self.make_fp32()
        _lib.call("efea_spmv", precision, self.Nn, self.dof_n, self.group, ptr(self.node_ptr), ptr(self.node_cols),
                  ptr(vals), ptr(known), ptr(x), ptr(y), ptr(self.spmv_scratch), stream_ptr())
        return y

    def to_scipy(self) -> sp.csr_matrix:
        h = self.host
        m = sp.csr_matrix((self.values.detach().cpu().numpy(), h.indices, h.indptr), shape=(self.Ndof, self.Ndof))
        m.has_canonical_format = True
        return m


# The FP32 matrix copy pays only when the Krylov SpMV is bandwidth-bound:
# measured step-1 kernel times are equal in FP32 and FP64 for matrices of a
# few MB (latency floor) and 23% apart at 517 MB (64^3 HEXA8).
MIXED_MIN_MATRIX_BYTES = 128 * 2**20


@dataclass
class SolveInfo:
    iterations: int
    rel_residual: float
    replacements: int
    status: int
    mode: str
    fallback: bool = False


class Workspace:
    """Reusable device scratch (grown on demand)."""

    def __init__(self):
        self.buf: Optional[torch.Tensor] = None
        self.small = torch.zeros(1100, dtype=torch.float64, device="cuda")

    def get(self, nbytes: int) -> torch.Tensor:
        if self.buf is None or self.buf.numel() < nbytes:
            self.buf = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        return self.buf


def _solve_once(A: DeviceCsr, b, x, known, m: int, ws: Workspace, tol_residual, tol_anorm, replace_ratio, maxit,
                check_every):
    Minv = torch.empty(A.Nn * A.dof_n * A.dof_n, dtype=torch.float64 if m == 64 else torch.float32,
                       device="cuda")
    _lib.call("efea_bj_setup", m, A.Nn, A.dof_n, ptr(A.node_ptr), ptr(A.node_cols), ptr(A.values), ptr(known),
              ptr(Minv), stream_ptr())
    v32 = A.make_fp32() if m == 32 else None
    nbytes = _lib.load().efea_pcg_workspace_bytes(A.Nn, A.dof_n, A.group)
    buf = ws.get(nbytes)
    info = (C.c_double * 4)()
    rc = _lib.call("efea_solve_pcg", m, A.Nn, A.dof_n, A.group, ptr(A.node_ptr), ptr(A.node_cols), ptr(A.values),
                   ptr(v32), ptr(known), ptr(Minv), ptr(b), ptr(x), ptr(buf), nbytes, tol_residual, tol_anorm,
                   replace_ratio, maxit, check_every, stream_ptr(), C.cast(info, C.c_void_p))
    return rc, SolveInfo(int(info[0]), float(info[1]), int(info[2]), int(info[3]), "fp64" if m == 64 else "mixed")


def solve(A: DeviceCsr, b: torch.Tensor, x: torch.Tensor, known: Optional[torch.Tensor], mode: str,
          ws: Workspace, tol_residual: float = 1e-12, tol_anorm: float = 1e-9, replace_ratio: float = 1e-3,
          maxit: int = 100000, check_every: int = 32, raise_on_fail: bool = True) -> SolveInfo:
    """Solves the Dirichlet-reduced system on the device (see ``efea_solve_pcg``).

    ``x`` carries the Dirichlet values at the known DOFs and the initial guess
    elsewhere; it is overwritten with the solution.  A mixed-precision solve
    that breaks down, diverges or stalls is redone in FP64 from the same
    initial guess (``SolveInfo.fallback``)."""
    assert mode in ("fp64", "mixed")
    m = 64 if mode == "fp64" else 32
    if m == 32 and A.nnz * 8 < MIXED_MIN_MATRIX_BYTES:
        # latency-bound system: the FP32 copy saves no time per iteration
        # (the step kernels sit at their ~5 us floor) and costs iterations
        m = 64
    x0 = x.clone() if m == 32 else None
    rc, si = _solve_once(A, b, x, known, m, ws, tol_residual, tol_anorm, replace_ratio, maxit, check_every)
    if rc != 0 and m == 32:
        x.copy_(x0)
        rc, si = _solve_once(A, b, x, known, 64, ws, tol_residual, tol_anorm, replace_ratio, maxit, check_every)
        si.fallback = True
    if rc != 0 and raise_on_fail:
        raise _lib.EfeaError("efea_solve_pcg", rc, _lib.last_error())
    return si
