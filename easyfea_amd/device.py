"""Device-resident mesh data, assembled systems and solves.

PyTorch provides device memory and the stream; every computation is a call
into ``libefea_hip.so``.  Nothing here falls back to the CPU: on a machine
without a GPU the constructors raise.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from ._utils import ElemType, MatrixType
from .assembly_map import CsrMap, VecMap, build_csr_map, build_vec_map
from .coarse import aggregate, coarse_pairs, default_target, n_modes
from .quadrature import N_pg, dN_pg, gauss_points


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("easyfea_amd needs a ROCm GPU (torch.cuda.is_available() is False); no CPU fallback")
    _lib.load()
    return torch.device("cuda", torch.cuda.current_device())


def _stream_ptr_public() -> int:
    return torch.cuda.current_stream().cuda_stream


def _stream_ptr_raw() -> int:
    # torch's current raw stream of the current device, without the Python
    # stream object (several microseconds per call, paid before every
    # launch the host issues while the GPU waits between solves)
    return torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice())


stream_ptr = _stream_ptr_raw if hasattr(torch._C, "_cuda_getCurrentRawStream") else _stream_ptr_public


def ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


def to_dev(a: np.ndarray, dtype=None) -> torch.Tensor:
    t = torch.from_numpy(np.require(a, requirements=["C", "W"]))  # read-only views (broadcasts) are copied
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda", non_blocking=False)


def to_host(*tensors: torch.Tensor) -> list:
    """Device tensors as host arrays of their own: asynchronous copies into
    fresh pinned buffers (torch's pinned-block cache) and one wait, instead
    of one synchronised pageable copy per tensor."""
    outs = []
    for t in tensors:
        h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
        h.copy_(t, non_blocking=True)
        outs.append(h)
    torch.cuda.current_stream().synchronize()
    return [h.numpy() for h in outs]


def table_block(elemType: ElemType, matrixType: MatrixType) -> np.ndarray:
    """[w (nPg)] [N (nPg*nPe)] [dN (nPg*dim*nPe)] as one FP64 block."""
    _, w = gauss_points(elemType, matrixType)
    return np.concatenate([w, N_pg(elemType, matrixType).ravel(), dN_pg(elemType, matrixType).ravel()])


class _MeshData:
    """Device data that depends on the mesh only (connectivity, coordinates,
    integration tables, sparsity patterns and scatter maps), shared by every
    simulation on the same ``Mesh`` object."""

    def __init__(self, mesh):
        g = mesh.groupElem
        self.connect_np = g.connect
        self.connect = to_dev(g.connect.astype(np.int32))
        self.coord = to_dev(mesh.coord.astype(np.float64))
        self.tables = {}
        self.patterns = {}
        self.vec_maps = {}

    @staticmethod
    def of(mesh) -> "_MeshData":
        data = mesh.__dict__.get("_efea_device")
        if data is None:
            data = _MeshData(mesh)
            mesh.__dict__["_efea_device"] = data
        return data


class DeviceMesh:
    """Connectivity, coordinates and integration tables of the main group,
    plus the global matrices of one simulation."""

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
        self._shared = _MeshData.of(mesh)
        self.connect_np = self._shared.connect_np
        self.connect = self._shared.connect
        self.coord = self._shared.coord
        self._maps = {}

    def tables(self, matrixType: MatrixType):
        mt = MatrixType(matrixType)
        t = self._shared.tables
        if mt not in t:
            t[mt] = (to_dev(table_block(self.elemType, mt)), gauss_points(self.elemType, mt)[1].size)
        return t[mt]

    def node_incidence(self):
        """(node_ptr, node_elems) int32 on the device: the elements holding
        each node, in increasing element order (the node -> element incidence
        of ``_GroupElem.Get_connect_n_e``)."""
        c = self._shared.__dict__.get("incidence")
        if c is None:
            Ne, nPe = self.connect_np.shape
            rows = self.connect_np.ravel().astype(np.int64)
            cols = np.repeat(np.arange(Ne, dtype=np.int64), nPe)
            order = np.lexsort((cols, rows))
            ptr_ = np.zeros(self.Nn + 1, dtype=np.int32)
            ptr_[1:] = np.cumsum(np.bincount(rows, minlength=self.Nn))
            c = self._shared.__dict__["incidence"] = (to_dev(ptr_), to_dev(cols[order].astype(np.int32)))
        return c

    def node_values(self, values_e: torch.Tensor) -> torch.Tensor:
        """Element values (Ne, k) -> node values (Nn, k) (``Mesh.Get_Node_Values``)."""
        v = values_e.reshape(self.Ne, -1).contiguous()
        out = torch.empty((self.Nn, v.shape[1]), dtype=torch.float64, device="cuda")
        n_ptr, n_el = self.node_incidence()
        _lib.call("efea_node_average", self.Nn, v.shape[1], ptr(n_ptr), ptr(n_el), ptr(v), ptr(out), stream_ptr())
        return out

    def pattern(self, dof_n: int) -> "CsrPattern":
        p = self._shared.patterns
        if dof_n not in p:
            p[dof_n] = CsrPattern(build_csr_map(self.connect_np, self.Nn, dof_n), self.Ne, self.mesh.coord,
                                  self.dim)
        return p[dof_n]

    def csr_map(self, dof_n: int) -> "DeviceCsr":
        """This simulation's matrix with ``dof_n`` DOFs per node (values of
        its own over the mesh's shared pattern)."""
        return self.matrix("K", dof_n)

    def matrix(self, name: str, dof_n: int) -> "DeviceCsr":
        """This simulation's matrix ``name`` (e.g. "K", "M", "C", the scheme
        matrix "A") with ``dof_n`` DOFs per node, over the shared pattern."""
        key = (name, dof_n)
        if key not in self._maps:
            self._maps[key] = DeviceCsr(self.pattern(dof_n))
        return self._maps[key]

    def vec_map(self, dof_n: int) -> "DeviceVecMap":
        v = self._shared.vec_maps
        if dof_n not in v:
            v[dof_n] = DeviceVecMap(build_vec_map(self.connect_np, self.Nn, dof_n))
        return v[dof_n]


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


class CsrPattern:
    """Sparsity pattern and scatter map of a global matrix on the device: the
    node adjacency the kernels use, the scatter segments and permutation, and
    the coarse spaces built on it."""

    def __init__(self, cmap: CsrMap, Ne: int, coord: Optional[np.ndarray] = None, dim: Optional[int] = None):
        self.host = cmap
        self.coord = coord
        self.dim = dim
        self._coarse = {}
        self.Nn = cmap.Nn
        self.dof_n = cmap.dof_n
        self.nnz = cmap.nnz
        self.Ne = Ne
        self.node_ptr = to_dev(cmap.node_ptr)
        self.node_cols = to_dev(cmap.node_cols)
        self.seg = to_dev(cmap.seg_ptr)
        self.perm = to_dev(cmap.perm)
        self.group = _lib.load().efea_default_group(cmap.dof_n, cmap.max_neighbours, cmap.Nn)
        nd = cmap.nPe * cmap.dof_n
        self.n_entries = Ne * nd * nd

    def coarse(self, target: Optional[int] = None) -> "CoarseSpace":
        target = default_target(self.dim) if target is None else int(target)
        if target not in self._coarse:
            assert self.coord is not None, "the coarse space needs node coordinates"
            self._coarse[target] = CoarseSpace(self, target)
        return self._coarse[target]


class CoarseSpace:
    """Aggregates, rigid-body modes and the pair lists of ``Ac = P^T A P`` on
    the device (mesh data: built once per pattern and aggregate size)."""

    def __init__(self, pattern: "CsrPattern", target: int, singletons: Optional[np.ndarray] = None):
        h = pattern.host
        self.target = target
        self.dof_n = pattern.dof_n
        self.n_isolated = 0 if singletons is None else int(len(singletons))
        self.ag = aggregate(pattern.coord, h.node_ptr, h.node_cols, target, singletons)
        cp = coarse_pairs(self.ag, h.node_ptr, h.node_cols)
        self.nagg = self.ag.nagg
        self.nm = n_modes(pattern.dim, pattern.dof_n)
        self.ld = _lib.call("efea_coarse_ld", self.nagg, self.dof_n)
        self.agg_ptr = to_dev(self.ag.agg_ptr)
        self.agg_nodes = to_dev(self.ag.agg_nodes)
        self.rel = to_dev(self.ag.rel)
        self.ncb = int(cp.blk_I.size)
        self.pair_ptr = to_dev(cp.pair_ptr)
        self.pair_node = to_dev(cp.pair_node)
        self.pair_pos = to_dev(cp.pair_pos)
        self.blk_I = to_dev(cp.blk_I)
        self.blk_J = to_dev(cp.blk_J)
        self.scratch = torch.empty(_lib.load().efea_coarse_scratch_bytes(self.ld) // 8, dtype=torch.float64,
                                   device="cuda")

    def pipeline(self, pattern: "CsrPattern", group: int) -> "PipeChunks":
        """Node chunks of the pipelined solve's product pass (built once per
        lanes-per-node value)."""
        c = self.__dict__.setdefault("_pipe", {})
        if group not in c:
            c[group] = PipeChunks(self, pattern, group)
        return c[group]

    def setup(self, A: "DeviceCsr", known: Optional[torch.Tensor], shift: float = 0.0) -> torch.Tensor:
        """``Ac^-1`` (ld x ld, FP64) of ``A``'s current FP64 values, the
        diagonal of ``Ac`` scaled by ``1 + shift`` first."""
        Ainv = torch.empty(self.ld * self.ld, dtype=torch.float64, device="cuda")
        _lib.call("efea_coarse_setup", self.dof_n, self.ncb, ptr(self.pair_ptr), ptr(self.pair_node),
                  ptr(self.pair_pos), ptr(self.blk_I), ptr(self.blk_J), ptr(A.node_ptr), ptr(A.node_cols),
                  ptr(A.values), ptr(known), ptr(self.rel), self.ld, shift, ptr(Ainv), ptr(self.scratch),
                  stream_ptr())
        return Ainv


class PipeChunks:
    """Chunks of the pipelined two-level solve (``efea_solve_pcg2p``): runs of
    at most W = 256 / group nodes of one aggregate in aggregate order, so that
    each chunk's partial of ``P^T w`` belongs to one aggregate.  Chunk c owns
    the W positions ``c W + t`` of ``rows`` (node, then its row range; node
    -1 pads a short chunk), so a lane reaches its row without first reading
    the chunk's bounds; ``first`` lists each aggregate's first 256 nodes
    (-1 past its end) for the coarse kernel's one-thread-per-node pass."""

    def __init__(self, cs: "CoarseSpace", pattern: "CsrPattern", group: int):
        ag = cs.ag
        npb = 256 // int(group)
        sizes = np.diff(ag.agg_ptr).astype(np.int64)
        self.maxch = int(max(1, ((sizes + npb - 1) // npb).max()))
        k0, k1, slot = [], [], []
        for a in range(ag.nagg):
            lo, hi = int(ag.agg_ptr[a]), int(ag.agg_ptr[a + 1])
            for j, s in enumerate(range(lo, hi, npb)):
                k0.append(s)
                k1.append(min(s + npb, hi))
                slot.append(a * self.maxch + j)
        self.nchunk = len(k0)
        self.nslot = ag.nagg * self.maxch
        self.width = npb
        nodes = ag.agg_nodes.astype(np.int64)
        node_ptr = pattern.host.node_ptr.astype(np.int64)
        rows = np.zeros((self.nchunk, npb, 3), dtype=np.int32)
        rows[:, :, 0] = -1
        for c, (lo, hi) in enumerate(zip(k0, k1)):
            nd = nodes[lo:hi]
            rows[c, :hi - lo, 0] = nd
            rows[c, :hi - lo, 1] = node_ptr[nd]
            rows[c, :hi - lo, 2] = node_ptr[nd + 1]
        first = np.full((ag.nagg, 256), -1, dtype=np.int32)
        for a in range(ag.nagg):
            lo, hi = int(ag.agg_ptr[a]), int(ag.agg_ptr[a + 1])
            m = min(hi - lo, 256)
            first[a, :m] = nodes[lo:lo + m]
        self.rows = to_dev(rows.ravel())
        self.slot = to_dev(np.asarray(slot, dtype=np.int32))
        self.first = to_dev(first.ravel())


class DeviceCsr:
    """A global matrix on the device: canonical CSR values (FP64, optional
    FP32 copy) over a shared ``CsrPattern``."""

    def __init__(self, pattern: CsrPattern):
        self.pattern = pattern
        self.host = pattern.host
        self.Nn = pattern.Nn
        self.dof_n = pattern.dof_n
        self.nnz = pattern.nnz
        self.Ne = pattern.Ne
        self.node_ptr = pattern.node_ptr
        self.node_cols = pattern.node_cols
        self.seg = pattern.seg
        self.perm = pattern.perm
        self.group = pattern.group
        self.n_entries = pattern.n_entries
        self.dim = pattern.dim
        self.values = torch.empty(self.nnz, dtype=torch.float64, device="cuda")
        self.values32: Optional[torch.Tensor] = None
        self.corr32: Optional[torch.Tensor] = None
        self._spmv_scratch: Optional[torch.Tensor] = None
        self.version = 0  # counts value changes (scatter / combine)

    def coarse(self, target: Optional[int] = None) -> CoarseSpace:
        return self.pattern.coarse(target)

    @property
    def spmv_scratch(self) -> torch.Tensor:
        if self._spmv_scratch is None:
            self._spmv_scratch = torch.empty(_lib.load().efea_spmv_scratch_bytes(self.Nn, self.dof_n),
                                             dtype=torch.uint8, device="cuda")
        return self._spmv_scratch

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
        self.corr32 = None
        self.version += 1

    def combine(self, terms: list) -> None:
        """values = sum of c * X.values over ``terms`` [(c, DeviceCsr)], 1 to 4
        matrices on this pattern (``efea_lincomb``)."""
        assert 1 <= len(terms) <= 4 and all(X.pattern is self.pattern for _, X in terms)
        args = []
        for k in range(4):
            c, X = terms[k] if k < len(terms) else (0.0, None)
            args += [float(c), ptr(X.values) if X is not None else None]
        _lib.call("efea_lincomb", self.nnz, *args, ptr(self.values), stream_ptr())
        self.values32 = None
        self.corr32 = None
        self.version += 1

    def make_fp32(self) -> torch.Tensor:
        """FP32 copy of the values and its row-sum correction (``efea_split_f32``)."""
        if self.values32 is None:
            self.values32 = torch.empty(self.nnz, dtype=torch.float32, device="cuda")
            self.corr32 = torch.empty(self.Nn * self.dof_n * self.dof_n, dtype=torch.float32, device="cuda")
            _lib.call("efea_split_f32", self.Nn, self.dof_n, ptr(self.node_ptr), ptr(self.values),
                      ptr(self.values32), ptr(self.corr32), stream_ptr())
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


def lincomb(out: torch.Tensor, terms: list) -> torch.Tensor:
    """out = sum of c * x over ``terms`` [(c, x)], 1 to 4 FP64 device
    vectors of out's size (``efea_lincomb``; out may be one of the x)."""
    assert 1 <= len(terms) <= 4
    args = []
    for k in range(4):
        c, x = terms[k] if k < len(terms) else (0.0, None)
        if x is not None:
            assert x.dtype == torch.float64 and x.numel() == out.numel() and x.is_contiguous()
        args += [float(c), ptr(x)]
    _lib.call("efea_lincomb", out.numel(), *args, ptr(out), stream_ptr())
    return out


def effective_precision(A: "DeviceCsr", mode: str, adaptive: bool = True) -> int:
    """64 or 32: the value precision of the Krylov products for ``mode``
    ("fp64" | "mixed").  With ``adaptive``, a mixed solve of a matrix below
    ``MIXED_MIN_MATRIX_BYTES`` runs in FP64: the step kernels sit at their
    ~5 us floor there, so the FP32 copy saves no time and costs iterations."""
    assert mode in ("fp64", "mixed"), mode
    if mode == "fp64":
        return 64
    if adaptive and A.nnz * 8 < MIXED_MIN_MATRIX_BYTES:
        return 64
    return 32


def choose_preconditioner(A: "DeviceCsr", precond: str = "auto") -> str:
    """"bj" (nodal block Jacobi) or "two-level" (block Jacobi + rigid-body
    coarse correction).  auto: two-level for vector problems with at least
    four aggregates of the default size, block Jacobi otherwise (scalar
    damage problems converge in tens of iterations with it)."""
    assert precond in ("auto", "bj", "two-level"), precond
    if precond != "auto":
        return precond
    if A.dof_n >= 2 and A.dim is not None and A.Nn >= 4 * default_target(A.dim):
        return "two-level"
    return "bj"


@dataclass
class SolveInfo:
    iterations: int
    rel_residual: float
    replacements: int  # residual replacements; outer steps of a refinement solve
    status: int
    mode: str
    fallback: bool = False
    precond: str = "bj"
    coarse_reused: bool = False
    refine: bool = False  # FP32 inner PCG inside FP64 iterative refinement
    pipelined: bool = False  # pipelined PCG (block Jacobi: one launch per iteration; two-level: two)
    isolated: int = 0  # nodes given aggregates of their own after a stalled solve


# Mixed-precision strategies of a single two-level solve:
#   "values": FP32 matrix values (and block Jacobi) in the Krylov products,
#             FP64 vectors, FP64 residual replacement at every state read-back;
#   "refine": FP64 iterative refinement (r = b - A64 x and x += c in FP64)
#             around an FP32 inner PCG (FP32 vectors, values, block Jacobi and
#             coarse inverse), each inner solve to REFINE_INNER_TOL;
#   "auto":   "refine" for matrices of at least REFINE_MIN_MATRIX_BYTES (the
#             products are bandwidth-bound), "values" below.
# Block-Jacobi solves and the continuous-batching engine use "values" (the
# serving requests ask for it, so that a request alone and in a batch run the
# same algorithm).
REFINE_INNER_TOL = 1e-4
REFINE_MIN_MATRIX_BYTES = 128 * 2**20
# FP64 two-level solves of latency-bound systems (matrix under
# REFINE_MIN_MATRIX_BYTES, at most PIPELINE_MAX_COARSE coarse unknowns) run
# the pipelined PCG (efea_solve_pcg2p): two launches per iteration instead of
# three, at 8 instead of 5 work vectors.
PIPELINE_MAX_COARSE = 4096
# Block-Jacobi solves take the pipelined PCG (efea_solve_pcgp, one launch per
# iteration) under "auto" when the previous solve of their sequence took at
# most this many iterations (the staggered damage solves: about 40-50).  Its
# recurrences drift on long ill-conditioned solves: on the 360 x 39
# cantilever (about 1,200 block-Jacobi iterations) the pipelined solve
# stopped 5e-9 to 6e-8 from the direct solution depending on the
# replacement period, against 3e-10 for the classic iteration.
PIPELINE_BJ_MAX_EXPECTED = 128
# A phase-field displacement solve that stalls with fully damaged nodes is
# redone with those nodes as singleton aggregates; the attempts before stop
# after this many iterations (or 4x the previous solve's), the coarse space
# of the retry holds at most ISOLATE_MAX_COARSE unknowns.
ISOLATE_FIRST_MAXIT = 5000
ISOLATE_MAX_COARSE = 8192
ISOLATE_SHIFTS = (1e-8, 1e-6, 1e-4)


# Lagged coarse inverse.  Any symmetric positive definite Ac^-1 keeps
# B^-1 + P Ac^-1 P^T a valid preconditioner, so a sequence of solves on one
# slowly changing matrix (the staggered phase-field increments) can keep the
# inverse of an earlier Ac: the solution still meets the same stopping test,
# only the iteration count can grow.  The inverse is rebuilt when a solve
# needs COARSE_REUSE_GROWTH times the iterations of the first solve after
# the last rebuild, after COARSE_REUSE_MAX reuses, or when the Dirichlet mask
# changes.
COARSE_REUSE_GROWTH = 1.2
COARSE_REUSE_MAX = 100


@dataclass
class _CoarseCache:
    Ainv: torch.Tensor
    known: Optional[torch.Tensor]
    ref_iterations: Optional[int] = None
    uses: int = 0
    stale: bool = False
    Ainv32: Optional[torch.Tensor] = None


class Workspace:
    """Reusable device scratch (grown on demand)."""

    def __init__(self):
        self.buf: Optional[torch.Tensor] = None

    def get(self, nbytes: int) -> torch.Tensor:
        if self.buf is None or self.buf.numel() < nbytes:
            self.buf = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        return self.buf


class SolveJob:
    """One Dirichlet-reduced solve: ``x`` carries the Dirichlet values at the
    known DOFs and the initial guess elsewhere and receives the solution."""

    def __init__(self, A: "DeviceCsr", b: torch.Tensor, x: torch.Tensor, known: Optional[torch.Tensor],
                 mode: str = "mixed", tol_residual: float = 1e-12, tol_anorm: float = 1e-9,
                 replace_ratio: float = 0.0, maxit: int = 100000, check_every: int = 32, adaptive: bool = True,
                 owner=None, precond: str = "auto", coarse_target: Optional[int] = None, max_blocks: int = 0,
                 expected_iterations: Optional[int] = None, strategy: str = "auto", pipelined="auto"):
        self.A, self.b, self.x, self.known = A, b, x, known
        self.expected_iterations = expected_iterations
        self.precond = choose_preconditioner(A, precond)
        self.max_blocks = int(max_blocks)
        self.coarse_target = coarse_target
        self.cs: Optional[CoarseSpace] = None
        self.Ainv: Optional[torch.Tensor] = None
        self.mode = mode
        self.tol_residual, self.tol_anorm, self.replace_ratio = tol_residual, tol_anorm, replace_ratio
        self.maxit, self.check_every = maxit, check_every
        self.precision = effective_precision(A, mode, adaptive)
        self.owner = owner  # the simulation (lagged coarse inverse cache)
        self.request = None  # the serving request waiting for this solve
        self.fallback = False
        self.x0 = x.clone() if self.precision == 32 else None  # mixed: start of an FP64 redo
        self.Minv: Optional[torch.Tensor] = None
        self.coarse_reused = False
        self._cache: Optional[_CoarseCache] = None
        assert strategy in ("auto", "values", "refine"), strategy
        self.strategy = strategy
        self.refine = False  # decided by run(): single two-level solves only
        self.Ainv32: Optional[torch.Tensor] = None
        assert pipelined in ("auto", True, False), pipelined
        self.pipelined_mode = pipelined
        # "pcg", or "bicgstab" for a non-symmetric matrix (FP64, block Jacobi)
        self.method = "pcg"
        self.pipe_replace_every = 0  # pipelined block Jacobi: residual replacement period (0: the library's)
        self.pipelined = False  # decided by run(), as refine
        # defer: the single-system driver may start the solve without waiting
        # for it (start(): a PendingSolve the owner finishes later)
        self.defer = False
        # before_solve(): called once the set-up is queued, just before the
        # solve; True when the matrix values were assembled again meanwhile,
        # and the set-up that read them is then redone
        self.before_solve: Optional[Callable[[], bool]] = None
        # isolate(): the nodes whose elements have lost their stiffness (a
        # phase-field displacement solve); set, a two-level solve that stalls
        # is redone with those nodes as aggregates of their own
        self.isolate: Optional[Callable[[], np.ndarray]] = None
        self.n_isolated = 0

    def setup(self, precision: int) -> None:
        """Block-Jacobi inverse (and the FP32 value copy) for ``precision``."""
        A = self.A
        self.precision = precision
        self.Minv = torch.empty(A.Nn * A.dof_n * A.dof_n,
                                dtype=torch.float64 if precision == 64 else torch.float32, device="cuda")
        _lib.call("efea_bj_setup", precision, A.Nn, A.dof_n, ptr(A.node_ptr), ptr(A.node_cols), ptr(A.values),
                  ptr(self.known), ptr(self.Minv), stream_ptr())
        if precision == 32:
            A.make_fp32()
        if self.precond == "two-level" and self.Ainv is None:
            self.cs = A.coarse(self.coarse_target)
            self._coarse_inverse()

    def _coarse_inverse(self) -> None:
        """``Ac^-1`` of the current values, or the owner's lagged one."""
        caches = None if self.owner is None or not getattr(self.owner, "coarse_reuse", False) else \
            self.owner.__dict__.setdefault("_coarse_caches", {})
        key = (id(self.A), self.cs.target)
        c = caches.get(key) if caches is not None else None
        if c is not None and not c.stale and c.known is self.known and c.Ainv.numel() == self.cs.ld ** 2:
            c.uses += 1
            self.Ainv, self.coarse_reused, self._cache = c.Ainv, True, c
            return
        self.Ainv = self.cs.setup(self.A, self.known)
        if caches is not None:
            self._cache = caches[key] = _CoarseCache(self.Ainv, self.known)

    def record(self, info: "SolveInfo") -> None:
        """Updates the lagged coarse inverse's rebuild test with a finished
        solve's iteration count."""
        c = self._cache
        if c is None:
            return
        if c.ref_iterations is None:
            c.ref_iterations = info.iterations
        elif info.iterations > COARSE_REUSE_GROWTH * c.ref_iterations or c.uses >= COARSE_REUSE_MAX:
            c.stale = True

    def restart_fp64(self) -> None:
        """After a failed mixed solve: the initial guess back, FP64 setup."""
        self.x.copy_(self.x0)
        self.fallback = True
        self.setup(64)

    def info(self, raw) -> SolveInfo:
        return SolveInfo(int(raw[0]), float(raw[1]), int(raw[2]), int(raw[3]),
                         "fp64" if self.precision == 64 else "mixed", self.fallback, self.precond,
                         self.coarse_reused, self.refine and self.precision == 32,
                         self.pipelined, self.n_isolated)

    def _wants_pipeline(self) -> bool:
        if self.pipelined_mode is False or self._wants_refine():
            return False
        if self.precond == "bj":
            # one launch per iteration; its grid covers the node groups in one pass
            if self.A.Nn * self.A.group > _lib.load().efea_pcgp_max_lanes():
                return False
            if self.pipelined_mode is True:
                return True
            n = self.expected_iterations
            return n is not None and n <= PIPELINE_BJ_MAX_EXPECTED and self.A.nnz * 8 < REFINE_MIN_MATRIX_BYTES
        if self.precond != "two-level":
            return False
        cs = self.A.coarse(self.coarse_target)
        if cs.nagg * cs.nm > PIPELINE_MAX_COARSE or cs.pipeline(self.A.pattern, self.A.group).nchunk > 1024:
            return False
        return self.pipelined_mode is True or self.A.nnz * 8 < REFINE_MIN_MATRIX_BYTES

    def _wants_refine(self) -> bool:
        if self.precision != 32 or self.precond != "two-level" or self.strategy == "values":
            return False
        return self.strategy == "refine" or self.A.nnz * 8 >= REFINE_MIN_MATRIX_BYTES

    def _coarse_inverse32(self) -> torch.Tensor:
        """FP32 copy of the coarse inverse (kept with a lagged inverse)."""
        c = self._cache
        if c is not None and c.Ainv32 is not None and c.Ainv is self.Ainv:
            return c.Ainv32
        a32 = torch.empty(self.Ainv.numel(), dtype=torch.float32, device="cuda")
        _lib.call("efea_cast_f64_f32", self.Ainv.numel(), ptr(self.Ainv), ptr(a32), stream_ptr())
        if c is not None and c.Ainv is self.Ainv:
            c.Ainv32 = a32
        return a32

    def _polls(self) -> tuple:
        """(check_every, first_check) of the host's state read-backs.  With
        an expected iteration count (the previous solve of a sequence), the
        first read-back comes just before it and later ones every 8
        iterations; classic FP32 solves keep the regular polls, where their
        residual replacements happen (the pipelined solve replaces on the
        device)."""
        n = self.expected_iterations
        if (self.precision == 64 or self.pipelined) and n:
            return min(self.check_every, 8), max(1, int(0.95 * n))
        return self.check_every, 0

    def _pcgp_args(self, ws: Workspace, check_every: int, first_check: int) -> tuple:
        """The arguments of ``efea_solve_pcgp`` (and of its deferred form)
        before the stream."""
        A = self.A
        v32 = A.values32 if self.precision == 32 else None
        c32 = A.corr32 if self.precision == 32 else None
        nbytes = _lib.load().efea_pcgp_workspace_bytes(A.Nn, A.dof_n)
        buf = ws.get(nbytes)
        return (self.precision, A.Nn, A.dof_n, A.group, ptr(A.node_ptr), ptr(A.node_cols), ptr(A.values), ptr(v32),
                ptr(c32), ptr(self.known), ptr(self.Minv), ptr(self.b), ptr(self.x), ptr(buf), nbytes,
                self.tol_residual, self.tol_anorm, self.pipe_replace_every, self.maxit, check_every, first_check)

    def _pcg2p_args(self, ws: Workspace, check_every: int, first_check: int) -> tuple:
        """The arguments of ``efea_solve_pcg2p`` (and of its deferred form)
        before the stream."""
        A, cs = self.A, self.cs
        v32 = A.values32 if self.precision == 32 else None
        c32 = A.corr32 if self.precision == 32 else None
        pc = cs.pipeline(A.pattern, A.group)
        nbytes = _lib.load().efea_pcg2p_workspace_bytes(A.Nn, A.dof_n, pc.nslot)
        buf = ws.get(nbytes)
        return (self.precision, A.Nn, A.dof_n, A.group, ptr(A.node_ptr), ptr(A.node_cols), ptr(A.values), ptr(v32),
                ptr(c32), ptr(self.known), ptr(self.Minv), cs.nagg, ptr(cs.agg_ptr), ptr(cs.agg_nodes), ptr(cs.rel),
                ptr(self.Ainv), cs.ld, pc.nchunk, ptr(pc.rows), ptr(pc.slot), pc.maxch, ptr(pc.first), ptr(self.b),
                ptr(self.x), ptr(buf), nbytes, self.tol_residual, self.tol_anorm, self.maxit, check_every,
                first_check)

    def _solve_once(self, ws: Workspace):
        A = self.A
        raw = (C.c_double * 4)()
        check_every, first_check = self._polls()
        v32 = A.values32 if self.precision == 32 else None
        c32 = A.corr32 if self.precision == 32 else None
        if self.precond == "bj" and self.pipelined:
            args = self._pcgp_args(ws, check_every, first_check)
            rc = _lib.call("efea_solve_pcgp", *args, stream_ptr(), C.cast(raw, C.c_void_p))
            return rc, self.info(raw)
        if self.precond == "two-level" and self.pipelined:
            args = self._pcg2p_args(ws, check_every, first_check)
            rc = _lib.call("efea_solve_pcg2p", *args, stream_ptr(), C.cast(raw, C.c_void_p))
            return rc, self.info(raw)
        if self.precond == "two-level":
            cs = self.cs
            nbytes = _lib.load().efea_pcg2_workspace_bytes(A.Nn, A.dof_n, cs.nagg)
            buf = ws.get(nbytes)
            refine = self.refine and self.precision == 32
            # mode 33 (refinement): FP32 block Jacobi and coarse inverse; the
            # replace-ratio slot carries the inner solves' accuracy
            rc = _lib.call("efea_solve_pcg2", 33 if refine else self.precision, A.Nn, A.dof_n, A.group,
                           ptr(A.node_ptr), ptr(A.node_cols), ptr(A.values), ptr(v32), ptr(c32), ptr(self.known),
                           ptr(self.Minv), cs.nagg,
                           ptr(cs.agg_ptr), ptr(cs.agg_nodes), ptr(cs.rel),
                           ptr(self._coarse_inverse32() if refine else self.Ainv), cs.ld, ptr(self.b),
                           ptr(self.x), ptr(buf), nbytes, self.tol_residual, self.tol_anorm,
                           REFINE_INNER_TOL if refine else self.replace_ratio,
                           self.maxit, check_every, first_check, self.max_blocks, stream_ptr(),
                           C.cast(raw, C.c_void_p))
            return rc, self.info(raw)
        nbytes = _lib.load().efea_pcg_workspace_bytes(A.Nn, A.dof_n, A.group)
        buf = ws.get(nbytes)
        rc = _lib.call("efea_solve_pcg", self.precision, A.Nn, A.dof_n, A.group, ptr(A.node_ptr),
                       ptr(A.node_cols), ptr(A.values), ptr(v32), ptr(c32), ptr(self.known), ptr(self.Minv),
                       ptr(self.b),
                       ptr(self.x), ptr(buf), nbytes, self.tol_residual, self.tol_anorm, self.replace_ratio,
                       self.maxit, check_every, first_check, self.max_blocks, stream_ptr(),
                       C.cast(raw, C.c_void_p))
        return rc, self.info(raw)

    def run(self, ws: Optional[Workspace] = None, raise_on_fail: bool = True) -> SolveInfo:
        """Solves now with ``efea_solve_pcg``; a mixed solve that breaks down,
        diverges or stalls is redone in FP64 from the same initial guess.
        With ``isolate`` set, the attempts before the last are capped at
        ``ISOLATE_FIRST_MAXIT`` (or 4x the expected count) iterations and a
        solve that still fails is redone, two-level, with the nodes
        ``isolate()`` names as aggregates of their own."""
        ws = ws or Workspace()
        if self.method == "bicgstab":
            return self._run_bicgstab(ws, raise_on_fail)
        self._prepare()
        if self.before_solve is not None and self.before_solve():
            self._redo_setup()
        rc, si = self._solve_once(ws)
        return self._conclude(ws, rc, si, raise_on_fail)

    def _redo_setup(self) -> None:
        """The set-up again after the matrix values changed: block-Jacobi
        inverse, FP32 copy and, if this job built it, the coarse inverse (a
        lagged one it reused stays, as it would have)."""
        if self.precond == "two-level" and not self.coarse_reused:
            if self._cache is not None:
                self._cache.stale = True
            self.Ainv = None
        self.setup(self.precision)

    def _prepare(self) -> None:
        self.refine = self._wants_refine()
        self.pipelined = self._wants_pipeline()
        isolating = self.isolate is not None and self.A.pattern.coord is not None
        if (self.pipelined or isolating) and self.x0 is None:
            self.x0 = self.x.clone()
        self._maxit = self.maxit
        if isolating:
            self.maxit = min(self.maxit, max(ISOLATE_FIRST_MAXIT, 4 * (self.expected_iterations or 0)))
        self.setup(self.precision)

    def _conclude(self, ws: Workspace, rc: int, si: SolveInfo, raise_on_fail: bool) -> SolveInfo:
        """The fallbacks of a solve that did not converge (see ``run``)."""
        isolating = self.isolate is not None and self.A.pattern.coord is not None
        if rc != 0 and self.pipelined:
            # a pipelined breakdown (rounding in its recurrences): the classic
            # iteration from the same start
            self.pipeline_failure = (rc, _lib.last_error(), si)
            self.x.copy_(self.x0)
            self.pipelined = False
            self.fallback = True
            rc, si = self._solve_once(ws)
        if rc != 0 and self.precision == 32:
            self.restart_fp64()
            rc, si = self._solve_once(ws)
        self.maxit = self._maxit
        if rc != 0 and isolating:
            rc, si = self._isolated_retry(ws, rc, si)
        if rc != 0 and raise_on_fail:
            raise _lib.EfeaError("efea_solve_pcg", rc, _lib.last_error())
        return si

    def start(self, ws: Optional[Workspace] = None) -> Optional["PendingSolve"]:
        """Starts a pipelined solve (block Jacobi or two-level) with an
        expected iteration count without waiting for it (``efea_pcgp_start`` /
        ``efea_pcg2p_start``): the start-up and ``deferred_budget`` iterations
        are queued on the current stream and the caller's host work overlaps
        them.  None (nothing queued) for any other solve, which ``run`` then
        does at once."""
        if self.method != "pcg" or not self.expected_iterations:
            return None
        self.refine = self._wants_refine()
        if not self._wants_pipeline():
            return None
        if self.precond == "bj":
            if self.isolate is not None:
                return None
            kind = "pcgp"
        elif self.precond == "two-level":
            kind = "pcg2p"
        else:
            return None
        ws = ws or Workspace()
        self._prepare()
        if self.before_solve is not None and self.before_solve():
            self._redo_setup()
        check_every, _ = self._polls()
        budget = deferred_budget(self.expected_iterations, kind)
        args = (self._pcgp_args if kind == "pcgp" else self._pcg2p_args)(ws, check_every, budget)
        s = stream_ptr()
        _lib.call(f"efea_{kind}_start", *args, s)
        return PendingSolve(self, ws, kind, args, s)

    def _run_bicgstab(self, ws: Workspace, raise_on_fail: bool) -> SolveInfo:
        """A non-symmetric system: right-preconditioned BiCGStab with the
        FP64 block-Jacobi inverse (``efea_solve_bicgstab``), to a relative
        residual of ``tol_residual``."""
        self.precond, self.pipelined, self.refine = "bj", False, False
        self.setup(64)
        if self.before_solve is not None and self.before_solve():
            self.setup(64)
        A = self.A
        raw = (C.c_double * 4)()
        nbytes = _lib.load().efea_bicgstab_workspace_bytes(A.Nn, A.dof_n)
        buf = ws.get(nbytes)
        rc = _lib.call("efea_solve_bicgstab", A.Nn, A.dof_n, A.group, ptr(A.node_ptr), ptr(A.node_cols),
                       ptr(A.values), ptr(self.known), ptr(self.Minv), ptr(self.b), ptr(self.x), ptr(buf), nbytes,
                       self.tol_residual, self.maxit, self.check_every, stream_ptr(), C.cast(raw, C.c_void_p))
        si = self.info(raw)
        if rc != 0 and raise_on_fail:
            raise _lib.EfeaError("efea_solve_bicgstab", rc, _lib.last_error())
        return si

    def _isolated_retry(self, ws: Workspace, rc: int, si: SolveInfo):
        """A stalled two-level solve of a body with fully damaged nodes: their
        elements keep only the residual stiffness, so local mechanisms of the
        broken band (and the rigid motions of separated pieces) have near-zero
        energy that block Jacobi and the aggregate modes cannot see.  Given
        aggregates of their own, those nodes' translations enter the coarse
        space and the dense coarse solve resolves them.  The retry starts
        again from the initial guess, in FP64, without the lagged inverse."""
        nodes = np.asarray(self.isolate(), dtype=np.int64)
        cs0 = self.A.coarse(self.coarse_target)
        budget = (ISOLATE_MAX_COARSE - cs0.nagg * cs0.nm) // max(cs0.nm, 1)
        if nodes.size == 0 or budget <= 0:
            return rc, si
        nodes = nodes[:budget]
        key = (self.coarse_target, nodes.tobytes())
        cache = self.A.__dict__.setdefault("_isolated_coarse", {})
        if key not in cache:
            cache.clear()
            cache[key] = CoarseSpace(self.A.pattern, cs0.target, nodes)
        self.x.copy_(self.x0)
        self.fallback = True
        self.pipelined = False
        self.precision = 64
        self.refine = False
        self.precond = "two-level"
        self.cs = cache[key]
        self.n_isolated = int(nodes.size)
        self.coarse_reused, self._cache = False, None
        self.Minv = torch.empty(self.A.Nn * self.A.dof_n * self.A.dof_n, dtype=torch.float64, device="cuda")
        _lib.call("efea_bj_setup", 64, self.A.Nn, self.A.dof_n, ptr(self.A.node_ptr), ptr(self.A.node_cols),
                  ptr(self.A.values), ptr(self.known), ptr(self.Minv), stream_ptr())
        expected, self.expected_iterations = self.expected_iterations, None
        # Ac is nearly singular here (the isolated modes keep the residual
        # stiffness): the blocked unpivoted Gauss-Jordan inverse has a
        # backward error of a few 1e-7 of max |Ac| and can lose definiteness
        # (a breakdown); a larger diagonal shift is tried then
        for shift in ISOLATE_SHIFTS:
            self.Ainv = self.cs.setup(self.A, self.known, shift)
            rc, si = self._solve_once(ws)
            if rc == 0 or si.status != 2:
                break
            self.x.copy_(self.x0)
        self.expected_iterations = expected
        return rc, si


def deferred_budget(expected: int, kind: str = "pcgp") -> int:
    """Iterations a deferred solve queues before its first read-back: the
    expected count and a margin, so that the read-back rarely finds the
    solve unfinished.  An iteration past convergence returns at once (about
    2 us per launch): block Jacobi (one launch) takes an eighth more, the
    two-level solve (two launches, iteration counts that change by a few from
    one increment to the next) a thirty-second."""
    n = int(expected)
    return n + max(4, n // 8 if kind == "pcgp" else n // 32)


class PendingSolve:
    """A solve queued by ``SolveJob.start``; ``finish`` waits for it and
    continues it if the queued iterations were not enough.  ``resumed`` is
    then True: the solution changed after the queued iterations (further
    iterations, or a fallback solve), and work queued from it in between must
    be redone."""

    def __init__(self, job: SolveJob, ws: Workspace, kind: str, args: tuple, stream: int):
        self.job, self.ws, self.kind, self.args, self.stream = job, ws, kind, args, stream
        self.resumed = False

    def finish(self, raise_on_fail: bool = True) -> SolveInfo:
        job = self.job
        raw = (C.c_double * 4)()
        resumed = C.c_int(0)
        rc = _lib.call(f"efea_{self.kind}_finish", *self.args, self.stream, C.cast(raw, C.c_void_p),
                       C.byref(resumed))
        si = job._conclude(self.ws, rc, job.info(raw), raise_on_fail)
        self.resumed = bool(resumed.value) or rc != 0
        return si


def solve(A: DeviceCsr, b: torch.Tensor, x: torch.Tensor, known: Optional[torch.Tensor], mode: str,
          ws: Workspace, tol_residual: float = 1e-12, tol_anorm: float = 1e-9, replace_ratio: float = 0.0,
          maxit: int = 100000, check_every: int = 32, raise_on_fail: bool = True) -> SolveInfo:
    """Solves the Dirichlet-reduced system on the device (``efea_solve_pcg``).

    ``x`` carries the Dirichlet values at the known DOFs and the initial guess
    elsewhere; it is overwritten with the solution.  A mixed-precision solve
    that breaks down, diverges or stalls is redone in FP64 from the same
    initial guess (``SolveInfo.fallback``)."""
    job = SolveJob(A, b, x, known, mode, tol_residual, tol_anorm, replace_ratio, maxit, check_every)
    return job.run(ws, raise_on_fail)


def drive(gen):
    """Runs a generator of ``SolveJob`` s to completion, solving each job
    at once (the single-system path), except a job marked ``defer`` that
    ``SolveJob.start`` can queue: the generator then receives the
    ``PendingSolve`` and finishes it itself.  Returns the generator's value."""
    try:
        job = next(gen)
        while True:
            pending = job.start() if job.defer else None
            job = gen.send(job.run() if pending is None else pending)
    except StopIteration as stop:
        return stop.value


class BatchSolver:
    """Continuous batching over ``efea_batch_*``: the jobs of one (precision,
    dof_n, group) class advance together, one launch per PCG step, on the
    batch's own stream.  A job's inputs come from the caller's current
    stream, which the batch stream waits for when the job joins."""

    def __init__(self, precision: int, dof_n: int, group: int, precond: str = "bj", capacity: int = 256):
        self.precision, self.dof_n, self.group, self.capacity = precision, dof_n, group, capacity
        self.precond = precond
        self.stream = torch.cuda.Stream()
        self._inflight = False
        h = C.c_void_p()
        _lib.call("efea_batch_create", precision, dof_n, group, int(precond == "two-level"), capacity, C.byref(h))
        self._h = h
        self._jobs = {}

    def __len__(self) -> int:
        return len(self._jobs)

    def full(self) -> bool:
        return len(self._jobs) >= self.capacity

    def add(self, job: SolveJob) -> int:
        A = job.A
        assert job.precision == self.precision and A.dof_n == self.dof_n and A.group == self.group
        assert job.precond == self.precond
        cs = job.cs
        if cs is not None:
            nbytes = _lib.load().efea_pcg2_workspace_bytes(A.Nn, A.dof_n, cs.nagg)
            coarse = (cs.nagg, ptr(cs.agg_ptr), ptr(cs.agg_nodes), ptr(cs.rel), ptr(job.Ainv), cs.ld)
        else:
            nbytes = _lib.load().efea_pcg_workspace_bytes(A.Nn, A.dof_n, A.group)
            coarse = (0, None, None, None, None, 0)
        job.ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        v32 = A.values32 if self.precision == 32 else None
        c32 = A.corr32 if self.precision == 32 else None
        ready = torch.cuda.Event()
        ready.record(torch.cuda.current_stream())  # the job's set-up and inputs
        self.stream.wait_event(ready)
        jid = _lib.call("efea_batch_add", self._h, A.Nn, ptr(A.node_ptr), ptr(A.node_cols), ptr(A.values), ptr(v32),
                        ptr(c32), ptr(job.known), ptr(job.Minv), *coarse, ptr(job.b), ptr(job.x), ptr(job.ws), nbytes,
                        job.tol_residual, job.tol_anorm, job.replace_ratio, job.maxit, job.max_blocks,
                        self.stream.cuda_stream)
        self._jobs[jid] = job
        return jid

    def launch(self, iterations: int) -> bool:
        """Enqueues ``iterations`` PCG iterations of every job (returns at
        once); ``wait`` collects them."""
        if not self._jobs or self._inflight:
            return False
        _lib.call("efea_batch_launch", self._h, iterations, self.stream.cuda_stream)
        self._inflight = True
        return True

    def wait(self) -> list:
        """Waits for the launched chunk; the finished ``(job, SolveInfo)``
        pairs (their results are complete on the device)."""
        if not self._inflight:
            return []
        self._inflight = False
        n = _lib.call("efea_batch_wait", self._h, self.stream.cuda_stream)
        return self._pop(n)

    def run(self, iterations: int) -> list:
        """``iterations`` PCG iterations of every job, then the finished
        ``(job, SolveInfo)`` pairs."""
        self.launch(iterations)
        return self.wait()

    def _pop(self, n: int) -> list:
        out = []
        jid = C.c_int()
        raw = (C.c_double * 4)()
        for _ in range(n):
            _lib.call("efea_batch_pop", self._h, C.byref(jid), C.cast(raw, C.c_void_p))
            job = self._jobs.pop(jid.value)
            job.ws = None
            out.append((job, job.info(raw)))
        return out

    def close(self) -> None:
        if self._h is not None and _lib._lib is not None:
            _lib.call("efea_batch_destroy", self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchExecutor:
    """Routes solve jobs to one ``BatchSolver`` per (precision, dof_n, group,
    preconditioner) class and runs them ``check_every`` iterations at a time.  A mixed job
    that fails is restarted in FP64 from its initial guess, as the single
    path does."""

    def __init__(self, capacity: int = 256, check_every: int = 32):
        self.capacity = capacity
        self.check_every = check_every
        self.batches = {}
        self.waiting = {}  # jobs of a full batch, in submission order

    def submit(self, job: SolveJob) -> None:
        if job.method != "pcg":
            raise NotImplementedError(f"batched solves are PCG solves; a {job.method} job runs alone (job.run())")
        key = (job.precision, job.A.dof_n, job.A.group, job.precond)
        b = self.batches.get(key)
        if b is None:
            b = self.batches[key] = BatchSolver(*key, capacity=self.capacity)
            self.waiting[key] = []
        if b.full() or self.waiting[key]:
            self.waiting[key].append(job)
            return
        job.setup(job.precision)
        b.add(job)

    def active(self) -> int:
        return sum(len(b) for b in self.batches.values()) + sum(len(w) for w in self.waiting.values())

    def launch(self) -> bool:
        """Enqueues one chunk of every non-empty batch (returns at once)."""
        launched = False
        for b in self.batches.values():
            launched |= b.launch(self.check_every)
        return launched

    def step(self) -> list:
        """One chunk of every non-empty batch, then ``collect``."""
        self.launch()
        return self.collect()

    def collect(self) -> list:
        """Waits for the launched chunks; returns the completed
        ``(job, SolveInfo)`` pairs (fallback restarts are resubmitted);
        waiting jobs take the slots that freed up."""
        done = []
        for key, b in list(self.batches.items()):
            for job, si in b.wait():
                if si.status != 1 and job.precision == 32:
                    job.restart_fp64()
                    self.submit(job)
                    continue
                done.append((job, si))
            w = self.waiting[key]
            while w and not b.full():
                job = w.pop(0)
                job.setup(job.precision)
                b.add(job)
        return done

    def close(self) -> None:
        for b in self.batches.values():
            b.close()
        self.batches = {}
