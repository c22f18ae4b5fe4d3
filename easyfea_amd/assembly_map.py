"""Sparsity pattern and deterministic scatter map (one-time, per mesh).

EasyFEA builds its CSR pattern by canonicalising the COO triplets of every
element entry with scipy and maps each entry to its slot with
``searchsorted`` (``Simulations/_simu.py:1062-1102``).  The pattern is fully
determined by node adjacency: row ``n*dof_n + a`` holds the columns
``m*dof_n + b`` of every node ``m`` sharing an element with ``n`` (itself
included), sorted.  This module derives, from the connectivity alone,

* the node adjacency ``node_ptr`` / ``node_cols`` (the SpMV index arrays),
* scipy's canonical ``indptr`` / ``indices`` (int32) for export,
* the gather-form scatter map: ``seg_ptr`` (slot -> range) and ``perm``
  (element entries of each slot in increasing entry index ``e*nd^2 + i*nd + j``),

working at node-pair granularity (``Ne * nPe^2`` pairs) and expanding to DOF
entries arithmetically, instead of sorting all ``Ne * nd^2`` entries.
"""

from dataclasses import dataclass

import numpy as np


@dataclass
class CsrMap:
    Nn: int
    dof_n: int
    nPe: int
    node_ptr: np.ndarray  # int32 (Nn+1,)
    node_cols: np.ndarray  # int32 (nnz_node,)
    indptr: np.ndarray  # int32 (Ndof+1,)
    indices: np.ndarray  # int32 (nnz,)
    seg_ptr: np.ndarray  # int32 (nnz+1,)
    perm: np.ndarray  # int32 (Ne*nd*nd,)
    max_neighbours: int

    @property
    def nnz(self) -> int:
        return int(self.indices.size)

    @property
    def Ndof(self) -> int:
        return self.Nn * self.dof_n


def node_adjacency(connect: np.ndarray, Nn: int):
    """Sorted unique neighbour nodes per node (self included) and, for each
    element node pair ``(e, i, j)``, the index of ``connect[e, j]`` in the
    adjacency of ``connect[e, i]``."""
    connect = np.asarray(connect, dtype=np.int64)
    Ne, nPe = connect.shape
    rows = np.repeat(connect, nPe, axis=1).ravel()  # (e, i, j) -> connect[e, i]
    cols = np.tile(connect, (1, nPe)).ravel()  # (e, i, j) -> connect[e, j]
    key = rows * Nn + cols
    ukey = np.unique(key)
    node_rows = (ukey // Nn).astype(np.int64)
    node_cols = (ukey % Nn).astype(np.int32)
    counts = np.bincount(node_rows, minlength=Nn)
    node_ptr = np.zeros(Nn + 1, dtype=np.int64)
    np.cumsum(counts, out=node_ptr[1:])
    pair_slot = np.searchsorted(ukey, key)  # global index into node_cols
    return node_ptr, node_cols, pair_slot


def build_csr_map(connect: np.ndarray, Nn: int, dof_n: int) -> CsrMap:
    connect = np.asarray(connect, dtype=np.int64)
    Ne, nPe = connect.shape
    dn = dof_n
    nd = nPe * dn
    node_ptr, node_cols, pair_slot = node_adjacency(connect, Nn)
    cnt = np.diff(node_ptr)
    nnz_node = node_cols.size
    nnz = nnz_node * dn * dn
    if nnz >= 2**31 or Ne * nd * nd >= 2**31:
        raise ValueError("system too large for int32 scatter maps")

    # DOF-level canonical pattern
    row_node = np.repeat(np.arange(Nn), cnt)  # row node of each node slot
    Ndof = Nn * dn
    indptr = np.zeros(Ndof + 1, dtype=np.int64)
    rowlen = np.repeat(cnt * dn, dn)  # each of the dn rows of node n has cnt*dn entries
    np.cumsum(rowlen, out=indptr[1:])
    # indices: for node n, row a: for pos: for b -> node_cols*dn + b
    seg_cols = (node_cols.astype(np.int64)[:, None] * dn + np.arange(dn)[None, :]).ravel()  # per node slot, b
    # order rows: for n, for a, repeat the node's seg_cols block
    indices = np.empty(nnz, dtype=np.int64)
    # position of node n's block for row a starts at indptr[n*dn + a]
    base_node = dn * dn * node_ptr[:-1]
    for a in range(dn):
        # destination for node slot q (row node n, local pos p) and component b
        q = np.arange(nnz_node)
        n = row_node
        p = q - node_ptr[n]
        dst = base_node[n] + a * dn * cnt[n] + p * dn
        for b in range(dn):
            indices[dst + b] = node_cols[q].astype(np.int64) * dn + b
    del seg_cols

    # contributing element pairs per node slot, in increasing element order
    order = np.argsort(pair_slot, kind="stable")
    cntp = np.bincount(pair_slot, minlength=nnz_node)
    pstart = np.zeros(nnz_node + 1, dtype=np.int64)
    np.cumsum(cntp, out=pstart[1:])
    pe = order // (nPe * nPe)
    pli = (order // nPe) % nPe
    plj = order % nPe

    # slot s of (node slot q, a, b): base_node[n] + a*dn*cnt[n] + p*dn + b
    q = np.arange(nnz_node)
    n = row_node
    p = q - node_ptr[n]
    seglen = np.empty(nnz, dtype=np.int64)
    slot_q = np.empty(nnz, dtype=np.int64)
    slot_a = np.empty(nnz, dtype=np.int64)
    slot_b = np.empty(nnz, dtype=np.int64)
    for a in range(dn):
        for b in range(dn):
            s = base_node[n] + a * dn * cnt[n] + p * dn + b
            seglen[s] = cntp[q]
            slot_q[s] = q
            slot_a[s] = a
            slot_b[s] = b
    seg_ptr = np.zeros(nnz + 1, dtype=np.int64)
    np.cumsum(seglen, out=seg_ptr[1:])
    E = int(seg_ptr[-1])
    assert E == Ne * nd * nd
    # expand: entry t of slot s is pair order[pstart[q] + t]
    slot_of = np.repeat(np.arange(nnz), seglen)
    t = np.arange(E) - seg_ptr[slot_of]
    qq = slot_q[slot_of]
    pidx = pstart[qq] + t
    del t, qq
    aa = slot_a[slot_of]
    bb = slot_b[slot_of]
    del slot_of
    perm = pe[pidx] * (nd * nd) + (pli[pidx] * dn + aa) * nd + plj[pidx] * dn + bb
    del pidx, aa, bb
    return CsrMap(
        Nn=Nn,
        dof_n=dn,
        nPe=nPe,
        node_ptr=node_ptr.astype(np.int32),
        node_cols=node_cols.astype(np.int32),
        indptr=indptr.astype(np.int32),
        indices=indices.astype(np.int32),
        seg_ptr=seg_ptr.astype(np.int32),
        perm=perm.astype(np.int32),
        max_neighbours=int(cnt.max()) if cnt.size else 0,
    )


@dataclass
class VecMap:
    """Scatter map of element vectors into a global vector (``isMatrix=False``):
    slot = DOF; ``used`` lists the DOFs that appear (the (Ndof, 1) CSR pattern)."""

    Ndof: int
    seg_ptr: np.ndarray
    perm: np.ndarray
    used: np.ndarray


def build_vec_map(connect: np.ndarray, Nn: int, dof_n: int) -> VecMap:
    connect = np.asarray(connect, dtype=np.int64)
    Ne, nPe = connect.shape
    asm = np.empty((Ne, nPe * dof_n), dtype=np.int64)
    for d in range(dof_n):
        asm[:, d::dof_n] = connect * dof_n + d
    rows = asm.ravel()
    Ndof = Nn * dof_n
    perm = np.argsort(rows, kind="stable")
    counts = np.bincount(rows, minlength=Ndof)
    seg = np.zeros(Ndof + 1, dtype=np.int64)
    np.cumsum(counts, out=seg[1:])
    return VecMap(Ndof, seg.astype(np.int32), perm.astype(np.int32), np.nonzero(counts)[0].astype(np.int32))
