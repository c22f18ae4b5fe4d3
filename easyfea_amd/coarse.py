"""Coarse space of the two-level preconditioner: node aggregates with their
rigid-body modes.

The PCG preconditioner ``M^-1 = B^-1 + P Ac^-1 P^T`` adds to block Jacobi
(``B``: each node's ``dof_n x dof_n`` diagonal block) a coarse correction over
aggregates of nodes.  ``P`` holds, per aggregate, the rigid-body modes of its
nodes (2D: two translations and the rotation about the aggregate centroid;
3D: three translations and three rotations; scalar problems: the constant),
restricted to the unknown DOFs; ``Ac = P^T A P`` is assembled and inverted on
the device at each new matrix.  This is the additive two-level Schwarz
method with an aggregation coarse space (smoothed-aggregation style
tentative prolongator).  EasyFEA solves directly (``Solvers.py:225-394``), so
the preconditioner changes iteration counts, not the solution.

Aggregates: recursive coordinate bisection of the nodes into parts of at
most ``target`` nodes, then connected components of each part in the node
graph (a part that straddles an open crack splits into its two lips), then
components smaller than ``target // 4`` merged into a neighbouring aggregate.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components


def n_modes(dim: int, dof_n: int) -> int:
    if dof_n == 1:
        return 1
    return 3 if dim == 2 else 6


@dataclass
class Aggregation:
    agg: np.ndarray        # (Nn,) int32 aggregate of each node
    nagg: int
    agg_ptr: np.ndarray    # (nagg+1,) int32 offsets into agg_nodes
    agg_nodes: np.ndarray  # (Nn,) int32 nodes grouped by aggregate, increasing within one
    rel: np.ndarray        # (Nn, 3) FP64 node coordinates relative to their aggregate centroid
    cadj_ptr: np.ndarray   # (nagg+1,) int32 coarse adjacency (aggregates sharing a matrix block)
    cadj: np.ndarray       # int32, sorted per aggregate, includes the aggregate itself


def _rcb(idx: np.ndarray, coord: np.ndarray, target: int, out: list) -> None:
    stack = [idx]
    while stack:
        ids = stack.pop()
        if ids.size <= target:
            out.append(ids)
            continue
        c = coord[ids]
        ax = int(np.argmax(c.max(0) - c.min(0)))
        order = np.argsort(c[:, ax], kind="stable")
        h = ids.size // 2
        stack.append(ids[order[h:]])
        stack.append(ids[order[:h]])


def aggregate(coord: np.ndarray, node_ptr: np.ndarray, node_cols: np.ndarray, target: int,
              singletons: Optional[np.ndarray] = None) -> Aggregation:
    """``singletons``: nodes that become aggregates of their own (after the
    bisection, the component split and the merge), so that the coarse space
    holds their translations exactly (fully damaged nodes)."""
    Nn = coord.shape[0]
    parts: list = []
    _rcb(np.arange(Nn), coord, max(int(target), 1), parts)
    part = np.empty(Nn, dtype=np.int64)
    for k, ids in enumerate(parts):
        part[ids] = k
    rows = np.repeat(np.arange(Nn), np.diff(node_ptr))
    cols = node_cols.astype(np.int64)
    same = part[rows] == part[cols]
    G = sp.csr_matrix((np.ones(int(same.sum())), (rows[same], cols[same])), shape=(Nn, Nn))
    ncomp, comp = connected_components(G, directed=False)
    # merge small components into the neighbouring component they touch most
    sizes = np.bincount(comp, minlength=ncomp)
    small = sizes < max(target // 4, 1)
    if small.any():
        GA = sp.csr_matrix((np.ones(rows.size), (comp[rows], comp[cols])), shape=(ncomp, ncomp)).tolil()
        target_of = np.arange(ncomp)
        for c in np.nonzero(small)[0]:
            nb = np.array(GA.rows[c], dtype=np.int64)
            w = np.array(GA.data[c])
            keep = (nb != c) & ~small[nb]
            if keep.any():
                target_of[c] = nb[keep][np.argmax(w[keep])]
        comp = target_of[comp]
    if singletons is not None and len(singletons):
        # the other nodes of an aggregate that loses nodes may fall apart
        # (two lips of a broken band): its pieces become aggregates
        single = np.zeros(Nn, dtype=bool)
        single[np.asarray(singletons, dtype=np.int64)] = True
        keep = (comp[rows] == comp[cols]) & ~single[rows] & ~single[cols]
        G2 = sp.csr_matrix((np.ones(int(keep.sum())), (rows[keep], cols[keep])), shape=(Nn, Nn))
        _, comp = connected_components(G2, directed=False)
        comp = comp.astype(np.int64)
        comp[single] = comp.max() + 1 + np.arange(int(single.sum()))
    uniq, agg = np.unique(comp, return_inverse=True)
    nagg = uniq.size
    agg = agg.astype(np.int32)
    order = np.argsort(agg, kind="stable").astype(np.int32)
    agg_ptr = np.zeros(nagg + 1, dtype=np.int32)
    agg_ptr[1:] = np.cumsum(np.bincount(agg, minlength=nagg))
    cnt = np.bincount(agg, minlength=nagg).astype(float)
    cen = np.stack([np.bincount(agg, weights=coord[:, k], minlength=nagg) / cnt for k in range(3)], 1)
    rel = np.ascontiguousarray(coord - cen[agg])
    CA = sp.csr_matrix((np.ones(rows.size), (agg[rows], agg[cols])), shape=(nagg, nagg))
    CA.sum_duplicates()
    CA.sort_indices()
    return Aggregation(agg, nagg, agg_ptr, order, rel, CA.indptr.astype(np.int32), CA.indices.astype(np.int32))


def default_target(dim: int) -> int:
    """Largest part of the bisection: parts hold target/2 .. target nodes
    (2D: ~12 x 12, 3D: ~9 x 9 x 9), which keeps the dense coarse inverse at a
    few hundred (2D) to a few thousand (3D) unknowns on the measured meshes."""
    return 256 if dim == 2 else 1024


def prolongator(ag: Aggregation, dim: int, dof_n: int, known: np.ndarray) -> sp.csr_matrix:
    """``P`` (Ndof x nagg*nm) as a scipy matrix (host reference of the device
    prolongation)."""
    Nn = ag.agg.size
    nm = n_modes(dim, dof_n)
    x, y, z = ag.rel[:, 0], ag.rel[:, 1], ag.rel[:, 2]
    one, zero = np.ones(Nn), np.zeros(Nn)
    if dof_n == 1:
        modes = [[one]]
    elif dim == 2:
        modes = [[one, zero], [zero, one], [-y, x]]
    else:
        modes = [[one, zero, zero], [zero, one, zero], [zero, zero, one],
                 [zero, -z, y], [z, zero, -x], [-y, x, zero]]
    rows, cols, vals = [], [], []
    n = np.arange(Nn)
    for m, comps in enumerate(modes):
        for c, v in enumerate(comps):
            rows.append(n * dof_n + c)
            cols.append(ag.agg.astype(np.int64) * nm + m)
            vals.append(v)
    P = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(Nn * dof_n, ag.nagg * nm))
    P = sp.diags((~np.asarray(known, dtype=bool)).astype(float)) @ P
    P.eliminate_zeros()
    return P.tocsr()


@dataclass
class CoarsePairs:
    """Node pairs (node i, position of neighbour j in i's adjacency) grouped
    by the aggregate block (agg(i), agg(j)) they contribute to in
    ``Ac = P^T A P``; within a block in increasing (i, position) order."""
    pair_ptr: np.ndarray   # (ncb+1,) int32
    pair_node: np.ndarray  # int32
    pair_pos: np.ndarray   # int32
    blk_I: np.ndarray      # (ncb,) int32
    blk_J: np.ndarray      # (ncb,) int32


def coarse_pairs(ag: Aggregation, node_ptr: np.ndarray, node_cols: np.ndarray) -> CoarsePairs:
    Nn = ag.agg.size
    cnt = np.diff(node_ptr)
    node = np.repeat(np.arange(Nn, dtype=np.int64), cnt)
    pos = np.arange(node.size, dtype=np.int64) - np.repeat(node_ptr[:-1].astype(np.int64), cnt)
    I = ag.agg[node].astype(np.int64)
    J = ag.agg[node_cols].astype(np.int64)
    key = I * ag.nagg + J
    order = np.argsort(key, kind="stable")  # stable: (i, pos) order kept inside a block
    key_s = key[order]
    ub, start = np.unique(key_s, return_index=True)
    pair_ptr = np.append(start, key_s.size).astype(np.int32)
    return CoarsePairs(pair_ptr, node[order].astype(np.int32), pos[order].astype(np.int32),
                       (ub // ag.nagg).astype(np.int32), (ub % ag.nagg).astype(np.int32))
