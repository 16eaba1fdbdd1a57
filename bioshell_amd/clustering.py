"""Hierarchical clustering and row reductions over cRMSD matrices.

The host mirror of ``bioshell_clustering::hierarchical``
(bioshell-clustering/src/hierarchical/hierarchical.rs).  The O(n^3)
agglomeration loop -- closest-pair search plus the Lance-Williams row update
-- runs on the GPU (``bs_hierarchical_clustering``); the row minima and maxima
that outlier detection and the min-max medoid reduce to run on the GPU as well
(``bs_rmsd_row_stats``, ``bs_rmsd_members_row_max``).  This module rebuilds the
clustering tree from the device's merge sequence exactly as the reference's
loop does (hierarchical.rs:22-80) and provides the reference's tree
functions on top of it, with the same names and results.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import Callable, Sequence

import numpy as np

from . import _native
from .crmsd import Context, RmsdMatrix, check, default_context

LINKAGES = ("single", "complete", "average", "median", "centroid", "wards")


def linkage_of_cli(complete=False, average=False, wards=False, median=False,
                   centroid=False) -> str:
    """The linkage the reference's ``clust`` tool selects for its flags
    (bioshell-clustering/src/bin/clust.rs:66-80): ``--median`` runs average
    linkage and ``--centroid`` runs Ward's method."""
    if complete:
        return "complete"
    if average:
        return "average"
    if wards:
        return "wards"
    if median:
        return "average"
    if centroid:
        return "wards"
    return "single"


@dataclass
class ClusteringTreeNode:
    """A node of the clustering tree (``BinaryTreeNode<HierarchicalCluster>``):
    leaves carry the data index as ``id``; inner nodes get ids n, n + 1, ... in
    merge order."""
    id: int
    cluster_size: int
    merging_distance: float
    left: "ClusteringTreeNode | None" = field(default=None, repr=False)
    right: "ClusteringTreeNode | None" = field(default=None, repr=False)

    def is_leaf(self) -> bool:
        return self.left is None and self.right is None

    def rotate(self) -> None:
        """Mirrors the subtree: children swapped at every node (tree.rs:106-118)."""
        stack = [self]
        while stack:
            nd = stack.pop()
            if nd.left is not None:
                stack.append(nd.left)
            if nd.right is not None:
                stack.append(nd.right)
            nd.left, nd.right = nd.right, nd.left


def tree_from_merges(n: int, merge_i, merge_j, merge_d) -> ClusteringTreeNode:
    """Replays a merge sequence with the reference's index bookkeeping: the
    merged cluster takes index i, the last active cluster moves to index j
    (hierarchical.rs:44-78)."""
    if n < 1:
        raise ValueError("nothing to cluster")
    clusters = {i: ClusteringTreeNode(i, 1, 0.0) for i in range(n)}
    new_id = n
    order = n
    for i, j, d in zip(merge_i, merge_j, merge_d):
        i, j = int(i), int(j)
        ci, cj = clusters.pop(i), clusters.pop(j)
        clusters[i] = ClusteringTreeNode(new_id, ci.cluster_size + cj.cluster_size, float(d), ci, cj)
        last = order - 1
        if j < last:
            clusters[j] = clusters.pop(last)
        order -= 1
        new_id += 1
    return clusters[0]


def _packed(n: int, distance: Callable[[int, int], float]) -> np.ndarray:
    """Packed upper triangle from d(i, j), called with i > j as the reference
    fills its matrix (clustering_matrix.rs:11-21)."""
    up = np.empty(n * (n - 1) // 2, dtype=np.float32)
    k = 0
    for i in range(n):
        for j in range(i + 1, n):
            up[k] = distance(j, i)
            k += 1
    return up


def _upper_of(data) -> tuple[int, np.ndarray]:
    if isinstance(data, RmsdMatrix):
        return data.n_elements(), np.ascontiguousarray(data.upper, dtype=np.float32)
    raise TypeError("expected an RmsdMatrix")


def hierarchical_clustering(data, distance: Callable[[int, int], float] | None = None,
                            linkage: str = "single", ctx: Context | None = None,
                            device_upper: int | None = None) -> ClusteringTreeNode:
    """Agglomerative clustering (hierarchical.rs:22-80) on the GPU.

    ``data`` is an :class:`RmsdMatrix`, or the number of elements together with
    ``distance(i, j)``; with ``device_upper`` (a device pointer to the packed
    f32 triangle, e.g. the output of ``crmsd_all_vs_all_device``) ``data`` is
    the number of elements and the matrix never leaves HBM.  ``linkage`` is one
    of ``LINKAGES`` (strategies/mod.rs).
    """
    if linkage not in LINKAGES:
        raise ValueError(f"unknown linkage {linkage!r}")
    ctx = ctx or default_context()
    if device_upper is not None:
        n = int(data)
        up = None
    elif isinstance(data, RmsdMatrix):
        n, up = _upper_of(data)
    else:
        n = int(data)
        if distance is None:
            raise ValueError("a distance function is needed with an element count")
        up = _packed(n, distance)
    if n < 1:
        raise ValueError("nothing to cluster")
    m = max(n - 1, 0)
    mi = np.zeros(max(m, 1), dtype=np.uint32)
    mj = np.zeros(max(m, 1), dtype=np.uint32)
    md = np.zeros(max(m, 1), dtype=np.float32)
    args = (n, LINKAGES.index(linkage), mi.ctypes.data_as(_native.c_u32p),
            mj.ctypes.data_as(_native.c_u32p), md.ctypes.data_as(_native.c_fp))
    if device_upper is not None:
        check(_native.lib().bs_hierarchical_clustering_device(ctx.handle, ctypes.c_void_p(device_upper), *args))
    else:
        check(_native.lib().bs_hierarchical_clustering(ctx.handle, up.ctypes.data_as(_native.c_fp), *args))
    return tree_from_merges(n, mi[:m], mj[:m], md[:m])


def retrieve_data_id(cluster: ClusteringTreeNode) -> list[int]:
    """Leaf ids in depth-first preorder (hierarchical.rs:230-238)."""
    out, stack = [], [cluster]
    while stack:
        nd = stack.pop()
        if nd.is_leaf():
            out.append(nd.id)
        if nd.right is not None:
            stack.append(nd.right)
        if nd.left is not None:
            stack.append(nd.left)
    return out


def retrieve_data(cluster: ClusteringTreeNode, all_data: Sequence) -> list:
    """Data items in tree order (hierarchical.rs:244-250)."""
    return [all_data[i] for i in retrieve_data_id(cluster)]


def retrieve_clusters(root: ClusteringTreeNode, max_distance: float) -> list[ClusteringTreeNode]:
    """Subtrees cut at ``max_distance`` (hierarchical.rs:137-172)."""
    out: list[ClusteringTreeNode] = []

    def rec(node):
        if node.is_leaf():
            out.append(node)
        elif node.merging_distance > max_distance:
            for ch in (node.left, node.right):
                if ch is None:
                    continue
                if ch.merging_distance <= max_distance:
                    out.append(ch)
                else:
                    rec(ch)

    if root.merging_distance <= max_distance:
        out.append(root)
    else:
        rec(root)
    return out


def _leftmost(c):
    while c.left is not None:
        c = c.left
    return c.id


def _rightmost(c):
    while c.right is not None:
        c = c.right
    return c.id


def _if_rotate(c, d) -> tuple[bool, bool]:
    """hierarchical.rs:276-313"""
    if c.is_leaf():
        return False, False
    left, right = c.left, c.right
    if right.is_leaf() and left.is_leaf():
        return False, False
    if right.is_leaf():
        return (d(right.id, _leftmost(left)) < d(right.id, _rightmost(left))), False
    if left.is_leaf():
        return False, (d(left.id, _leftmost(right)) > d(left.id, _rightmost(right)))
    rr, rl, lr, ll = _rightmost(right), _leftmost(right), _rightmost(left), _leftmost(left)
    ds = [d(lr, rl), d(ll, rl), d(lr, rr), d(ll, rr)]
    k = int(np.argmin(np.asarray(ds, dtype=np.float32)))  # first minimum
    return k in (1, 3), k in (2, 3)


def balance_clustering_tree(root: ClusteringTreeNode, distance: Callable[[int, int], float]) -> None:
    """Rotates subtrees bottom-up so neighbouring leaves are close
    (hierarchical.rs:82-98)."""
    # post-order without recursion
    stack, order = [root], []
    while stack:
        nd = stack.pop()
        order.append(nd)
        if nd.left is not None:
            stack.append(nd.left)
        if nd.right is not None:
            stack.append(nd.right)
    for nd in reversed(order):
        rl, rr = _if_rotate(nd, distance)
        if rl:
            nd.left.rotate()
        if rr:
            nd.right.rotate()


def medoid_by_min_max(cluster: ClusteringTreeNode, distance, ctx: Context | None = None) -> int:
    """The member whose largest distance to the other members is smallest
    (first such member in tree order; hierarchical.rs:105-131).  With an
    :class:`RmsdMatrix` the row maxima are reduced on the GPU."""
    members = retrieve_data_id(cluster)
    if len(members) == 1:
        return members[0]
    if isinstance(distance, RmsdMatrix):
        rmax = members_row_max(distance, members, ctx=ctx)
    else:
        rmax = np.array([max(distance(a, b) for b in members if b != a) for a in members],
                        dtype=np.float32)
    best, best_i = np.float32(np.finfo(np.float32).max), 0
    for i, v in enumerate(rmax):
        if v < best:
            best, best_i = v, i
    return members[best_i]


def row_stats(matrix: RmsdMatrix, ctx: Context | None = None) -> tuple[np.ndarray, np.ndarray]:
    """Row minimum and maximum over j != i, on the GPU (bs_rmsd_row_stats)."""
    n, up = _upper_of(matrix)
    ctx = ctx or default_context()
    lo = np.zeros(n, dtype=np.float32)
    hi = np.zeros(n, dtype=np.float32)
    check(_native.lib().bs_rmsd_row_stats(ctx.handle, up.ctypes.data_as(_native.c_fp), n,
                                           lo.ctypes.data_as(_native.c_fp), hi.ctypes.data_as(_native.c_fp)))
    return lo, hi


def members_row_max(matrix: RmsdMatrix, members: Sequence[int], ctx: Context | None = None) -> np.ndarray:
    """For each member, its largest distance to the other members (GPU)."""
    n, up = _upper_of(matrix)
    ctx = ctx or default_context()
    mem = np.ascontiguousarray(members, dtype=np.uint32)
    out = np.zeros(len(mem), dtype=np.float32)
    check(_native.lib().bs_rmsd_members_row_max(ctx.handle, up.ctypes.data_as(_native.c_fp), n,
                                                 mem.ctypes.data_as(_native.c_u32p), len(mem),
                                                 out.ctypes.data_as(_native.c_fp)))
    return out


def retrieve_outliers(data, distance: Callable[[int, int], float] | None = None,
                      cutoff: float = 0.0, ctx: Context | None = None) -> list[int]:
    """Elements with no other element within ``cutoff`` (hierarchical.rs:198-217).
    With an :class:`RmsdMatrix` the row minima are reduced on the GPU."""
    if isinstance(data, RmsdMatrix):
        if data.n_elements() < 2:
            return []
        lo, _ = row_stats(data, ctx=ctx)
        return [int(i) for i in np.nonzero(lo > np.float32(cutoff))[0]]
    n = int(data)
    if n < 2:
        return []
    out = []
    for i in range(n):
        m = min(distance(i, j) for j in range(n) if j != i)
        if m > cutoff:
            out.append(i)
    return out


def write_tsv(matrix: RmsdMatrix, path: str, labels: Sequence[str] | None = None, decimals: int = -1,
              both_triangles: bool = False, n_threads: int = 0) -> None:
    """Writes the matrix as the TSV that the reference's ``clust`` tool reads
    (bs_write_distance_tsv): ``label_i<TAB>label_j<TAB>value`` for every pair
    i < j in row order (every ordered pair with ``both_triangles``).  Values
    are the shortest text that reads back to the same f32 unless
    ``decimals`` >= 0 is given."""
    n, up = _upper_of(matrix)
    arr = None
    if labels is not None:
        if len(labels) != n:
            raise ValueError(f"{len(labels)} labels for {n} elements")
        enc = [str(l).encode() for l in labels]
        if any(b"\t" in e or b"\n" in e for e in enc):
            raise ValueError("labels may not contain tabs or newlines")
        arr = (ctypes.c_char_p * n)(*enc)
    check(_native.lib().bs_write_distance_tsv(up.ctypes.data_as(_native.c_fp), n, arr, int(decimals),
                                               int(bool(both_triangles)), int(n_threads), os.fsencode(path)))


def read_tsv(path: str, symmetric: bool = True, default_distance: float = 100.0) -> tuple[RmsdMatrix, list[str]]:
    """Reads ``label_i label_j value`` lines into a matrix and its labels,
    numbered in order of first appearance, as ``clust`` loads its input
    (clust.rs:102-107 with DataMatrixBuilder, symmetric as with --symmetric;
    pairs absent from the file take ``default_distance``; distances are used
    as f32, distance_matrix.rs:71-74)."""
    ids: dict[str, int] = {}
    vals: dict[tuple[int, int], float] = {}
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if len(parts) < 3:
                raise ValueError(f"expected 'label label value': {line!r}")
            a, b, v = parts[0], parts[1], float(parts[2])
            for x in (a, b):
                ids.setdefault(x, len(ids))
            vals[(ids[a], ids[b])] = v
            if symmetric:
                vals[(ids[b], ids[a])] = v
    n = len(ids)
    up = np.empty(n * (n - 1) // 2, dtype=np.float32)
    k = 0
    for i in range(n):
        for j in range(i + 1, n):
            # hierarchical_clustering asks for d(j, i) with j > i (clustering_matrix.rs:11-21)
            up[k] = np.float32(vals.get((j, i), vals.get((i, j), default_distance)) if symmetric
                               else vals.get((j, i), default_distance))
            k += 1
    return RmsdMatrix(n, up), list(ids)


def show_clustering_tree(root: ClusteringTreeNode) -> str:
    """hierarchical.rs:296-319"""
    lines, stack = [], [(root, 0)]
    while stack:
        nd, ind = stack.pop()
        ls = str(nd.left.id) if nd.left is not None else "-"
        rs = str(nd.right.id) if nd.right is not None else "-"
        lines.append("  " * ind + f"{nd.id} : {ls} {rs} @ {nd.merging_distance}")
        if nd.right is not None:
            stack.append((nd.right, ind + 1))
        if nd.left is not None:
            stack.append((nd.left, ind + 1))
    return "\n".join(lines) + "\n"
