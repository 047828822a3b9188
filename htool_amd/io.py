"""Interchange formats of the reference, read and written from Python.

* `save_plot` / `load_plot`: the leaf list CSV of HMatrix::save_plot
  (include/htool/types/hmatrix.hpp:1695-1712): a first line "nr,nc", then one
  line "offset_i,rows,offset_j,cols,rank" per leaf, dense leaves (rank -1)
  first, file name `<name>_<rank>.csv` (rank 0 here, one process).
* `cluster_tree`, `save_cluster` / `read_cluster`: the cluster tree of a point
  set (the tree the H-matrix builds use) and the two CSV files of
  Cluster::save_cluster / read_cluster (include/htool/clustering/cluster.hpp:
  433-583): `<name>_permutation.csv`, one line of comma-separated indices, and
  `<name>_tree.csv`, one line per depth, nodes in depth-first order separated by
  ',', each node "nb_sons|rank|offset|size|rad|ctr_x|ctr_y|ctr_z" with numbers
  printed as a C++ stream prints them (6 significant digits).
* `matrix_to_bytes` / `bytes_to_matrix` (include/htool/types/matrix.hpp:466-509):
  int32 rows, int32 cols, then the column-major entries; `vector_to_bytes` /
  `bytes_to_vector` (include/htool/types/vector.hpp:129-161): int32 size, then
  the entries. Entry types are the caller's (float64 or complex128).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import capi


# -- leaf list (save_plot) ------------------------------------------------------
def save_plot(H, outputname: str) -> str:
    """HMatrix::save_plot: writes `<outputname>_0.csv`, returns its path."""
    leaves = H.get_leaves()
    path = f"{outputname}_0.csv"
    with open(path, "w") as f:
        f.write(f"{H.nb_rows()},{H.nb_cols()}\n")
        for oi, m, oj, n, k in leaves[leaves[:, 4] < 0]:
            f.write(f"{oi},{m},{oj},{n},-1\n")
        for oi, m, oj, n, k in leaves[leaves[:, 4] >= 0]:
            f.write(f"{oi},{m},{oj},{n},{k}\n")
    return path


def load_plot(path: str):
    """(nr, nc, leaves) from a save_plot file; leaves sorted by (offset_i, offset_j)."""
    with open(path) as f:
        nr, nc = (int(v) for v in f.readline().split(","))
        rows = [[int(v) for v in line.split(",")] for line in f if line.strip()]
    leaves = np.array(rows, dtype=np.int32).reshape(-1, 5)
    order = np.lexsort((leaves[:, 2], leaves[:, 0]))
    return nr, nc, leaves[order]


# -- cluster trees ------------------------------------------------------------------
@dataclass
class ClusterTree:
    """Nodes in breadth order of creation: node 0 is the root, the sons of a node
    are consecutive from `first_son`. perm: cluster position -> original index."""
    perm: np.ndarray       # (n,) int32
    offset: np.ndarray     # per node
    size: np.ndarray
    depth: np.ndarray
    rank: np.ndarray
    nb_sons: np.ndarray
    first_son: np.ndarray  # -1 for leaves
    ctr: np.ndarray        # (nodes, 3)
    rad: np.ndarray

    def sons(self, i: int):
        return range(self.first_son[i], self.first_son[i] + self.nb_sons[i]) if self.nb_sons[i] else range(0)

    def preorder(self) -> list[int]:
        """Node indices depth first, sons in order (the order save_cluster walks)."""
        out, stack = [], [0]
        while stack:
            i = stack.pop()
            out.append(i)
            stack.extend(reversed(list(self.sons(i))))
        return out


def cluster_tree(points, minclustersize: int = 10) -> ClusterTree:
    """The PCA cluster tree of a point set ((n, 3) array), as the builds make it
    (Cluster<PCAGeometricClustering>::build, cluster.hpp:57-92)."""
    x = np.ascontiguousarray(points, dtype=np.float64)
    n = x.shape[0]
    L = capi.lib()
    cnt = C.c_int(0)
    perm = np.zeros(n, np.int32)
    capi.check(L.htg_cluster_build(n, capi.ptr(x), int(minclustersize), capi.ptr(perm), C.byref(cnt), None, None))
    ni = np.zeros((cnt.value, 6), np.int32)
    nd = np.zeros((cnt.value, 4), np.float64)
    capi.check(L.htg_cluster_build(n, capi.ptr(x), int(minclustersize), None, C.byref(cnt), capi.ptr(ni),
                                   capi.ptr(nd)))
    return ClusterTree(perm, ni[:, 0].copy(), ni[:, 1].copy(), ni[:, 2].copy(), ni[:, 3].copy(), ni[:, 4].copy(),
                       ni[:, 5].copy(), nd[:, :3].copy(), nd[:, 3].copy())


def _num(v) -> str:
    """A number as a C++ output stream prints it by default (precision 6, %g)."""
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return format(float(v), ".6g")


def save_cluster(tree: ClusterTree, filename: str) -> None:
    """Cluster::save_cluster: `<filename>_permutation.csv` and `<filename>_tree.csv`."""
    with open(filename + "_permutation.csv", "w") as f:
        f.write(",".join(str(int(p)) for p in tree.perm))
    levels: list[list[str]] = [[] for _ in range(int(tree.depth.max()) + 1)]
    stack = [0]
    while stack:  # depth first, sons in order
        i = stack.pop()
        info = [_num(tree.nb_sons[i]), _num(tree.rank[i]), _num(tree.offset[i]), _num(tree.size[i]),
                _num(tree.rad[i])] + [_num(c) for c in tree.ctr[i]]
        levels[tree.depth[i]].append("|".join(info))
        stack.extend(reversed(list(tree.sons(i))))
    with open(filename + "_tree.csv", "w") as f:
        for lv in levels:
            f.write(",".join(lv) + "\n")


def read_cluster(file_permutation: str, file_tree: str) -> ClusterTree:
    """Cluster::read_cluster: the tree as written (radii and centres at the file's
    precision); nodes numbered depth first, the sons of a node consecutive."""
    with open(file_permutation) as f:
        perm = np.array([int(float(v)) for v in f.readline().strip().split(",")], dtype=np.int32)
    with open(file_tree) as f:
        levels = [line.rstrip("\n").split(",") for line in f if line.strip()]
    fields = [[s.split("|") for s in lv] for lv in levels]
    off, size, depth, rank, nsons, first, ctr, rad = [], [], [], [], [], [], [], []
    used = [0] * (len(fields) + 1)  # nodes consumed per depth

    def add(d, k):
        v = fields[d][k]
        off.append(int(v[2])), size.append(int(v[3])), depth.append(d), rank.append(int(v[1]))
        nsons.append(int(v[0])), first.append(-1), rad.append(float(v[4]))
        ctr.append([float(c) for c in v[5:8]])
        return len(off) - 1

    root = add(0, 0)
    used[0] = 1
    # the sons of a node are the next nb_sons entries of the next level, parents
    # taken depth first (read_cluster's counter_offset)
    stack = [root]
    while stack:
        i = stack.pop()
        d = depth[i]
        kids = []
        for _ in range(nsons[i]):
            kids.append(add(d + 1, used[d + 1]))
            used[d + 1] += 1
        if kids:
            first[i] = kids[0]
        stack.extend(reversed(kids))
    return ClusterTree(perm, np.array(off, np.int32), np.array(size, np.int32), np.array(depth, np.int32),
                       np.array(rank, np.int32), np.array(nsons, np.int32), np.array(first, np.int32),
                       np.array(ctr, np.float64).reshape(-1, 3), np.array(rad, np.float64))


# -- dense matrices and vectors ---------------------------------------------------------
def matrix_to_bytes(A, file: str) -> None:
    A = np.asarray(A)
    with open(file, "wb") as f:
        f.write(np.array([A.shape[0], A.shape[1]], dtype=np.int32).tobytes())
        f.write(np.asfortranarray(A).tobytes(order="F"))


def bytes_to_matrix(file: str, dtype=np.float64) -> np.ndarray:
    with open(file, "rb") as f:
        rows, cols = np.frombuffer(f.read(8), dtype=np.int32)
        data = np.frombuffer(f.read(), dtype=dtype, count=int(rows) * int(cols))
    return data.reshape((int(rows), int(cols)), order="F").copy()


def vector_to_bytes(v, file: str) -> None:
    v = np.ascontiguousarray(v)
    with open(file, "wb") as f:
        f.write(np.array([v.size], dtype=np.int32).tobytes())
        f.write(v.tobytes())


def bytes_to_vector(file: str, dtype=np.float64) -> np.ndarray:
    with open(file, "rb") as f:
        size = int(np.frombuffer(f.read(4), dtype=np.int32)[0])
        return np.frombuffer(f.read(), dtype=dtype, count=size).copy()
