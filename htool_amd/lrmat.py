"""Batched leaf compression: the GPU counterpart of calling
VirtualLowRankGenerator::copy_low_rank_approximation once per block
(include/htool/lrmat/virtual_lrmat_generator.hpp:15, partialACA.hpp:39-171,
sympartialACA.hpp:38-209)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import capi
from .hmatrix import Generator


@dataclass
class Block:
    """One block to compress: original point indices of its rows and columns,
    and the target/source cluster centres and offsets the ACA uses for its first
    pivot and (sympartialACA) its orientation."""
    rows: np.ndarray
    cols: np.ndarray
    ctr_t: tuple
    ctr_s: tuple
    off_t: int = 0
    off_s: int = 0


def aca_batch(gen: Generator, blocks: list, xt: np.ndarray, xs: np.ndarray | None = None, *, epsilon: float = 1e-6,
              reqrank: int = -1, compressor: str = "partialACA", cap: int = 64, noalloc: bool = False):
    """Returns a list of (rank, U, V): U (m, k), V (k, n); rank -1 means the block is
    not compressible (the reference then stores it dense). noalloc: run through
    htg_aca_batch_noalloc (no device allocation; the library's 64 MiB scratch)."""
    L = capi.lib()
    xt = np.ascontiguousarray(xt, dtype=np.float64)
    same = xs is None
    xs_ = None if same else np.ascontiguousarray(xs, dtype=np.float64)
    nb = len(blocks)
    M = np.array([len(b.rows) for b in blocks], np.int32)
    N = np.array([len(b.cols) for b in blocks], np.int32)
    rows = np.ascontiguousarray(np.concatenate([np.asarray(b.rows, np.int32) for b in blocks]) if nb else
                                np.zeros(0, np.int32))
    cols = np.ascontiguousarray(np.concatenate([np.asarray(b.cols, np.int32) for b in blocks]) if nb else
                                np.zeros(0, np.int32))
    row_off = np.zeros(nb, np.int32)
    col_off = np.zeros(nb, np.int32)
    if nb:
        row_off[1:] = np.cumsum(M)[:-1]
        col_off[1:] = np.cumsum(N)[:-1]
    ctr_t = np.ascontiguousarray([b.ctr_t for b in blocks], dtype=np.float64).reshape(-1)
    ctr_s = np.ascontiguousarray([b.ctr_s for b in blocks], dtype=np.float64).reshape(-1)
    off_t = np.array([b.off_t for b in blocks], np.int32)
    off_s = np.array([b.off_s for b in blocks], np.int32)
    u_off = np.zeros(nb, np.int64)
    v_off = np.zeros(nb, np.int64)
    if nb:
        u_off[1:] = np.cumsum(M.astype(np.int64) * cap)[:-1]
        v_off[1:] = np.cumsum(N.astype(np.int64) * cap)[:-1]
    U = np.zeros(int((M.astype(np.int64) * cap).sum()) or 1)
    V = np.zeros(int((N.astype(np.int64) * cap).sum()) or 1)
    ranks = np.zeros(nb, np.int32)
    g = gen._c(same)
    fn = L.htg_aca_batch_noalloc if noalloc else L.htg_aca_batch
    capi.check(fn(C.byref(g), capi.SYM_PARTIAL_ACA if compressor == "sympartialACA" else
                               capi.PARTIAL_ACA, float(epsilon), int(reqrank), xt.shape[0], capi.ptr(xt),
                               xt.shape[0] if same else xs_.shape[0], capi.ptr(xs_), nb, capi.ptr(M), capi.ptr(N),
                               capi.ptr(rows), capi.ptr(row_off), capi.ptr(cols), capi.ptr(col_off), capi.ptr(ctr_t),
                               capi.ptr(ctr_s), capi.ptr(off_t), capi.ptr(off_s), int(cap), capi.ptr(ranks),
                               capi.ptr(U), capi.ptr(u_off), capi.ptr(V), capi.ptr(v_off)))
    out = []
    for b in range(nb):
        k = int(ranks[b])
        if k <= 0:
            out.append((k, None, None))
            continue
        m, n = int(M[b]), int(N[b])
        u = U[u_off[b]:u_off[b] + m * k].reshape(k, m).T.copy()
        v = V[v_off[b]:v_off[b] + k * n].reshape(n, k).T.copy()
        out.append((k, u, v))
    return out
