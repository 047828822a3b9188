"""ctypes binding of oracle/liboracle.so (see oracle/oracle.h). Test infrastructure."""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_VP, _I, _D = C.c_void_p, C.c_int, C.c_double
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built: run `make -C oracle oracle`")
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_kernel_block": (None, [_I, _D, _D, _I, _VP, _VP, _I, _I, _VP, _VP, _VP]),
            "orc_cluster_build": (_I, [_I, _VP, _I, _VP, _VP, _VP]),
            "orc_aca": (_I, [_I, _D, _I, _I, _VP, _VP, _I, _D, _D, _I, _VP, _VP, _VP, _I, _VP, _I, _I, _I, _VP, _VP]),
            "orc_hmat_build": (_VP, [_I, _VP, _I, _VP, _I, _D, _D, _I, _D, _D, _I, _I, _D, _I]),
            "orc_hmat_nleaves": (_I, [_VP]),
            "orc_hmat_leaves": (None, [_VP, _VP]),
            "orc_hmat_data_size": (C.c_longlong, [_VP]),
            "orc_hmat_data": (None, [_VP, _VP]),
            "orc_hmat_perms": (None, [_VP, _VP, _VP]),
            "orc_hmat_false_positives": (_I, [_VP]),
            "orc_hmat_mvprod": (None, [_VP, _VP, _VP, _I]),
            "orc_hmat_free": (None, [_VP]),
            "orc_leaf_mvprod": (None, [_I, _I, _VP, _VP, _I, _VP, _VP, _VP, _VP, _I]),
            "orc_dense_rows": (None, [_I, _D, _D, _I, _VP, _VP, _I, _I, _VP, _VP, _I, _VP]),
            "orc_mlp_loss_grad": (None, [_VP, _VP, _VP, _VP, _VP, _I, _VP, _VP]),
            "orc_hmat_is_complex": (_I, [_VP]),
            "orc_leaf_mvprod_c": (None, [_I, _I, _VP, _VP, _I, _VP, _VP, _VP, _VP, _I]),
            "orc_leaf_mvprod_sym": (None, [_I, _VP, _I, _VP, _VP, _VP, _VP, _I, _I, _I]),
            "orc_dense_rows_c": (None, [_I, _D, _D, _I, _VP, _VP, _I, _I, _VP, _VP, _I, _VP]),
            "orc_mlp_predict": (None, [_VP, _VP, _VP, _VP, _I, _VP]),
            "orc_set_ref_int_overflow": (None, [_I]),
            "orc_set_min_depths": (None, [_I, _I]),
            "orc_leaf_mvprod_herm": (None, [_I, _VP, _I, _VP, _VP, _VP, _VP, _I, _I]),
            "orc_aca_size_bound": (C.c_longlong, [_I, _I]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class ref_int_overflow:
    """Context manager: inside it, the oracle's ACA evaluates q (M + N) > M N in
    the reference's 32-bit int arithmetic (partialACA.hpp:74), outside in exact
    64-bit arithmetic."""

    def __init__(self, on=True):
        self.on = bool(on)

    def __enter__(self):
        lib().orc_set_ref_int_overflow(int(self.on))
        return self

    def __exit__(self, *exc):
        lib().orc_set_ref_int_overflow(0)
        return False


class min_depths:
    """Context manager: block trees of builds inside it use set_mintargetdepth /
    set_minsourcedepth (hmatrix.hpp:215-216)."""

    def __init__(self, mintargetdepth=0, minsourcedepth=0):
        self.d = (int(mintargetdepth), int(minsourcedepth))

    def __enter__(self):
        lib().orc_set_min_depths(*self.d)
        return self

    def __exit__(self, *exc):
        lib().orc_set_min_depths(0, 0)
        return False


def aca_size_bound(M, N):
    """Right side of the q (M + N) > M N test under the current arithmetic."""
    return int(lib().orc_aca_size_bound(int(M), int(N)))


def kernel_block(kernel, delta, diag, xt, xs, rows, cols, same=False):
    xt = np.ascontiguousarray(xt, np.float64)
    xs = xt if same else np.ascontiguousarray(xs, np.float64)
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    out = np.zeros(len(rows) * len(cols))
    lib().orc_kernel_block(kernel, delta, diag, int(same), _p(xt), _p(xs), len(rows), len(cols), _p(rows), _p(cols),
                           _p(out))
    return out.reshape(len(cols), len(rows)).T


@dataclass
class ClusterTree:
    perm: np.ndarray
    nodes: np.ndarray   # (n, 6): offset, size, depth, rank, nb_sons, first_son
    geom: np.ndarray    # (n, 4): ctr0..2, rad


def cluster_build(x, minclustersize=10) -> ClusterTree:
    x = np.ascontiguousarray(x, np.float64)
    n = x.shape[0]
    perm = np.zeros(n, np.int32)
    ni = np.zeros((2 * n + 2, 6), np.int32)
    nd = np.zeros((2 * n + 2, 4))
    cnt = lib().orc_cluster_build(n, _p(x), minclustersize, _p(perm), _p(ni), _p(nd))
    return ClusterTree(perm, ni[:cnt].copy(), nd[:cnt].copy())


def aca(xt, xs, rows, cols, *, kernel=0, delta=0.0, diag=0.0, same=False, eps=1e-6, sym=False, ctr_t=None,
        ctr_s=None, off_t=0, off_s=0, reqrank=-1, cap=None):
    xt = np.ascontiguousarray(xt, np.float64)
    xs = xt if same else np.ascontiguousarray(xs, np.float64)
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    M, N = len(rows), len(cols)
    cap = cap or (min(M, N) + 1)
    U = np.zeros(M * cap)
    V = np.zeros(N * cap)
    ct = np.ascontiguousarray(ctr_t, np.float64)
    cs = np.ascontiguousarray(ctr_s if ctr_s is not None else ctr_t, np.float64)
    k = lib().orc_aca(int(sym), eps, M, N, _p(rows), _p(cols), kernel, delta, diag, int(same), _p(xt), _p(xs),
                      _p(ct), off_t, _p(cs), off_s, reqrank, cap, _p(U), _p(V))
    if k <= 0:
        return k, None, None
    return k, U[:M * k].reshape(k, M).T.copy(), V[:N * k].reshape(N, k).T.copy()


@dataclass
class OracleHMatrix:
    nr: int
    nc: int
    perm_t: np.ndarray
    perm_s: np.ndarray
    leaves: np.ndarray
    data: np.ndarray
    false_positive: int
    uplo: str = "N"
    symmetry: str = "S"  # with uplo 'L' / 'U': 'S' symmetric or 'H' Hermitian storage

    def mvprod(self, x):
        if self.uplo != "N" and self.symmetry == "H" and np.iscomplexobj(self.data):
            return leaf_mvprod_herm(self.nr, self.perm_t, self.leaves, self.data, x, self.uplo)
        if self.uplo != "N":
            return leaf_mvprod_sym(self.nr, self.perm_t, self.leaves, self.data, x, self.uplo)
        return leaf_mvprod(self.nr, self.nc, self.perm_t, self.perm_s, self.leaves, self.data, x)


def hmat_build(xt, xs=None, *, kernel=0, delta=0.0, diag=0.0, minclustersize=10, eta=10.0, eps=1e-6, sym=False,
               reqrank=-1, maxblocksize=1e6, wavenumber=0.0, uplo="N", symmetry="S") -> OracleHMatrix:
    """Complex kernel families (kernel 4, 5, 7): data is complex128 and mvprod takes
    complex vectors; the Helmholtz wavenumber travels in `delta`. uplo 'L' / 'U':
    symmetric storage (HMatrix(..., symmetry, uplo), one point set): the leaves of
    one triangle (the same for 'S' and 'H'), and a product that also applies
    their transposes ('S') or conjugate transposes ('H')."""
    if uplo != "N" and xs is not None:
        raise ValueError("symmetric storage needs one point set")
    if kernel == 4 and wavenumber:
        delta = wavenumber
    L = lib()
    xt = np.ascontiguousarray(xt, np.float64)
    xs_ = None if xs is None else np.ascontiguousarray(xs, np.float64)
    nt = xt.shape[0]
    ns = nt if xs is None else xs_.shape[0]
    h = L.orc_hmat_build(nt, _p(xt), ns, _p(xs_), kernel, delta, diag, minclustersize, eta, eps, int(sym), reqrank,
                         maxblocksize, ord(uplo))
    try:
        n = L.orc_hmat_nleaves(h)
        leaves = np.zeros((n, 5), np.int32)
        L.orc_hmat_leaves(h, _p(leaves))
        cplx = L.orc_hmat_is_complex(h) != 0
        data = np.zeros(L.orc_hmat_data_size(h) * (2 if cplx else 1))
        L.orc_hmat_data(h, _p(data))
        if cplx:
            data = data.view(np.complex128)
        pt = np.zeros(nt, np.int32)
        ps = np.zeros(ns, np.int32)
        L.orc_hmat_perms(h, _p(pt), _p(ps))
        fp = L.orc_hmat_false_positives(h)
    finally:
        L.orc_hmat_free(h)
    return OracleHMatrix(nt, ns, pt, ps, leaves, data, fp, uplo, symmetry)


def leaf_mvprod(nr, nc, perm_t, perm_s, leaves, data, x):
    cplx = np.iscomplexobj(data)
    dt = np.complex128 if cplx else np.float64
    x = np.asarray(x, dt)
    vec = x.ndim == 1
    xm = np.asfortranarray(x.reshape(nc, -1))
    mu = xm.shape[1]
    y = np.zeros((nr, mu), dtype=dt, order="F")
    leaves = np.ascontiguousarray(leaves, np.int32)
    data = np.ascontiguousarray(data, dt)
    perm_t = np.ascontiguousarray(perm_t, np.int32)  # kept alive across the call
    perm_s = np.ascontiguousarray(perm_s, np.int32)
    fn = lib().orc_leaf_mvprod_c if cplx else lib().orc_leaf_mvprod
    fn(nr, nc, _p(perm_t), _p(perm_s), leaves.shape[0], _p(leaves), _p(data), _p(xm), _p(y), mu)
    return y[:, 0].copy() if vec else np.ascontiguousarray(y)


def leaf_mvprod_sym(n, perm, leaves, data, x, uplo):
    """Product of a symmetric-storage leaf list (one triangle): kept leaves, the
    transposes of the off-diagonal ones, symv on the diagonal leaves."""
    cplx = np.iscomplexobj(data)
    dt = np.complex128 if cplx else np.float64
    x = np.asarray(x, dt)
    vec = x.ndim == 1
    xm = np.asfortranarray(x.reshape(n, -1))
    mu = xm.shape[1]
    y = np.zeros((n, mu), dtype=dt, order="F")
    leaves = np.ascontiguousarray(leaves, np.int32)
    data = np.ascontiguousarray(data, dt)
    perm = np.ascontiguousarray(perm, np.int32)
    lib().orc_leaf_mvprod_sym(n, _p(perm), leaves.shape[0], _p(leaves), _p(data), _p(xm), _p(y), mu, ord(uplo),
                              int(cplx))
    return y[:, 0].copy() if vec else np.ascontiguousarray(y)


def leaf_mvprod_herm(n, perm, leaves, data, x, uplo):
    """Product of a Hermitian-storage leaf list (one triangle, complex): kept
    leaves, conjugate transposes of the off-diagonal ones, hemv on the
    diagonal leaves."""
    x = np.asarray(x, np.complex128)
    vec = x.ndim == 1
    xm = np.asfortranarray(x.reshape(n, -1))
    mu = xm.shape[1]
    y = np.zeros((n, mu), dtype=np.complex128, order="F")
    leaves = np.ascontiguousarray(leaves, np.int32)
    data = np.ascontiguousarray(data, np.complex128)
    perm = np.ascontiguousarray(perm, np.int32)
    lib().orc_leaf_mvprod_herm(n, _p(perm), leaves.shape[0], _p(leaves), _p(data), _p(xm), _p(y), mu, ord(uplo))
    return y[:, 0].copy() if vec else np.ascontiguousarray(y)


def dense_rows_c(xt, xs, rows, x, *, kernel=4, wavenumber=0.0, diag=0.0, same=False):
    """Brute-force complex dense rows (kernel 4 or 5)."""
    xt = np.ascontiguousarray(xt, np.float64)
    xs = xt if same else np.ascontiguousarray(xs, np.float64)
    rows = np.ascontiguousarray(rows, np.int32)
    x = np.asarray(x, np.complex128)
    vec = x.ndim == 1
    xm = np.asfortranarray(x.reshape(xs.shape[0], -1))
    mu = xm.shape[1]
    out = np.zeros((len(rows), mu), dtype=np.complex128, order="F")
    lib().orc_dense_rows_c(kernel, wavenumber, diag, int(same), _p(xt), _p(xs), xs.shape[0], len(rows), _p(rows),
                           _p(xm), mu, _p(out))
    return out[:, 0].copy() if vec else np.ascontiguousarray(out)


def dense_rows(xt, xs, rows, x, *, kernel=0, delta=0.0, diag=0.0, same=False):
    xt = np.ascontiguousarray(xt, np.float64)
    xs = xt if same else np.ascontiguousarray(xs, np.float64)
    rows = np.ascontiguousarray(rows, np.int32)
    x = np.asarray(x, np.float64)
    vec = x.ndim == 1
    xm = np.asfortranarray(x.reshape(xs.shape[0], -1))
    mu = xm.shape[1]
    out = np.zeros((len(rows), mu), order="F")
    lib().orc_dense_rows(kernel, delta, diag, int(same), _p(xt), _p(xs), xs.shape[0], len(rows), _p(rows), _p(xm),
                         mu, _p(out))
    return out[:, 0].copy() if vec else np.ascontiguousarray(out)


def iter_leaf_data(leaves, data):
    """Yields (off_i, m, off_j, n, k, A or (U, V)) in leaf order."""
    pos = 0
    for oi, m, oj, n, k in leaves:
        if k < 0:
            yield oi, m, oj, n, k, data[pos:pos + m * n].reshape(n, m).T
            pos += m * n
        else:
            u = data[pos:pos + m * k].reshape(k, m).T
            pos += m * k
            v = data[pos:pos + k * n].reshape(n, k).T
            pos += k * n
            yield oi, m, oj, n, k, (u, v)


MLP_NPARAMS = 4801


def mlp_loss_grad(params, mean, std, feat, target):
    """FP64 loss and gradient of the rank predictor network (oracle/oracle.h)."""
    params = np.ascontiguousarray(params, dtype=np.float64)
    mean = np.ascontiguousarray(mean, dtype=np.float64)
    std = np.ascontiguousarray(std, dtype=np.float64)
    feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1, 8)
    target = np.ascontiguousarray(target, dtype=np.float32)
    loss = C.c_double()
    grad = np.zeros(MLP_NPARAMS)
    lib().orc_mlp_loss_grad(_p(params), _p(mean), _p(std), _p(feat), _p(target), feat.shape[0], C.byref(loss),
                            _p(grad))
    return loss.value, grad


def mlp_predict(params, mean, std, feat):
    params = np.ascontiguousarray(params, dtype=np.float64)
    mean = np.ascontiguousarray(mean, dtype=np.float64)  # kept alive across the call
    std = np.ascontiguousarray(std, dtype=np.float64)
    feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1, 8)
    out = np.zeros(feat.shape[0])
    lib().orc_mlp_predict(_p(params), _p(mean), _p(std), _p(feat), feat.shape[0], _p(out))
    return out
