"""Krylov solvers on device-resident vectors, with the GPU H-matrix product as
the operator: the caller side of the hot path in a solver loop (the role of
HPDDM's GMV, which applies the H-matrix through mymvprod_local_to_local,
include/htool/wrappers/wrapper_hpddm.hpp:41-83).

Every Krylov vector lives on the GPU (torch tensors); each iteration is one
`HMatrix.mvprod_device` on the solver's stream plus a few BLAS-1/2 calls. Only
the small Hessenberg (GMRES) or the step scalars (CG) come to the host. The
right-hand side is copied to the device once and the solution back once.
"""
from __future__ import annotations

import numpy as np


class _Op:
    """y = H x on the solver's device. An HMatrix runs its GPU product on the
    current stream; any other operator object supplies `apply(x, y)` and a
    `device` (used to check the Krylov recurrences against dense algebra)."""

    def __init__(self, H, dtype, torch):
        self.H, self.torch = H, torch
        self.n = H.nb_rows()
        if H.nb_cols() != self.n:
            raise ValueError("the solvers need a square H-matrix")
        self.device = getattr(H, "device", "cuda")
        if self.device == "cuda" and not torch.cuda.is_available():
            raise RuntimeError("the solvers need a GPU (HIP device)")
        self.stream = torch.cuda.current_stream().cuda_stream if self.device == "cuda" else None
        self.dtype = dtype

    def __call__(self, x, y):
        if hasattr(self.H, "apply"):
            self.H.apply(x, y)
        else:
            self.H.mvprod_device(x.data_ptr(), y.data_ptr(), 1, self.stream)


def _setup(H, b, x0):
    import torch
    dtype = torch.complex128 if H.is_complex else torch.float64
    op = _Op(H, dtype, torch)
    bd = (b if torch.is_tensor(b) else torch.as_tensor(np.asarray(b))).to(op.device, dtype)
    xd = torch.zeros_like(bd) if x0 is None else torch.as_tensor(np.asarray(x0)).to(op.device, dtype)
    return torch, op, bd.contiguous(), xd.contiguous()


def gmres(H, b, x0=None, *, tol: float = 1e-8, restart: int = 50, maxiter: int = 1000):
    """Restarted GMRES(restart) with classical Gram-Schmidt applied twice.
    Returns (x as a numpy array, {"iterations", "residual", "converged"}) with
    residual = ||b - H x|| / ||b|| of the final iterate (recomputed)."""
    torch, op, bd, x = _setup(H, b, x0)
    n = op.n
    bnorm = float(torch.linalg.vector_norm(bd))
    if bnorm == 0:
        return np.zeros(n, dtype=np.complex128 if H.is_complex else np.float64), \
            {"iterations": 0, "residual": 0.0, "converged": True}
    m = int(restart)
    V = torch.zeros((m + 1, n), dtype=op.dtype, device=op.device)
    w = torch.empty_like(bd)
    r = torch.empty_like(bd)
    it = 0
    converged = False
    cplx = H.is_complex
    while it < maxiter:
        op(x, r)
        r = bd - r
        beta = float(torch.linalg.vector_norm(r))
        if beta / bnorm <= tol:
            converged = True
            break
        V[0] = r / beta
        Hs = np.zeros((m + 1, m), dtype=np.complex128 if cplx else np.float64)
        cs = np.zeros(m, dtype=Hs.dtype)
        sn = np.zeros(m, dtype=Hs.dtype)
        g = np.zeros(m + 1, dtype=Hs.dtype)
        g[0] = beta
        k = 0
        for j in range(m):
            op(V[j], w)
            h = torch.zeros(j + 1, dtype=op.dtype, device=op.device)
            for _ in range(2):  # CGS2
                c = V[:j + 1].conj() @ w
                w = w - c @ V[:j + 1]
                h = h + c
            hn = float(torch.linalg.vector_norm(w))
            col = np.zeros(m + 1, dtype=Hs.dtype)
            col[:j + 1] = h.cpu().numpy()
            col[j + 1] = hn
            for i in range(j):  # earlier rotations
                a, bb = col[i], col[i + 1]
                col[i] = np.conj(cs[i]) * a + np.conj(sn[i]) * bb
                col[i + 1] = -sn[i] * a + cs[i] * bb
            den = np.hypot(abs(col[j]), abs(col[j + 1]))
            cs[j] = col[j] / den if den else 1.0
            sn[j] = col[j + 1] / den if den else 0.0
            col[j] = den
            col[j + 1] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = np.conj(cs[j]) * g[j]
            Hs[:, j] = col
            it += 1
            k = j + 1
            if hn > 0 and j + 1 < m + 1:
                V[j + 1] = w / hn
            if abs(g[j + 1]) / bnorm <= tol or hn == 0 or it >= maxiter:
                break
        y = np.linalg.solve(np.triu(Hs[:k, :k]), g[:k])
        x = x + torch.as_tensor(y, dtype=op.dtype, device=op.device) @ V[:k]
    op(x, r)
    res = float(torch.linalg.vector_norm(bd - r)) / bnorm
    return x.cpu().numpy(), {"iterations": it, "residual": res, "converged": converged or res <= tol}


def cg(H, b, x0=None, *, tol: float = 1e-8, maxiter: int = 1000):
    """Conjugate gradients for a symmetric (Hermitian) positive definite H.
    Returns (x, {"iterations", "residual", "converged"})."""
    torch, op, bd, x = _setup(H, b, x0)
    r = torch.empty_like(bd)
    op(x, r)
    r = bd - r
    p = r.clone()
    q = torch.empty_like(bd)
    bnorm = float(torch.linalg.vector_norm(bd))
    rr = float(torch.vdot(r, r).real)
    it = 0
    while it < maxiter and np.sqrt(rr) / max(bnorm, 1e-300) > tol:
        op(p, q)
        alpha = rr / float(torch.vdot(p, q).real)
        x = x + alpha * p
        r = r - alpha * q
        rr_new = float(torch.vdot(r, r).real)
        p = r + (rr_new / rr) * p
        rr = rr_new
        it += 1
    op(x, q)
    res = float(torch.linalg.vector_norm(bd - q)) / max(bnorm, 1e-300)
    return x.cpu().numpy(), {"iterations": it, "residual": res, "converged": res <= tol}
