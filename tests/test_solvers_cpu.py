"""CPU: the Krylov recurrences of htool_amd/solvers.py against numpy's dense
solve, with a dense operator standing in for the GPU product (the same code
path apart from the product itself)."""
import numpy as np
import pytest
import torch

from htool_amd import solvers


class DenseOp:
    device = "cpu"

    def __init__(self, A):
        self.A = torch.as_tensor(A)
        self.is_complex = np.iscomplexobj(A)

    def nb_rows(self): return self.A.shape[0]
    def nb_cols(self): return self.A.shape[1]

    def apply(self, x, y):
        y.copy_(self.A @ x)


def spd(n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (Q * np.linspace(1.0, 50.0, n)) @ Q.T


@pytest.mark.parametrize("complex_", [False, True])
def test_gmres_matches_dense_solve(complex_):
    rng = np.random.default_rng(3)
    n = 120
    A = np.eye(n) * 4 + rng.standard_normal((n, n)) / np.sqrt(n)
    if complex_:
        A = A + 1j * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n) + (1j * rng.standard_normal(n) if complex_ else 0)
    x, info = solvers.gmres(DenseOp(A), b, tol=1e-12, restart=10, maxiter=500)
    assert info["converged"], info
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(x) < 1e-10


def test_cg_matches_dense_solve():
    A = spd(150, 4)
    b = np.random.default_rng(5).standard_normal(150)
    x, info = solvers.cg(DenseOp(A), b, tol=1e-12, maxiter=1000)
    assert info["converged"], info
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(x) < 1e-10
