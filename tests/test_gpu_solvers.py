"""GPU: Krylov solvers with device-resident vectors and the GPU H-matrix
product as operator (htool_amd/solvers.py; the solver-loop caller of the
product, HPDDM's GMV through mymvprod_local_to_local,
include/htool/wrappers/wrapper_hpddm.hpp:41-83). Solutions are checked against
numpy's dense solve of the same kernel matrix (closed form), to the accuracy the
H-matrix approximation allows. The diagonal is the kernel's row-sum scale
N/(4 pi) (second-kind-like systems): with the bare 1/(4 pi h) diagonal and
random points, near pairs make the collocation matrix too ill-conditioned to
say anything about the solver."""
import math

import numpy as np
import pytest

from htool_amd import geometry as geo
from htool_amd import solvers

pytestmark = pytest.mark.gpu


def dense_laplace(x, diag):
    r = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)
    np.fill_diagonal(r, 1.0)
    A = 1.0 / (4 * math.pi * r)
    np.fill_diagonal(A, diag)
    return A


def dense_helmholtz(x, k, diag):
    r = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)
    np.fill_diagonal(r, 1.0)
    A = np.exp(1j * k * r) / (4 * math.pi * r)
    np.fill_diagonal(A, diag)
    return A


@pytest.mark.parametrize("method", ["gmres", "cg"])
def test_laplace_solve_matches_dense(htg, method):
    print(f"solver {method}", flush=True)
    n = 3000
    x = geo.sphere(n, seed=51)
    diag = n / (4 * math.pi)
    H = htg.HMatrix(epsilon=1e-9, eta=10.0, minclustersize=32)
    H.set_compression("partialACA")
    H.build(htg.Generator(geo.KERNEL_LAPLACE, 0.0, diag), x)
    b = np.random.default_rng(1).standard_normal(n)
    sol, info = (solvers.gmres(H, b, tol=1e-10, restart=40) if method == "gmres" else solvers.cg(H, b, tol=1e-10))
    assert info["converged"] and info["residual"] < 1e-10, info
    ref = np.linalg.solve(dense_laplace(x, diag), b)
    assert np.linalg.norm(sol - ref) / np.linalg.norm(ref) < 1e-6


def test_helmholtz_gmres_matches_dense(htg):
    n = 2500
    x = geo.sphere(n, seed=52)
    h = geo.sphere_spacing(n)
    k = 2 * math.pi / (10 * h)
    diag = n / (4 * math.pi)
    H = htg.HMatrix(epsilon=1e-9, eta=10.0, minclustersize=32)
    H.set_compression("partialACA")
    H.build(htg.Generator(geo.KERNEL_HELMHOLTZ, 0.0, diag, k), x)
    rng = np.random.default_rng(2)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    sol, info = solvers.gmres(H, b, tol=1e-10, restart=60, maxiter=2000)
    assert info["converged"], info
    ref = np.linalg.solve(dense_helmholtz(x, k, diag), b)
    assert np.linalg.norm(sol - ref) / np.linalg.norm(ref) < 1e-5
