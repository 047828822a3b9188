"""GPU: the multi-workgroup ACA of large blocks against the one-workgroup ACA
and the oracle. With `aca_wide_min` lowered, most admissible blocks of a small
problem take the multi-workgroup path (residual rows and columns cut into
2048-entry chunks, several chunks per vector at these sizes); the leaf list, the
ranks and the factors must be those of the default build (residual entries are
the same sums in the same order; only the Frobenius estimate is summed in
another order) and of the oracle's restatement of partialACA.hpp:39-171."""
import numpy as np
import pytest

from htool_amd import geometry as geo

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def build(htg, pts, gen, comp, wide_min, eps=1e-6, mcs=32):
    H = htg.HMatrix(epsilon=eps, eta=10.0, minclustersize=mcs)
    H.set_compression(comp)
    H.set_aca_wide_min(wide_min)
    H.build(gen, pts)
    return H


def compare(orc, H0, H1):
    l0, d0 = H0.get_leaves(with_data=True)
    l1, d1 = H1.get_leaves(with_data=True)
    assert np.array_equal(l0, l1), "leaf lists or ranks differ"
    worst = 0.0
    for (_, _, _, _, k, a), (_, _, _, _, _, b) in zip(orc.iter_leaf_data(l0, d0), orc.iter_leaf_data(l1, d1)):
        if k > 0:
            worst = max(worst, np.abs(a[0] - b[0]).max() / np.abs(a[0]).max(),
                        np.abs(a[1] - b[1]).max() / np.abs(a[1]).max())
    return worst


@pytest.mark.parametrize("comp", ["partialACA", "sympartialACA"])
def test_wide_aca_laplace_equals_one_workgroup(htg, orc, comp):
    print(f"wide aca laplace {comp}", flush=True)
    n = 20000
    pts = geo.sphere(n, seed=21)
    gen = htg.Generator(geo.KERNEL_LAPLACE, 0.0, geo.laplace_diag(geo.sphere_spacing(n)))
    H0 = build(htg, pts, gen, comp, -1)
    H1 = build(htg, pts, gen, comp, 400)
    print("wide blocks", H1.get_info().n_aca_wide, "of", H1.get_info().n_aca, flush=True)
    assert H0.get_info().n_aca_wide == 0
    assert H1.get_info().n_aca_wide > 100
    assert compare(orc, H0, H1) < 1e-13
    x = np.random.default_rng(3).standard_normal(n)
    assert rel(H1 @ x, H0 @ x) < 1e-13


def test_wide_aca_matches_oracle(htg, orc):
    print("wide aca oracle", flush=True)
    xt = geo.cylinder(6000, seed=5)
    gen = htg.Generator(geo.KERNEL_INV_DELTA_R, 1e-5, 0.0)
    H = build(htg, xt, gen, "partialACA", 300, mcs=10)
    print("wide blocks", H.get_info().n_aca_wide, "of", H.get_info().n_aca, flush=True)
    assert H.get_info().n_aca_wide > 10
    O = orc.hmat_build(xt, kernel=geo.KERNEL_INV_DELTA_R, delta=1e-5, minclustersize=10, eta=10.0, eps=1e-6)
    leaves, data = H.get_leaves(with_data=True)
    assert np.array_equal(leaves, O.leaves)
    x = np.random.default_rng(4).uniform(-1, 1, 6000)
    assert rel(H @ x, O.mvprod(x)) < 1e-12


def test_wide_aca_helmholtz_equals_one_workgroup(htg, orc):
    print("wide aca helmholtz", flush=True)
    n = 12000
    pts = geo.ellipsoid(n, (1.0, 1.0, 4.0), seed=7)
    h = np.sqrt(geo.prolate_spheroid_area(1.0, 4.0) / n)
    gen = htg.Generator(geo.KERNEL_HELMHOLTZ, 0.0, geo.laplace_diag(h), 2 * np.pi / (10 * h))
    H0 = build(htg, pts, gen, "partialACA", -1, eps=1e-5)
    H1 = build(htg, pts, gen, "partialACA", 600, eps=1e-5)
    print("wide blocks", H1.get_info().n_aca_wide, "of", H1.get_info().n_aca, flush=True)
    assert H1.get_info().n_aca_wide > 10
    assert compare(orc, H0, H1) < 1e-13
