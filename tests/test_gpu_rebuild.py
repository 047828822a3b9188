"""GPU: builds in one process share the device's ACA workspace and visited
flags (kept between builds, grown on demand, never cleared wholesale). A build
that follows other builds — larger ones, complex entries, the multi-workgroup
ACA — must give bit for bit the leaves, ranks, factors and product of the first
build of the same problem."""
import numpy as np
import pytest

from htool_amd import geometry as geo

pytestmark = pytest.mark.gpu


def build(htg, pts, gen, wide_min):
    H = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=32)
    H.set_compression("partialACA")
    H.set_aca_wide_min(wide_min)
    H.build(gen, pts)
    return H


def snapshot(H, x):
    leaves, data = H.get_leaves(with_data=True)
    return np.array(leaves, copy=True), np.array(data, copy=True), H @ x


def laplace(htg, n):
    return htg.Generator(geo.KERNEL_LAPLACE, 0.0, geo.laplace_diag(geo.sphere_spacing(n)))


def test_build_after_other_builds_is_identical(htg):
    print("rebuild identical", flush=True)
    n = 8000
    pts = geo.sphere(n, seed=31)
    x = np.random.default_rng(4).standard_normal(n)
    wides = (-1, 400)  # one-workgroup ACA only; most blocks on the multi-workgroup ACA
    first = [snapshot(build(htg, pts, laplace(htg, n), w), x) for w in wides]

    # builds in between: a larger real one and a complex one, both with wide blocks
    nb = 20000
    build(htg, geo.sphere(nb, seed=32), laplace(htg, nb), 400)
    nc = 6000
    pc = geo.sphere(nc, seed=33)
    h = geo.sphere_spacing(nc)
    k = 2 * np.pi / (10 * h)
    hc = build(htg, pc, htg.Generator(geo.KERNEL_HELMHOLTZ, 0.0, geo.laplace_diag(h), k), 400)
    assert hc.get_info().n_aca_wide > 0

    for w, (l0, d0, y0) in zip(wides, first):
        l1, d1, y1 = snapshot(build(htg, pts, laplace(htg, n), w), x)
        assert np.array_equal(l0, l1), f"leaves or ranks differ (aca_wide_min {w})"
        assert np.array_equal(d0, d1), f"leaf data differ (aca_wide_min {w})"
        assert np.array_equal(y0, y1), f"products differ (aca_wide_min {w})"
