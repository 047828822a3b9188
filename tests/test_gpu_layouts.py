"""GPU: every product path against the FP64 product of the same H-matrix, over
a grid of geometries, admissibility, precision classes, storages, right-hand-side
counts and work decompositions.

The FP64 product itself is held to the oracle elsewhere (1e-12 on the same
leaves); here the other paths are held to it: the single-RHS V chunk lists
(whole panels, column chunks), the multi-RHS chunk list, row blocks whole and
in parts, the 32-wide matrix-core passes (FP32 and 16-bit) and the narrower
passes, with each storage's error budget. Columns of a multi-RHS product also
match single-vector products of the same storage.
"""
import numpy as np
import pytest

from htool_amd import geometry as geo

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


GEOMETRIES = {
    "sphere": lambda n: (geo.sphere(n, seed=21), geo.KERNEL_LAPLACE, 0.0, geo.laplace_diag(geo.sphere_spacing(n))),
    "spheroid": lambda n: (geo.ellipsoid(n, (1.0, 1.0, 4.0), seed=22), geo.KERNEL_LAPLACE, 0.0,
                           geo.laplace_diag(np.sqrt(geo.prolate_spheroid_area(1.0, 4.0) / n))),
    "cylinder": lambda n: (geo.cylinder(n, seed=23), geo.KERNEL_INV_DELTA_R, 1e-5, 0.0),
}
# storage: bound against FP64 with trailing terms in the 16-bit class (lp) and without
BOUND = {"fp32": 1e-6, "fp32_fp16": 1e-6, "fp32_bf16": 8e-6}
LAYOUTS = [dict(), dict(rb_max=16, vrows=8), dict(rb_max=64, vchunk_el=256, vgroup_el=256, vgroup_tiles=1),
           dict(vrows=32, vchunk_el=4096, vgroup_el=8192, vgroup_tiles=8)]


@pytest.mark.parametrize("geom,n,eta", [("sphere", 9000, 10.0), ("spheroid", 7000, 2.0), ("cylinder", 5000, 10.0)])
@pytest.mark.parametrize("lowprec", [False, True])
def test_all_paths_against_fp64(htg, geom, n, eta, lowprec):
    print(f"layouts {geom} n={n} eta={eta} lowprec={lowprec}", flush=True)
    x, kernel, delta, diag = GEOMETRIES[geom](n)
    H = htg.HMatrix(epsilon=1e-6, eta=eta, minclustersize=32)
    H.set_compression("partialACA")
    if lowprec:
        H.set_recompression(2e-6)
        H.set_lowprec(5e-7)
    H.build(htg.Generator(kernel, delta, diag), x)
    info = H.get_info()
    assert (info.lowprec_entries > 0) == lowprec
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, 40))
    H.set_storage("fp64")
    Y64 = H.mvprod_global_to_global(X)
    for li, lay in enumerate(LAYOUTS):
        if lay:
            H.set_layout(**lay)
        for storage, bound in BOUND.items():
            if storage != "fp32" and not lowprec:
                continue
            H.set_storage(storage)
            for mu in (1, 3, 32, 40):
                Y = H.mvprod_global_to_global(X[:, :mu])
                e = rel(Y, Y64[:, :mu])
                assert e < bound, (li, storage, mu, e)
            y1 = H.mvprod_global_to_global(X[:, 33])
            assert rel(Y[:, 33], y1) < 1e-6, (li, storage)
        H.set_storage("fp64")
        assert rel(H.mvprod_global_to_global(X[:, :3]), Y64[:, :3]) < 1e-12, li


@pytest.mark.parametrize("all_terms", [False, True])
def test_complex_paths_against_fp64(htg, all_terms):
    """Complex Helmholtz H-matrix: single-vector, narrow multi-RHS, 16-wide,
    32-wide and (FP32 arithmetic) 64-wide matrix-core passes (FP32, and
    FP16/BF16 on the 16-bit cores; 40 RHS = one 32-wide pass and an 8-wide one,
    72 = a 64-wide pass and an 8-wide one, or two 32-wide and an 8-wide in FP64)
    against the complex FP64 product, for three work decompositions; columns of
    a 72-RHS product match single-vector products of the same storage, FP64
    included."""
    print(f"complex layouts all_terms={all_terms}", flush=True)
    n = 6000
    x = geo.ellipsoid(n, (1.0, 1.0, 4.0), seed=31)
    h = np.sqrt(geo.prolate_spheroid_area(1.0, 4.0) / n)
    k = 2 * np.pi / (10 * h)
    H = htg.HMatrix(epsilon=1e-5, eta=10.0, minclustersize=32)
    H.set_compression("partialACA")
    H.set_lowprec(0.0 if all_terms else 2e-5, all_terms=all_terms)
    H.build(htg.Generator(geo.KERNEL_HELMHOLTZ, 0.0, geo.laplace_diag(h), k), x)
    assert H.get_info().lowprec_entries > 0
    rng = np.random.default_rng(32)
    X = rng.standard_normal((n, 72)) + 1j * rng.standard_normal((n, 72))
    H.set_storage("fp64")
    Y64 = H.mvprod_global_to_global(X)
    for c in (0, 17, 31, 33, 39, 63, 71):
        assert rel(Y64[:, c], H.mvprod_global_to_global(X[:, c])) < 1e-12, c
    bounds = {"fp32": 2e-6, "fp32_fp16": 2e-3 if all_terms else 2e-5, "fp32_bf16": 2e-2 if all_terms else 2e-4}
    for li, lay in enumerate([dict(), dict(vchunk_el=256, vgroup_el=256, vgroup_tiles=1), dict(vrows=32, rb_max=64)]):
        if lay:
            H.set_layout(**lay)
        for storage, bound in bounds.items():
            H.set_storage(storage)
            for mu in (1, 3, 16, 20, 32, 40, 64):
                e = rel(H.mvprod_global_to_global(X[:, :mu]), Y64[:, :mu])
                assert e < bound, (li, storage, mu, e)
            Y = H.mvprod_global_to_global(X)
            assert rel(Y, Y64) < bound, (li, storage)
            for c in (0, 15, 17, 31, 39, 47, 63, 64, 71):
                assert rel(Y[:, c], H.mvprod_global_to_global(X[:, c])) < 1e-6, (li, storage, c)
