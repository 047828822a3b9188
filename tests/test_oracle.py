"""CPU checks that pin the oracle (oracle/oracle.cpp) before it is trusted:

* golden fixtures written by the reference Htool build (tests/golden/*.npz):
  permutations, leaf lists, ranks and leaf data must be identical, products
  within 1e-13;
* the reference driver itself where it is built (same checks on fresh cases);
* what the mathematics fixes: closed-form kernel entries, cluster-tree
  invariants, every matrix entry covered by exactly one leaf, admissibility of
  every low-rank leaf, the reference's lrmat bars (fixed rank 10 error < 1e-8,
  space saving 0.88 / (0.93, 0.96)), the Eckart-Young floor of a truncated SVD,
  the leaf product against a dense numpy reconstruction, brute-force dense rows.
"""
import glob
import math
import os

import numpy as np
import pytest

from htool_amd import geometry as geo

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


HMAT_FIXTURES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
                       if not os.path.basename(p).startswith(("lrmat", "intwrap")))
LRMAT_FIXTURES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "lrmat*.npz")))


def oracle_from_fixture(orc, g):
    xs = None if int(g["same"]) else g["xs"]
    uplo = chr(int(g["uplo"])) if "uplo" in g and int(g["uplo"]) else "N"
    symmetry = chr(int(g["symmetry"])) if "symmetry" in g else "S"
    return orc.hmat_build(g["xt"], xs, kernel=int(g["kernel"]), delta=float(g["delta"]), diag=float(g["diag"]),
                          minclustersize=int(g["mcs"]), eta=float(g["eta"]), eps=float(g["eps"]),
                          sym=bool(int(g["sym"])), uplo=uplo, symmetry=symmetry)


@pytest.mark.parametrize("name", HMAT_FIXTURES)
def test_oracle_matches_reference_fixture(orc, name):
    g = load(name)
    O = oracle_from_fixture(orc, g)
    assert np.array_equal(O.perm_t, g["perm_t"]) and np.array_equal(O.perm_s, g["perm_s"])
    assert np.array_equal(O.leaves, g["leaves"])
    assert O.false_positive == int(g["false_positive"])
    assert np.iscomplexobj(O.data) == np.iscomplexobj(g["data"])
    ours = list(orc.iter_leaf_data(O.leaves, O.data))
    cplx = np.iscomplexobj(g["data"])
    pos = 0
    for li in g["data_leaves"]:
        oi, m, oj, n, k, arr = ours[li]
        if k < 0:
            ref = g["data"][pos:pos + m * n].reshape(n, m).T
            pos += m * n
            assert np.array_equal(arr, ref)
        else:
            u = g["data"][pos:pos + m * k].reshape(k, m).T
            pos += m * k
            v = g["data"][pos:pos + k * n].reshape(n, k).T
            pos += k * n
            if cplx:
                # the reference's complex axpy (MKL zaxpy) and std::complex division
                # round differently from the restated FMA pairs and Smith's method
                assert np.max(np.abs(arr[0] - u)) <= 1e-8 * np.max(np.abs(u))
                assert np.max(np.abs(arr[1] - v)) <= 1e-8 * np.max(np.abs(v))
            else:
                assert np.array_equal(arr[0], u) and np.array_equal(arr[1], v)
    y = O.mvprod(g["x"])
    if int(g["kernel"]) == geo.KERNEL_ZERO:  # test_hmat_zero.hpp:163-166
        assert np.linalg.norm(g["y"]) == 0 and np.linalg.norm(y) < 1e-15
        assert (O.leaves[:, 4] == -1).all() and O.false_positive > 0  # every ACA found a zero first row
    else:
        assert rel(y, g["y"]) < 1e-13


@pytest.mark.parametrize("name", LRMAT_FIXTURES)
def test_oracle_aca_matches_reference_fixture(orc, name):
    g = load(name)
    T = orc.cluster_build(g["xt"], 10)
    S = orc.cluster_build(g["xs"], 10)
    assert np.array_equal(T.perm, g["perm_t"]) and np.array_equal(S.perm, g["perm_s"])
    k, U, V = orc.aca(g["xt"], g["xs"], T.perm, S.perm, kernel=geo.KERNEL_LAPLACE, eps=float(g["eps"]),
                      sym=bool(int(g["sym"])), ctr_t=T.geom[0, :3], ctr_s=S.geom[0, :3], reqrank=int(g["reqrank"]))
    assert k == int(g["rank"])
    assert np.array_equal(U, g["U"]) and np.array_equal(V, g["V"])


def test_oracle_matches_reference_driver(orc, ref):
    """Fresh cases straight from the reference build (this container only)."""
    cases = [
        (geo.cylinder(4000, seed=1), None, geo.KERNEL_INV_DELTA_R, 1e-5, 0.0, 10, 0),
        (geo.cylinder(4000, seed=1), None, geo.KERNEL_INV_DELTA_R, 1e-5, 0.0, 10, 1),
        (geo.sphere(8000, seed=3), None, geo.KERNEL_LAPLACE, 0.0, geo.laplace_diag(geo.sphere_spacing(8000)), 32, 0),
    ]
    for xt, xs, kernel, delta, diag, mcs, sym in cases:
        x = np.random.default_rng(5).standard_normal((xt.shape[0], 1))
        r = ref.run(xt, xs, x=x, kernel=kernel, delta=delta, diag=diag, minclustersize=mcs, eta=10.0, eps=1e-6,
                    compressor=sym)
        O = orc.hmat_build(xt, xs, kernel=kernel, delta=delta, diag=diag, minclustersize=mcs, eta=10.0, eps=1e-6,
                           sym=bool(sym))
        assert np.array_equal(O.perm_t, r.perm_t)
        order = np.lexsort((r.leaves[:, 2], r.leaves[:, 0]))
        assert np.array_equal(O.leaves, r.leaves[order])
        assert rel(O.mvprod(x), r.y) < 1e-13


def test_symmetric_oracle_matches_reference_driver(orc, ref):
    """HMatrix(..., 'S', UPLO) from the reference build: the kept triangle's leaf
    list and false positives, and y of its symmetric product (hmatrix.hpp:864-904)."""
    cases = [
        (geo.sphere(3000, seed=5), geo.KERNEL_LAPLACE, 0.0, geo.laplace_diag(geo.sphere_spacing(3000)), 0.0, 1, "L"),
        (geo.sphere(3000, seed=5), geo.KERNEL_LAPLACE, 0.0, geo.laplace_diag(geo.sphere_spacing(3000)), 0.0, 0, "U"),
        (geo.cylinder(2500, seed=6), geo.KERNEL_INV_DELTA_R, 1e-5, 0.0, 0.0, 1, "U"),
        (geo.sphere(2000, seed=7), 4, 0.0, 2.0, 9.0, 1, "L"),
        # Hermitian storage HMatrix(..., 'H', UPLO) with the reference's GeneratorTestComplexHermitian
        (geo.sphere(2500, seed=8), 7, 0.0, 0.0, 0.0, 1, "L", "H"),
        (geo.cylinder(2000, seed=9), 7, 0.0, 0.0, 0.0, 0, "U", "H"),
    ]
    for case in cases:
        xt, kernel, delta, diag, k, comp, uplo = case[:7]
        symmetry = case[7] if len(case) > 7 else "S"
        n = xt.shape[0]
        rng = np.random.default_rng(1)
        x = rng.standard_normal((n, 2)) + (1j * rng.standard_normal((n, 2)) if kernel in geo.COMPLEX_KERNELS else 0)
        r = ref.run(xt, None, x=x, kernel=kernel, delta=delta, diag=diag, wavenumber=k, minclustersize=10, eta=10.0,
                    eps=1e-6, compressor=comp, mu=2, uplo=uplo, symmetry=symmetry)
        O = orc.hmat_build(xt, kernel=kernel, delta=delta, diag=diag, wavenumber=k, minclustersize=10, eta=10.0,
                           eps=1e-6, sym=bool(comp), uplo=uplo, symmetry=symmetry)
        order = np.lexsort((r.leaves[:, 2], r.leaves[:, 0]))
        assert np.array_equal(O.leaves, r.leaves[order])
        assert O.false_positive == int(r.info["false_positive"])
        assert (O.leaves[:, 2] <= O.leaves[:, 0]).all() if uplo == "L" else (O.leaves[:, 2] >= O.leaves[:, 0]).all()
        assert rel(O.mvprod(x), r.y) < 1e-13


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_symmetric_storage_is_the_same_matrix(orc, uplo):
    """One triangle of blocks applied with its transposes is the whole matrix:
    the symmetric product agrees with the general H-matrix and with dense rows to
    the ACA tolerance, and the two triangles agree with each other; a symmetric
    leaf list covers each (i, j) with i >= j ('L') exactly once."""
    n = 1500
    xt = geo.sphere(n, seed=8)
    diag = geo.laplace_diag(geo.sphere_spacing(n))
    x = np.random.default_rng(2).standard_normal(n)
    Os = orc.hmat_build(xt, kernel=geo.KERNEL_LAPLACE, diag=diag, minclustersize=10, eps=1e-8, sym=True, uplo=uplo)
    Of = orc.hmat_build(xt, kernel=geo.KERNEL_LAPLACE, diag=diag, minclustersize=10, eps=1e-8, sym=True)
    ys = Os.mvprod(x)
    assert rel(ys, Of.mvprod(x)) < 1e-7
    rows = np.arange(0, n, 7, dtype=np.int32)
    assert rel(ys[rows], orc.dense_rows(xt, None, rows, x, kernel=geo.KERNEL_LAPLACE, diag=diag, same=True)) < 1e-7
    cover = np.zeros((n, n), np.int8)
    for oi, m, oj, nn, k in Os.leaves:
        cover[oi:oi + m, oj:oj + nn] += 1
        if oi != oj:
            cover[oj:oj + nn, oi:oi + m] += 1
    assert (cover == 1).all()


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_hermitian_storage_is_the_same_matrix(orc, uplo):
    """One triangle of a Hermitian matrix applied with the conjugate transposes
    of its off-diagonal blocks (and hemv on the diagonal leaves) is the whole
    matrix: the 'H' product agrees with the general H-matrix and with dense rows
    to the ACA tolerance; 'S' on the same leaves (plain transposes) does not."""
    n = 1500
    xt = geo.sphere(n, seed=10)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    K = geo.KERNEL_HERMITIAN_TEST
    Oh = orc.hmat_build(xt, kernel=K, minclustersize=10, eps=1e-8, sym=True, uplo=uplo, symmetry="H")
    Of = orc.hmat_build(xt, kernel=K, minclustersize=10, eps=1e-8, sym=True)
    yh = Oh.mvprod(x)
    err = rel(yh, Of.mvprod(x))
    assert err < 1e-7
    rows = np.arange(0, n, 7, dtype=np.int32)
    assert rel(yh[rows], orc.dense_rows_c(xt, None, rows, x, kernel=K, same=True)) < 1e-7
    Os = orc.hmat_build(xt, kernel=K, minclustersize=10, eps=1e-8, sym=True, uplo=uplo, symmetry="S")
    # (the 1e5 diagonal dominates the product, so the missing conjugations show at ~1e-4)
    assert np.array_equal(Os.leaves, Oh.leaves) and rel(Os.mvprod(x), yh) > 100 * err


@pytest.mark.parametrize("kernel", [0, 1, 2, 3, 6])
def test_kernel_entries_closed_form(orc, kernel):
    xt = geo.sphere(37, seed=1)
    xs = geo.disk(23, 2.0, seed=2)
    rows = np.arange(37, dtype=np.int32)[::-1].copy()
    cols = np.arange(23, dtype=np.int32)
    A = orc.kernel_block(kernel, 1e-5, 0.0, xt, xs, rows, cols)
    r = np.linalg.norm(xt[rows][:, None, :] - xs[cols][None, :, :], axis=2)
    ref = {0: 1 / (4 * math.pi * r), 1: 1 / (1e-5 + r), 2: 1 / (1e-5 + 4 * math.pi * r),
           3: 1 / (4 * math.pi * np.sqrt(1e-5 + r * r)), 6: np.zeros_like(r)}[kernel]
    assert np.allclose(A, ref, rtol=1e-14, atol=0)
    # diagonal rule of the singular Laplace kernel on a square same-point matrix
    D = orc.kernel_block(0, 0.0, 7.5, xt, None, rows, rows, same=True)
    assert np.all(np.diag(D) == 7.5)


@pytest.mark.parametrize("kernel", [4, 5, 7])
def test_complex_kernel_entries_closed_form(orc, kernel):
    """Columns of the complex families by brute force against numpy:
    e^{ikr}/(4 pi r) (diagonal value on a square matrix), (1 + i)/(4 pi r) and
    (1 + sign(x_i - x_j) i)/(1e-5 + 4 pi r), which is Hermitian entry for entry."""
    xt = geo.sphere(41, seed=3)
    # (1 + i)/(4 pi r) is singular on the diagonal: disjoint point sets, as the reference's tests use
    xs = None if kernel in (4, 7) else geo.disk(37, 2.0, seed=4)
    ns = 41 if xs is None else 37
    k, diag = 7.0, 3.25
    rows = np.arange(41, dtype=np.int32)
    A = np.zeros((41, ns), complex)
    for j in range(ns):
        e = np.zeros(ns, complex)
        e[j] = 1.0
        A[:, j] = orc.dense_rows_c(xt, xs, rows, e, kernel=kernel, wavenumber=k, diag=diag, same=xs is None)
    r = np.linalg.norm(xt[:, None, :] - (xt if xs is None else xs)[None, :, :], axis=2)
    with np.errstate(divide="ignore", invalid="ignore"):
        if kernel == 4:
            ref = np.exp(1j * k * r) / (4 * math.pi * r)
        elif kernel == 5:
            ref = (1 + 1j) / (4 * math.pi * r)
        else:
            ref = (1 + 1j * np.sign(xt[:, None, 0] - xt[None, :, 0])) / (1e-5 + 4 * math.pi * r)
    off = ~np.eye(41, ns, dtype=bool) if xs is None else np.ones((41, ns), bool)
    assert np.allclose(A[off], ref[off], rtol=1e-14, atol=0)
    if kernel == 4:
        assert np.all(np.diag(A) == diag)
    if kernel == 7:
        assert np.array_equal(A, A.conj().T) and np.all(np.diag(A) == 1.0 / 1e-5)


def test_complex_oracle_matches_reference_driver(orc, ref):
    """Complex H-matrices from the reference build (HMatrix<std::complex<double>>):
    same tree and leaves, dense leaves bitwise, products within 1e-13."""
    for kernel, xt, xs, k, diag, sym in ((4, geo.sphere(3000, seed=31), None, 9.0, 2.0, 0),
                                         (5, geo.disk(900, 0.0, seed=32), geo.disk(700, 1.3, seed=33), 0.0, 0.0, 1)):
        ns = xt.shape[0] if xs is None else xs.shape[0]
        rng = np.random.default_rng(kernel)
        x = rng.standard_normal((ns, 2)) + 1j * rng.standard_normal((ns, 2))
        r = ref.run(xt, xs, x=x, kernel=kernel, wavenumber=k, diag=diag, minclustersize=10, eta=10.0, eps=1e-6,
                    compressor=sym, mu=2)
        O = orc.hmat_build(xt, xs, kernel=kernel, wavenumber=k, diag=diag, minclustersize=10, eta=10.0, eps=1e-6,
                           sym=bool(sym))
        order = np.lexsort((r.leaves[:, 2], r.leaves[:, 0]))
        assert np.array_equal(O.leaves, r.leaves[order])
        assert rel(O.mvprod(x), r.y) < 1e-13


def test_cluster_tree_invariants(orc):
    x = geo.ellipsoid(3000, (1.0, 0.5, 2.0), seed=4)
    T = orc.cluster_build(x, 25)
    assert np.array_equal(np.sort(T.perm), np.arange(3000))
    nodes, geom = T.nodes, T.geom
    assert nodes[0, 1] == 3000 and nodes[0, 4] == 1  # root: all points, one son (one process)
    for c in range(len(nodes)):
        off, size, depth, rank, nsons, first = nodes[c]
        pts = x[T.perm[off:off + size]]
        assert np.allclose(geom[c, :3], pts.mean(axis=0), atol=1e-12)
        assert abs(geom[c, 3] - np.linalg.norm(pts - geom[c, :3], axis=1).max()) < 1e-12
        if nsons:
            sons = nodes[first:first + nsons]
            assert sons[0, 0] == off and sons[:, 1].sum() == size
            assert np.all(sons[1:, 0] == sons[:-1, 0] + sons[:-1, 1])  # contiguous
            if depth > 0:
                assert np.all(sons[:, 1] >= 25)
        # PCA split: son 0 lies on the positive side of the principal axis
        if nsons == 2:
            s0 = x[T.perm[nodes[first, 0]:nodes[first, 0] + nodes[first, 1]]]
            s1 = x[T.perm[nodes[first + 1, 0]:nodes[first + 1, 0] + nodes[first + 1, 1]]]
            w, vecs = np.linalg.eigh(np.cov((pts - geom[c, :3]).T, bias=True))
            d = vecs[:, -1]
            p0 = (s0 - geom[c, :3]) @ d
            p1 = (s1 - geom[c, :3]) @ d
            assert (np.all(p0 > -1e-9) and np.all(p1 < 1e-9)) or (np.all(p0 < 1e-9) and np.all(p1 > -1e-9))


def test_leaves_cover_matrix_once_and_far_leaves_admissible(orc):
    n = 1500
    x = geo.sphere(n, seed=6)
    eta = 10.0
    O = orc.hmat_build(x, kernel=0, diag=1.0, minclustersize=20, eta=eta, eps=1e-6)
    cover = np.zeros((n, n), np.int8)
    for oi, m, oj, nn, k in O.leaves:
        cover[oi:oi + m, oj:oj + nn] += 1
        if k >= 0:  # Rjasanow-Steinbach on the leaf's own clusters
            pt = x[O.perm_t[oi:oi + m]]
            ps = x[O.perm_s[oj:oj + nn]]
            ct, cs = pt.mean(axis=0), ps.mean(axis=0)
            rt = np.linalg.norm(pt - ct, axis=1).max()
            rs = np.linalg.norm(ps - cs, axis=1).max()
            assert 2 * min(rt, rs) < eta * max(np.linalg.norm(ct - cs) - rt - rs, 0.0) + 1e-12
    assert np.all(cover == 1)


def test_leaf_product_equals_dense_reconstruction(orc):
    n = 900
    x = geo.cylinder(n, seed=7)
    O = orc.hmat_build(x, kernel=1, delta=1e-5, minclustersize=10, eta=10.0, eps=1e-6)
    H = np.zeros((n, n))
    for oi, m, oj, nn, k, a in orc.iter_leaf_data(O.leaves, O.data):
        H[oi:oi + m, oj:oj + nn] = a if k < 0 else a[0] @ a[1]
    Hg = np.zeros((n, n))
    Hg[np.ix_(O.perm_t, O.perm_s)] = H  # back to original ordering
    v = np.random.default_rng(1).standard_normal((n, 3))
    assert rel(O.mvprod(v), Hg @ v) < 1e-13
    # and the H-matrix is the kernel matrix up to epsilon
    r = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)
    A = 1 / (1e-5 + r)
    assert np.linalg.norm(Hg - A) / np.linalg.norm(A) < 1e-5


def test_dense_rows_brute_force(orc):
    xt = geo.sphere(300, seed=8)
    xs = geo.disk(200, 3.0, seed=9)
    x = np.random.default_rng(2).standard_normal((200, 2))
    rows = np.array([0, 17, 299, 150], np.int32)
    out = orc.dense_rows(xt, xs, rows, x, kernel=0)
    A = 1 / (4 * math.pi * np.linalg.norm(xt[rows][:, None, :] - xs[None, :, :], axis=2))
    assert rel(out, A @ x) < 1e-14


@pytest.mark.parametrize("dist,seed", [(15, 100), (20, 101), (30, 102), (40, 200)])
def test_aca_reference_bars_and_svd_floor(orc, dist, seed):
    """tests/functional_tests/lrmat/test_lrmat.hpp:28-42, 83-88 with 500 x 100 disks. The
    reference draws its disks with C rand(); these are PCG64 disks (the ones of the
    lrmat fixtures for distances 15 and 40), on which its bars hold as well."""
    xt = geo.disk(500, 0.0, seed=seed)
    xs = geo.disk(100, float(dist), seed=seed + 1)
    T = orc.cluster_build(xt, 10)
    S = orc.cluster_build(xs, 10)
    A = orc.kernel_block(0, 0.0, 0.0, xt, xs, T.perm, S.perm)
    sv = np.linalg.svd(A, compute_uv=False)
    for sym in (False, True):
        k, U, V = orc.aca(xt, xs, T.perm, S.perm, kernel=0, eps=1e-4, sym=sym, ctr_t=T.geom[0, :3],
                          ctr_s=S.geom[0, :3], reqrank=10)
        assert k == 10
        err10 = np.linalg.norm(A - U @ V)
        assert err10 < 1e-8
        assert err10 >= np.sqrt((sv[10:] ** 2).sum()) * (1 - 1e-9)  # Eckart-Young: no rank-10 beats the SVD
        assert 0.87 < 1 - k * (1 / 500 + 1 / 100) < 0.89
        k, U, V = orc.aca(xt, xs, T.perm, S.perm, kernel=0, eps=1e-4, sym=sym, ctr_t=T.geom[0, :3],
                          ctr_s=S.geom[0, :3], reqrank=-1)
        assert 0.93 < 1 - k * (1 / 500 + 1 / 100) < 0.96
        assert np.linalg.norm(A - U @ V) / np.linalg.norm(A) < 1e-4


def test_aca_rejects_incompressible_block(orc):
    """A random dense block is not compressible: rank -1 (partialACA.hpp:74-77)."""
    # points far apart relative to each other within an 8 x 8 block: q(m+n) > mn quickly
    xt = geo.sphere(8, seed=3) * 5
    xs = geo.sphere(8, seed=4) * 5 + 0.01
    k, U, V = orc.aca(xt, xs, np.arange(8), np.arange(8), kernel=0, eps=1e-14, ctr_t=xt.mean(0), ctr_s=xs.mean(0))
    assert k == -1 and U is None


def test_aca_size_test_int_wrap_matches_reference(orc):
    """partialACA's q (M + N) > M N exit (partialACA.hpp:74) is evaluated in int
    by the reference: the fixture (written by the reference build) has two disks
    of 70000 points whose admissible 70000 x ~35000 blocks have M N > 2^31 - 1;
    the wrapped product is negative, the reference returns rank -1 at q = 1 and
    splits them (two false positives, four leaves). The oracle's 32-bit mode
    gives that leaf list and product; its default exact mode compresses the two
    blocks. The whole 70000 x 70000 block wraps to a positive 605032704, on which
    the reference's ACA runs normally (rank 14), as the oracle's 32-bit mode does."""
    g = load("intwrap_disks70000")
    n, seed, z = int(g["n"]), int(g["seed"]), float(g["z"])
    t = geo.disk(n, 0.0, seed=seed)
    s = geo.disk(n, z, seed=seed + 1)
    x = np.random.default_rng(seed + 2).standard_normal((n, 1))
    T = orc.cluster_build(t, 10)
    S = orc.cluster_build(s, 10)
    with orc.ref_int_overflow(True):
        assert orc.aca_size_bound(46340, 46340) == 46340 ** 2
        assert orc.aca_size_bound(46341, 46341) == 46341 ** 2 - 2 ** 32 < 0
        assert orc.aca_size_bound(n, n) == n * n - 2 ** 32 == 605032704
        O = orc.hmat_build(t, s, kernel=geo.KERNEL_LAPLACE, minclustersize=10, eta=10.0, eps=1e-6)
        k, _, _ = orc.aca(t, s, T.perm, S.perm, kernel=geo.KERNEL_LAPLACE, eps=1e-6, ctr_t=T.geom[0, :3],
                          ctr_s=S.geom[0, :3])
    assert np.array_equal(O.leaves, g["leaves"]) and O.false_positive == int(g["false_positive"]) == 2
    assert k == int(g["whole_block_rank"]) == 14
    assert rel(O.mvprod(x)[g["rows"], 0], g["y_rows"]) < 1e-12
    assert orc.aca_size_bound(46341, 46341) == 46341 ** 2  # back to exact arithmetic outside the context
    E = orc.hmat_build(t, s, kernel=geo.KERNEL_LAPLACE, minclustersize=10, eta=10.0, eps=1e-6)
    assert len(E.leaves) == 2 and E.false_positive == 0
    assert all(int(m) == n and int(m) * int(nn) > 2 ** 31 - 1 for _, m, _, nn, _ in E.leaves)
    assert rel(E.mvprod(x)[g["rows"], 0], g["y_rows"]) < 1e-5  # the same matrix within the ACA tolerance


def test_aca_size_test_int_wrap_reference_driver(orc, ref):
    """The reference's own leaf compression either side of M N = 2^31: a
    46340 x 46340 block compresses, a 46341 x 46341 block returns rank -1 at once;
    the oracle's 32-bit mode agrees, its exact mode compresses both."""
    for n, ref_rank in ((46340, None), (46341, -1)):
        t = geo.disk(n, 0.0, seed=301)
        s = geo.disk(n, 5.0, seed=302)
        r = ref.run(t, s, kernel=geo.KERNEL_LAPLACE, eps=1e-6, minclustersize=10, compressor=0, mode=1,
                    dump_data=False)
        T = orc.cluster_build(t, 10)
        S = orc.cluster_build(s, 10)
        kw = dict(kernel=geo.KERNEL_LAPLACE, eps=1e-6, ctr_t=T.geom[0, :3], ctr_s=S.geom[0, :3])
        with orc.ref_int_overflow(True):
            k32, _, _ = orc.aca(t, s, T.perm, S.perm, **kw)
        k64, _, _ = orc.aca(t, s, T.perm, S.perm, **kw)
        assert k32 == int(r.info["rank"])
        if ref_rank is None:
            assert k32 == k64 > 0
        else:
            assert k32 == ref_rank and k64 > 0


def test_min_depths_match_reference_driver(orc, ref):
    """HMatrix::set_mintargetdepth / set_minsourcedepth (hmatrix.hpp:215-216,
    blocks.hpp:51, 68): the reference's leaf list and product for depths (7, 2),
    (2, 7) and (6, 6), in the general and the symmetric block tree; the depths
    change the trees of this case (more, smaller blocks)."""
    n = 3000
    xt = geo.sphere(n, seed=12)
    diag = geo.laplace_diag(geo.sphere_spacing(n))
    x = np.random.default_rng(4).standard_normal((n, 1))
    counts = {}
    for uplo in ("N", "L"):
        for mt, ms in ((0, 0), (7, 2), (2, 7), (6, 6)):
            r = ref.run(xt, None, x=x, kernel=geo.KERNEL_LAPLACE, diag=diag, minclustersize=10, eta=10.0, eps=1e-6,
                        compressor=1, uplo=uplo, mintargetdepth=mt, minsourcedepth=ms)
            with orc.min_depths(mt, ms):
                O = orc.hmat_build(xt, kernel=geo.KERNEL_LAPLACE, diag=diag, minclustersize=10, eta=10.0, eps=1e-6,
                                   sym=True, uplo=uplo)
            order = np.lexsort((r.leaves[:, 2], r.leaves[:, 0]))
            assert np.array_equal(O.leaves, r.leaves[order]), (uplo, mt, ms)
            assert rel(O.mvprod(x), r.y) < 1e-13
            counts[uplo, mt, ms] = len(O.leaves)
        assert counts[uplo, 6, 6] > counts[uplo, 0, 0] and counts[uplo, 7, 2] > counts[uplo, 6, 6]
