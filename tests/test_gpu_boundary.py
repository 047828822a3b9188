"""GPU: the parts of the reference's HMatrix surface beyond the built-in kernel
families (include/htool/types/virtual_hmatrix.hpp, hmatrix.hpp).

* A user generator on the host (VirtualGenerator::copy_submatrix,
  virtual_generator.hpp:22) through htg_hmatrix_build_ex: with entries equal
  to a device family's, the leaves and factors are bit for bit those of the
  device build on the multi-workgroup ACA, and the product the oracle's.
* The caller's cluster trees, set_use_permutation(false) (the reference's
  test_hmat_permutation.cpp: the same product in the cluster ordering),
  set_mintargetdepth / set_minsourcedepth (leaf lists equal to the oracle's,
  itself pinned against the reference driver).
* mymvprod_local_to_local (cluster ordering, interleaved right-hand sides;
  HPDDM's GMV), mvprod_subrhs, get_local_dense(_perm), get_local_diagonal, and
  the infos nb_mat_vec_prod / total_time_mat_vec_prod."""
import numpy as np
import pytest

from htool_amd import geometry as geo
from htool_amd import io

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


class OracleLaplace:
    """Host generator with the oracle's entries of 1/(4 pi r) (bitwise the device family's)."""

    def __init__(self, orc, xt, xs, diag):
        self.orc, self.xt, self.xs, self.diag = orc, xt, xs, diag
        self.calls = 0

    def copy_submatrix(self, M, N, rows, cols, ptr):
        self.calls += 1
        same = self.xs is None
        ptr[:] = self.orc.kernel_block(geo.KERNEL_LAPLACE, 0.0, self.diag, self.xt, self.xs, rows, cols, same=same)


class ComplexTest:
    """(1 + i)/(4 pi r) on the host (GeneratorTestComplex), as the device evaluates it."""
    is_complex = True

    def __init__(self, xt, xs):
        self.xt, self.xs = xt, xs

    def copy_submatrix(self, M, N, rows, cols, ptr):
        d = self.xt[rows][:, None, :] - self.xs[cols][None, :, :]
        r = np.sqrt(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2])
        v = 1.0 / (4 * np.pi * r)
        ptr[:] = v + 1j * v


@pytest.mark.parametrize("comp,uplo", [("partialACA", "N"), ("sympartialACA", "N"), ("partialACA", "L")])
def test_host_generator_equals_device_family(htg, orc, comp, uplo):
    print(f"host generator {comp} {uplo}", flush=True)
    n = 4000
    xt = geo.sphere(n, seed=61)
    diag = geo.laplace_diag(geo.sphere_spacing(n))
    kw = dict(symmetry="S", UPLO=uplo) if uplo != "N" else {}
    D = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=32, **kw)
    D.set_compression(comp)
    D.set_aca_wide_min(1)  # every block on the multi-workgroup ACA, as host generators run
    D.build(htg.Generator(geo.KERNEL_LAPLACE, 0.0, diag), xt)
    gen = OracleLaplace(orc, xt, None, diag)
    Hh = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=32, **kw)
    Hh.set_compression(comp)
    Hh.build(gen, xt)
    assert gen.calls > 0
    ld, dd = D.get_leaves(with_data=True)
    lh, dh = Hh.get_leaves(with_data=True)
    assert np.array_equal(lh, ld) and np.array_equal(dh, dd)
    O = orc.hmat_build(xt, kernel=geo.KERNEL_LAPLACE, diag=diag, minclustersize=32, eta=10.0, eps=1e-6,
                       sym=comp == "sympartialACA", uplo=uplo)
    assert np.array_equal(lh, O.leaves)
    x = np.random.default_rng(2).standard_normal((n, 3))
    assert rel(Hh.mvprod_global_to_global(x), O.mvprod(x)) < 1e-12
    i = Hh.get_info()
    assert i.n_aca_wide == i.n_aca > 0


def test_complex_host_generator_on_two_disks(htg, orc):
    t = geo.disk(1500, 0.0, seed=62)
    s = geo.disk(1100, 1.0, seed=63)
    H = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=10)
    H.set_compression("partialACA")
    H.build(ComplexTest(t, s), t, s)
    assert H.is_complex
    O = orc.hmat_build(t, s, kernel=geo.KERNEL_COMPLEX_TEST, minclustersize=10, eta=10.0, eps=1e-6)
    assert np.array_equal(H.get_leaves(), O.leaves)
    v = np.random.default_rng(3).standard_normal((1100, 2)) + 1j * np.random.default_rng(4).standard_normal((1100, 2))
    assert rel(H.mvprod_global_to_global(v), O.mvprod(v)) < 1e-12


def test_host_generator_error_is_raised(htg):
    class Broken:
        def copy_submatrix(self, M, N, rows, cols, ptr):
            raise RuntimeError("generator failed")
    xt = geo.sphere(2000, seed=64)
    H = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=32)
    with pytest.raises(RuntimeError, match="generator failed"):
        H.build(Broken(), xt)


def test_given_cluster_trees_and_no_permutation(htg, orc):
    n = 5000
    xt = geo.sphere(n, seed=65)
    diag = geo.laplace_diag(geo.sphere_spacing(n))
    gen = htg.Generator(geo.KERNEL_LAPLACE, 0.0, diag)
    # a tree of leaf size 20 given to a build whose own setting is 10
    T20 = io.cluster_tree(xt, 20)
    H = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=10)
    H.set_compression("partialACA")
    H.build(gen, xt, cluster_t=T20)
    R = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=20)
    R.set_compression("partialACA")
    R.build(gen, xt)
    assert np.array_equal(H.get_leaves(), R.get_leaves())
    x = np.random.default_rng(5).standard_normal((n, 4))
    y = R.mvprod_global_to_global(x)
    assert rel(H.mvprod_global_to_global(x), y) < 1e-14
    # set_use_permutation(false): the same tree, points, generator and vectors in the cluster ordering
    perm = T20.perm
    P = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=20)
    P.set_compression("partialACA")
    P.set_use_permutation(False)
    P.build(gen, xt[perm], cluster_t=T20)
    assert np.array_equal(P.get_leaves(), R.get_leaves())
    pt, ps = P.get_perms()
    assert np.array_equal(pt, np.arange(n)) and np.array_equal(ps, np.arange(n))
    yp = P.mvprod_global_to_global(x[perm])
    assert rel(yp, y[perm]) < 1e-13
    with pytest.raises(Exception, match="cluster trees"):
        Q = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=20)
        Q.set_use_permutation(False)
        Q.build(gen, xt)
    bad = io.cluster_tree(xt, 20)
    bad.perm = bad.perm.copy()
    bad.perm[0] = bad.perm[1]
    with pytest.raises(Exception, match="permutation"):
        htg.HMatrix(epsilon=1e-6, eta=10.0).build(gen, xt, cluster_t=bad)


@pytest.mark.parametrize("depths", [(7, 2), (2, 7), (6, 6)])
def test_min_depths(htg, orc, depths):
    n = 3000
    xt = geo.sphere(n, seed=12)
    diag = geo.laplace_diag(geo.sphere_spacing(n))
    H = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=10)
    H.set_mintargetdepth(depths[0])
    H.set_minsourcedepth(depths[1])
    H.build(htg.Generator(geo.KERNEL_LAPLACE, 0.0, diag), xt)
    with orc.min_depths(*depths):
        O = orc.hmat_build(xt, kernel=geo.KERNEL_LAPLACE, diag=diag, minclustersize=10, eta=10.0, eps=1e-6,
                           sym=True)
    assert np.array_equal(H.get_leaves(), O.leaves)
    x = np.random.default_rng(6).standard_normal(n)
    assert rel(H @ x, O.mvprod(x)) < 1e-12


@pytest.mark.parametrize("kind", ["real", "complex_sym"])
def test_local_products_subrhs_dense_and_infos(htg, orc, kind):
    torch = pytest.importorskip("torch")
    n = 3000
    xt = geo.sphere(n, seed=66)
    diag = geo.laplace_diag(geo.sphere_spacing(n))
    if kind == "real":
        H = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=10)
        H.build(htg.Generator(geo.KERNEL_LAPLACE, 0.0, diag), xt)
        dt = np.float64
    else:
        H = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=10, symmetry="S", UPLO="U")
        H.build(htg.Generator(geo.KERNEL_HELMHOLTZ, 0.0, diag, 4.0), xt)
        dt = np.complex128
    pt, ps = H.get_perms()
    rng = np.random.default_rng(7)
    mu = 5
    X = rng.standard_normal((n, mu)).astype(dt)
    if dt == np.complex128:
        X = X + 1j * rng.standard_normal((n, mu))
    Y = H.mvprod_global_to_global(X)
    # cluster ordering, interleaved: x_c[j * mu + s] = X[ps[j], s]
    xc = np.ascontiguousarray(X[ps, :]).reshape(-1)
    yc = H.mymvprod_local_to_local(xc, mu)
    assert rel(yc.reshape(n, mu), Y[pt, :]) < 1e-14
    # device pointers, with and without a caller's workspace
    e = 2 if dt == np.complex128 else 1
    d_in = torch.from_numpy(xc.view(np.float64).copy()).cuda()
    d_out = torch.zeros(n * mu * e, dtype=torch.float64, device="cuda")
    work = torch.zeros(2 * n * mu * e, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    for w in (work.data_ptr(), 0):
        d_out.zero_()
        H.mymvprod_local_to_local_device(d_in.data_ptr(), d_out.data_ptr(), mu, w, st)
        torch.cuda.synchronize()
        assert rel(d_out.cpu().numpy().view(dt), yc) < 1e-15
    # dense export and diagonal against the oracle's reconstruction of the leaves
    leaves, data = H.get_leaves(with_data=True)
    Ac = H.get_local_dense()
    ref = np.zeros((n, n), dtype=dt)
    for oi, m, oj, nn, k, a in orc.iter_leaf_data(leaves, data):
        blk = a if k < 0 else (a[0] @ a[1] if k > 0 else 0)
        ref[oi:oi + m, oj:oj + nn] = blk
        if kind != "real" and oi != oj:
            ref[oj:oj + nn, oi:oi + m] = np.asarray(blk).T
    assert rel(Ac, ref) < 1e-14
    Ao = H.get_local_dense_perm()
    assert rel(Ao[np.ix_(pt, ps)], Ac) == 0.0
    assert rel(Ao @ X, Y) < 1e-12
    assert np.array_equal(H.get_local_diagonal(False), np.diag(Ac))  # cluster ordering
    assert np.array_equal(H.get_local_diagonal()[pt], np.diag(Ac))   # original ordering (permutation=True)
    # mvprod_subrhs: leaves meeting the column window [off, off + size], in from the window
    off, size, margin = 700, 400, 60
    win = np.ascontiguousarray(X[ps, :][off - margin:off + size + margin]).reshape(-1)
    ys = H.mvprod_subrhs(win, mu, off, size, margin).reshape(n, mu)
    xw = np.zeros((n, mu), dtype=dt)
    xw[off - margin:off + size + margin] = X[ps, :][off - margin:off + size + margin]
    exp = np.zeros((n, mu), dtype=dt)
    for oi, m, oj, nn, k, a in orc.iter_leaf_data(leaves, data):
        blk = a if k < 0 else (a[0] @ a[1] if k > 0 else None)
        if blk is None:
            continue
        if oj <= off + size and off <= oj + nn:
            exp[oi:oi + m] += blk @ xw[oj:oj + nn]
        if kind != "real" and oi != oj and oi <= off + size and off <= oi + m:
            exp[oj:oj + nn] += np.asarray(blk).T @ xw[oi:oi + m]
    assert rel(ys, exp) < 1e-13
    i = H.get_info()
    assert i.nb_mat_vec_prod >= 5 and i.total_time_mat_vec_prod > 0
    assert H.get_infos()["nb_mat_vec_prod"] == i.nb_mat_vec_prod
