"""GPU: complex kernel families (Helmholtz e^{ikr}/(4 pi r) and the reference's
GeneratorTestComplex (1 + i)/(4 pi r)).

The complex build must give the oracle's leaves (the oracle is pinned against
HMatrix<std::complex<double>> of the reference build); the FP64 product must
equal the oracle's product of the same leaves within 1e-12 and the dense
product within the ACA tolerance; FP32 and 16-bit storage stay within their
rounding budgets.
"""
import numpy as np
import pytest

from htool_amd import geometry as geo

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def cvec(n, mu, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, mu)) + 1j * rng.standard_normal((n, mu))


def helmholtz_case(n, seed, ppw=10.0):
    x = geo.sphere(n, seed=seed)
    h = geo.sphere_spacing(n)
    return x, 2 * np.pi / (ppw * h), geo.laplace_diag(h)


@pytest.mark.parametrize("sym", [False, True])
def test_helmholtz_build_and_product_match_oracle(htg, orc, sym):
    print(f"helmholtz sym={sym}", flush=True)
    x, k, diag = helmholtz_case(3000, 31)
    H = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=10)
    H.set_compression("sympartialACA" if sym else "partialACA")
    H.build(htg.Generator(geo.KERNEL_HELMHOLTZ, 0.0, diag, k), x)
    assert H.is_complex
    O = orc.hmat_build(x, kernel=4, wavenumber=k, diag=diag, minclustersize=10, eta=10.0, eps=1e-6, sym=sym)
    leaves, data = H.get_leaves(with_data=True)
    assert np.array_equal(leaves, O.leaves)
    assert np.iscomplexobj(data)
    # dense leaves (generator) to rounding of sin/cos, factors to the ACA's rounding
    for (oi, m, oj, n, kk, a), (_, _, _, _, _, b) in zip(orc.iter_leaf_data(leaves, data),
                                                          orc.iter_leaf_data(O.leaves, O.data)):
        if kk < 0:
            assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(b))
        elif kk > 0:
            assert np.max(np.abs(a[0] @ a[1] - b[0] @ b[1])) <= 1e-9 * np.max(np.abs(b[0] @ b[1]))
    v = cvec(3000, 3, 1)
    y = H.mvprod_global_to_global(v)
    assert rel(y, O.mvprod(v)) < 1e-12
    v16 = cvec(3000, 16, 4)  # one 16-RHS pass on the FP64 matrix cores
    Y16 = H.mvprod_global_to_global(v16)
    Yo = O.mvprod(v16)
    for c in range(16):
        assert rel(Y16[:, c], Yo[:, c]) < 1e-12, c
    pt, ps = H.get_perms()
    assert rel(y, orc.leaf_mvprod(3000, 3000, pt, ps, leaves, data, v)) < 1e-12
    rows = np.random.default_rng(2).choice(3000, 300, replace=False).astype(np.int32)
    yd = orc.dense_rows_c(x, None, rows, v[:, 0], kernel=4, wavenumber=k, diag=diag, same=True)
    assert rel(y[rows, 0], yd) < 1e-5


def test_complex_test_kernel_on_two_disks(htg, orc):
    t = geo.disk(1500, 0.0, seed=41)
    s = geo.disk(1200, 1.0, seed=42)
    H = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=10)
    H.set_compression("partialACA")
    H.build(htg.Generator(geo.KERNEL_COMPLEX_TEST), t, s)
    O = orc.hmat_build(t, s, kernel=5, minclustersize=10, eta=10.0, eps=1e-6)
    assert np.array_equal(H.get_leaves(), O.leaves)
    v = cvec(1200, 2, 3)
    assert rel(H.mvprod_global_to_global(v), O.mvprod(v)) < 1e-12


@pytest.mark.parametrize("mu", [1, 5, 16, 40])
def test_complex_multi_rhs_columns(htg, mu):
    print(f"complex mu={mu}", flush=True)
    x, k, diag = helmholtz_case(6000, 51)
    H = htg.HMatrix(epsilon=1e-5, eta=10.0, minclustersize=32)
    H.set_compression("partialACA")
    H.build(htg.Generator(geo.KERNEL_HELMHOLTZ, 0.0, diag, k), x)
    V = cvec(6000, mu, mu)
    Y = H.mvprod_global_to_global(V)
    for c in range(mu):
        assert rel(Y[:, c], H @ V[:, c]) < 1e-12, c


def test_complex_storage_modes(htg, orc):
    x, k, diag = helmholtz_case(8000, 61)
    H = htg.HMatrix(epsilon=1e-5, eta=10.0, minclustersize=32)
    H.set_compression("partialACA")
    H.set_lowprec(0.0, all_terms=True)  # every low-rank term 16-bit, per-term scaled
    H.build(htg.Generator(geo.KERNEL_HELMHOLTZ, 0.0, diag, k), x)
    info = H.get_info()
    assert info.is_complex and info.lowprec_entries == info.lowrank_entries > 0
    V = cvec(8000, 4, 7)
    Y64 = H.mvprod_global_to_global(V)
    rows = np.random.default_rng(8).choice(8000, 300, replace=False).astype(np.int32)
    yd = orc.dense_rows_c(x, None, rows, V[:, 0], kernel=4, wavenumber=k, diag=diag, same=True)
    assert rel(Y64[rows, 0], yd) < 1e-4
    errs = {}
    for storage, bound in (("fp32", 2e-6), ("fp32_fp16", 2e-3), ("fp32_bf16", 2e-2)):
        H.set_storage(storage)
        errs[storage] = rel(H.mvprod_global_to_global(V), Y64)
        assert errs[storage] < bound, (storage, errs[storage])
    print("complex storage errors", errs, flush=True)


@pytest.mark.parametrize("storage,bound", [("fp32", 1e-6), ("fp32_fp16", 1e-6), ("fp32_bf16", 1e-6)])
def test_complex_matrix_core_passes_match_single_rhs(htg, storage, bound):
    """40 complex right-hand sides in FP32 arithmetic: two 16-wide matrix-core
    passes and an 8-wide pass; every column agrees with the single-vector
    product of the same storage (FP32 arithmetic both, different order)."""
    print(f"complex mfma {storage}", flush=True)
    x, k, diag = helmholtz_case(7000, 71)
    H = htg.HMatrix(epsilon=1e-5, eta=10.0, minclustersize=32)
    H.set_compression("partialACA")
    H.set_lowprec(2e-5)  # trailing terms in the 16-bit class
    H.build(htg.Generator(geo.KERNEL_HELMHOLTZ, 0.0, diag, k), x)
    info = H.get_info()
    assert 0 < info.lowprec_entries < info.lowrank_entries
    H.set_storage(storage)
    V = cvec(7000, 40, 12)
    Y = H.mvprod_global_to_global(V)
    for c in (0, 5, 15, 16, 31, 32, 39):
        assert rel(Y[:, c], H @ V[:, c]) < bound, c


def test_complex_recompression_ranks_match_numpy_svd(htg, orc):
    """FP64 QR+SVD recompression of complex leaves (Hermitian Grams, unitary
    Jacobi rotations): the kept rank of every leaf of rank 2..36 is, within one,
    the smallest r whose dropped singular values have a Frobenius norm below
    tol ||A_leaf|| (the truncation of the reference's SVD compressor,
    lrmat/SVD.hpp:46-55, applied to the leaf); the recompressed leaves' product
    is the H-matrix's product (1e-12) and within the tolerance of the ACA leaves'."""
    x, k, diag = helmholtz_case(5000, 81)
    tol = 3e-6
    H0 = htg.HMatrix(epsilon=1e-5, eta=10.0, minclustersize=32)
    H0.set_compression("partialACA")
    H0.build(htg.Generator(geo.KERNEL_HELMHOLTZ, 0.0, diag, k), x)
    H1 = htg.HMatrix(epsilon=1e-5, eta=10.0, minclustersize=32)
    H1.set_compression("partialACA")
    H1.set_recompression(tol)
    H1.build(htg.Generator(geo.KERNEL_HELMHOLTZ, 0.0, diag, k), x)
    L0, D0 = H0.get_leaves(with_data=True)
    L1, D1 = H1.get_leaves(with_data=True)
    assert np.array_equal(L0[:, :4], L1[:, :4])
    checked = 0
    for (oi, m, oj, n, kk, a), k1 in zip(orc.iter_leaf_data(L0, D0), L1[:, 4]):
        if kk < 0:
            continue
        if 2 <= kk <= 36:
            s = np.linalg.svd(a[0] @ a[1], compute_uv=False)[:kk]
            norm = np.sqrt((s ** 2).sum())
            j, tail = len(s), 0.0
            while True:  # SVD.hpp: drop from the smallest while sqrt(tail) / norm < tol
                j -= 1
                tail += s[j] ** 2
                if not (j > 0 and np.sqrt(tail) / norm < tol):
                    break
            r = j + 1
            assert abs(int(k1) - r) <= 1, (oi, oj, kk, k1, r)
            checked += 1
        else:
            assert k1 == kk
    assert checked > 100
    i0, i1 = H0.get_info(), H1.get_info()
    assert i1.lowrank_entries < i0.lowrank_entries
    v = cvec(5000, 2, 9)
    y1 = H1.mvprod_global_to_global(v)
    pt, ps = H1.get_perms()
    assert rel(y1, orc.leaf_mvprod(5000, 5000, pt, ps, L1, D1, v)) < 1e-12
    assert rel(y1, H0.mvprod_global_to_global(v)) < 10 * tol
