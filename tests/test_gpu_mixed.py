"""GPU: FP64 recompression and the mixed-precision storage classes.

Recompression (QR + SVD of each low-rank leaf, no counterpart in the
reference) is checked three ways: the product of the recompressed H-matrix
equals the oracle's product of the same exported leaves (1e-12, FP64); the kept
ranks match a numpy SVD truncation of the ACA leaves at the same tolerance
(within one term, where a singular value sits on the threshold); the
H-matrix stays within 10 epsilon of dense FP64. Mixed storage (FP32 terms plus
FP16 trailing terms) is checked against the FP64 product of the same leaves
within the error budget it was built with.
"""
import numpy as np
import pytest

from htool_amd import geometry as geo

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def sphere_case(n=8000, seed=2):
    x = geo.sphere(n, seed=seed)
    return x, geo.laplace_diag(geo.sphere_spacing(n))


def build(htg, x, diag, rc=0.0, lp=0.0, lp_all=False):
    H = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=32)
    H.set_compression("partialACA")
    H.set_recompression(rc)
    H.set_lowprec(lp, lp_all)
    H.build(htg.Generator(geo.KERNEL_LAPLACE, 0.0, diag), x)
    return H


def test_recompressed_product_matches_oracle_on_same_leaves(htg, orc):
    x, diag = sphere_case()
    H = build(htg, x, diag, rc=2e-6)
    leaves, data = H.get_leaves(with_data=True)
    pt, ps = H.get_perms()
    v = np.random.default_rng(1).standard_normal((x.shape[0], 3))
    assert rel(H.mvprod_global_to_global(v), orc.leaf_mvprod(len(pt), len(ps), pt, ps, leaves, data, v)) < 1e-12
    rows = np.random.default_rng(2).choice(x.shape[0], 500, replace=False).astype(np.int32)
    yd = orc.dense_rows(x, None, rows, v[:, 0], kernel=geo.KERNEL_LAPLACE, diag=diag, same=True)
    assert rel(H.mvprod_global_to_global(v[:, 0])[rows], yd) < 1e-5


def test_recompression_ranks_match_numpy_svd(htg, orc):
    x, diag = sphere_case(6000, 3)
    tol = 2e-6
    H0 = build(htg, x, diag)
    H1 = build(htg, x, diag, rc=tol)
    L0, D0 = H0.get_leaves(with_data=True)
    L1 = H1.get_leaves()
    assert np.array_equal(L0[:, :4], L1[:, :4])  # same tree, same leaves
    off = 0
    checked = 0
    for (oi, m, oj, n, k, a), k1 in zip(orc.iter_leaf_data(L0, D0), L1[:, 4]):
        if k < 0:
            continue
        if 2 <= k <= 48:
            s = np.linalg.svd(a[0] @ a[1], compute_uv=False)[:k]
            tail = np.sqrt(np.cumsum((s ** 2)[::-1]))[::-1]
            r = k
            for j in range(1, k):
                if tail[j] <= tol * np.sqrt((s ** 2).sum()):
                    r = j
                    break
            assert abs(int(k1) - r) <= 1, (oi, oj, k, k1, r)
            checked += 1
        else:
            assert k1 == k
    assert checked > 100
    assert H1.get_info().lowrank_entries < H0.get_info().lowrank_entries


@pytest.mark.parametrize("storage,bound", [("fp32_fp16", 1e-6), ("fp32_bf16", 8e-6)])
def test_mixed_storage_within_budget(htg, orc, storage, bound):
    x, diag = sphere_case()
    H = build(htg, x, diag, rc=2e-6, lp=5e-7)
    info = H.get_info()
    assert 0 < info.lowprec_entries < info.lowrank_entries
    v = np.random.default_rng(4).standard_normal(x.shape[0])
    y64 = H @ v
    H.set_storage(storage)
    y = H @ v
    assert rel(y, y64) < bound
    H.set_storage("fp32")
    assert rel(H @ v, y64) < 1e-6


def test_all_lowrank_terms_16bit(htg, orc):
    """Every low-rank term in BF16/FP16 (per-term scaled): error stays at the 16-bit level."""
    x, diag = sphere_case(6000, 5)
    H = build(htg, x, diag, rc=1e-6, lp_all=True)
    info = H.get_info()
    assert info.lowprec_entries == info.lowrank_entries
    v = np.random.default_rng(6).standard_normal(x.shape[0])
    y64 = H @ v
    for storage, bound in (("fp32_fp16", 2e-3), ("fp32_bf16", 2e-2)):
        H.set_storage(storage)
        assert rel(H @ v, y64) < bound
    H.set_storage("fp64")
    X = np.random.default_rng(7).standard_normal((x.shape[0], 8))
    Y64 = H.mvprod_global_to_global(X)
    H.set_storage("fp32_fp16")
    assert rel(H.mvprod_global_to_global(X), Y64) < 2e-3


@pytest.mark.parametrize("storage,bound", [("fp32", 1e-6), ("fp32_fp16", 1e-6), ("fp32_bf16", 8e-6)])
def test_multi_rhs_matrix_core_path(htg, storage, bound):
    """40 right-hand sides in FP32 arithmetic run as one 32-wide matrix-core pass
    plus an 8-wide pass; every column agrees with the single-vector product of the
    same storage (both FP32 arithmetic, different summation order) and with FP64
    within the storage's budget."""
    x, diag = sphere_case(7000, 9)
    H = build(htg, x, diag, rc=2e-6, lp=5e-7)
    X = np.random.default_rng(11).standard_normal((x.shape[0], 40))
    Y64 = H.mvprod_global_to_global(X)
    H.set_storage(storage)
    Y = H.mvprod_global_to_global(X)
    for c in (0, 7, 31, 32, 39):
        assert rel(Y[:, c], H @ X[:, c]) < 1e-6, c
    for c in range(40):
        assert rel(Y[:, c], Y64[:, c]) < bound, c


@pytest.mark.parametrize("vchunk", [256, 1024, 8192])
def test_single_rhs_chunk_cuts(htg, vchunk):
    """Single-RHS products in FP32 arithmetic walk their own V chunk list
    (panels no wider than vchunk_el whole, wider ones cut by columns and summed
    in a second pass) and, with two precision classes, run row blocks in parts;
    multi-RHS products use the 16-row chunks and whole row blocks. Every cut
    gives the multi-RHS columns (FP32 arithmetic, other summation order) and the
    FP64 product within the storage's budget."""
    x, diag = sphere_case(7000, 12)
    H = build(htg, x, diag, rc=2e-6, lp=5e-7)
    X = np.random.default_rng(13).standard_normal((x.shape[0], 3))
    Y64 = H.mvprod_global_to_global(X)
    H.set_layout(vchunk_el=vchunk)
    for storage, bound in (("fp32", 1e-6), ("fp32_fp16", 1e-6), ("fp32_bf16", 8e-6)):
        H.set_storage(storage)
        Y = H.mvprod_global_to_global(X)
        for c in range(3):
            y1 = H @ X[:, c]
            assert rel(y1, Y[:, c]) < 1e-6, (storage, c)
            assert rel(y1, Y64[:, c]) < bound, (storage, c)


@pytest.mark.parametrize("spread", [False, True])
@pytest.mark.parametrize("scale", [1e-7, 1.0, 1e7])
def test_matrix_core_16bit_terms_any_magnitude(htg, scale, spread):
    """32-RHS products of 16-bit stored blocks run on the 16-bit matrix cores
    with the vector entries split into 16-bit terms (FP16: under a power-of-two
    scale per staged chunk of x, and one for the t entries): results stay
    finite and agree with the single-vector product (FP32 arithmetic) whatever
    the magnitude of x, far outside FP16's range, and for x whose entries
    spread over ten orders of magnitude (chunks of very different scales)."""
    x, diag = sphere_case(7000, 14)
    H = build(htg, x, diag, rc=2e-6, lp=5e-7)
    rng = np.random.default_rng(15)
    X = rng.standard_normal((x.shape[0], 32)) * scale
    if spread:
        X *= 10.0 ** rng.uniform(-5, 5, (x.shape[0], 1))
    for storage in ("fp32_fp16", "fp32_bf16"):
        H.set_storage(storage)
        Y = H.mvprod_global_to_global(X)
        assert np.isfinite(Y).all(), storage
        for c in (0, 17, 31):
            assert rel(Y[:, c], H @ X[:, c]) < 1e-6, (storage, c)


@pytest.mark.parametrize("spread", [False, True])
@pytest.mark.parametrize("scale", [1e-7, 1.0, 1e7])
def test_complex_matrix_core_16bit_any_magnitude(htg, scale, spread):
    """Complex 64- and 32-RHS products of 16-bit stored blocks (96 RHS: one
    pass of each; the vector entries split into 16-bit terms, FP16 under a
    power-of-two scale per work item and one for the t entries): finite and
    equal to the single-vector product of the same storage, for x far outside
    FP16's range and for x whose entries spread over ten orders of magnitude."""
    n = 5000
    x = geo.ellipsoid(n, (1.0, 1.0, 4.0), seed=41)
    h = np.sqrt(geo.prolate_spheroid_area(1.0, 4.0) / n)
    k = 2 * np.pi / (10 * h)
    H = htg.HMatrix(epsilon=1e-5, eta=10.0, minclustersize=32)
    H.set_compression("partialACA")
    H.set_lowprec(0.0, all_terms=True)
    H.build(htg.Generator(geo.KERNEL_HELMHOLTZ, 0.0, geo.laplace_diag(h), k), x)
    rng = np.random.default_rng(16)
    X = (rng.standard_normal((n, 96)) + 1j * rng.standard_normal((n, 96))) * scale
    if spread:
        X *= 10.0 ** rng.uniform(-5, 5, (n, 1))
    for storage in ("fp32_fp16", "fp32_bf16"):
        H.set_storage(storage)
        Y = H.mvprod_global_to_global(X)
        assert np.isfinite(Y).all(), storage
        for c in (0, 17, 40, 63, 64, 95):
            assert rel(Y[:, c], H.mvprod_global_to_global(X[:, c])) < 1e-5, (storage, c)
