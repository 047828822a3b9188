"""GPU: many independent H-matrices in one product, layouts, the autotuner and
graph replay.

A batch is the block-diagonal H-matrix of its instances; each instance's slice
of the batched product must equal the instance built alone (FP64, same leaves)
and the oracle's product of the instance (1e-12). Every layout is a different
work decomposition of the same sum, so products agree to rounding.
"""
import numpy as np
import pytest

from htool_amd import geometry as geo

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def instances(k, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        n = int(np.exp(rng.uniform(np.log(1500), np.log(6000))))
        axes = rng.uniform(0.5, 2.0, 3)
        pts = geo.ellipsoid(n, tuple(axes), seed=seed * 100 + i)
        h = np.sqrt(4 * np.pi * np.mean(axes) ** 2 / n)
        out.append((pts, geo.laplace_diag(h)))
    return out


def make(htg):
    H = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=32)
    H.set_compression("partialACA")
    return H


def test_batch_matches_instances_and_oracle(htg, orc):
    inst = instances(5, 3)
    B = make(htg).build_batch([htg.Generator(geo.KERNEL_LAPLACE, 0.0, d) for _, d in inst], [p for p, _ in inst])
    sizes = [p.shape[0] for p, _ in inst]
    assert B.nb_rows() == sum(sizes)
    x = np.random.default_rng(1).standard_normal((sum(sizes), 2))
    y = B.mvprod_global_to_global(x)
    leaves = B.get_leaves()
    off = 0
    for (pts, diag), n in zip(inst, sizes):
        print(f"instance n={n}", flush=True)
        H = make(htg).build(htg.Generator(geo.KERNEL_LAPLACE, 0.0, diag), pts)
        xi = x[off:off + n]
        assert rel(y[off:off + n], H.mvprod_global_to_global(xi)) < 1e-13
        O = orc.hmat_build(pts, kernel=geo.KERNEL_LAPLACE, diag=diag, minclustersize=32, eta=10.0, eps=1e-6)
        assert rel(y[off:off + n], O.mvprod(xi)) < 1e-12
        li = H.get_leaves().copy()
        li[:, 0] += off
        li[:, 2] += off
        mine = leaves[(leaves[:, 0] >= off) & (leaves[:, 0] < off + n)]
        assert np.array_equal(mine, li)
        off += n


@pytest.mark.parametrize("layout", [dict(rb_max=64), dict(rb_max=16, vrows=8), dict(vchunk_el=1024, vgroup_el=256),
                                    dict(vrows=64, vchunk_el=16384, vgroup_tiles=1)])
def test_layouts_give_the_same_product(htg, layout):
    print(f"layout {layout}", flush=True)
    x = geo.sphere(8000, seed=4)
    H = make(htg)
    H.set_recompression(2e-6)
    H.set_lowprec(5e-7)
    H.build(htg.Generator(geo.KERNEL_LAPLACE, 0.0, geo.laplace_diag(geo.sphere_spacing(8000))), x)
    v = np.random.default_rng(2).standard_normal((8000, 3))
    X = np.random.default_rng(3).standard_normal((8000, 32))
    y64 = H.mvprod_global_to_global(v)
    H.set_storage("fp32_fp16")
    Y16 = H.mvprod_global_to_global(X)
    H.set_storage("fp64")
    H.set_layout(**layout)
    assert {k: H.get_layout()[k] for k in layout} == layout
    assert rel(H.mvprod_global_to_global(v), y64) < 1e-13
    H.set_storage("fp32_fp16")
    assert rel(H.mvprod_global_to_global(X), Y16) < 1e-6


def test_autotune_picks_a_candidate_and_keeps_the_product(htg):
    inst = instances(4, 7)
    B = make(htg).build_batch([htg.Generator(geo.KERNEL_LAPLACE, 0.0, d) for _, d in inst], [p for p, _ in inst])
    x = np.random.default_rng(5).standard_normal(B.nb_cols())
    y = B @ x
    B.set_storage("fp32")
    y32 = B @ x
    start = B.get_layout()
    lay, ms, tried = B.autotune(mu=1, reps=3)
    assert tried == 25 and ms > 0
    assert lay == B.get_layout()
    assert lay == start or (lay["rb_max"] in (64, 128) and lay["vrows"] in (8, 16, 32) and lay["vgroup_tiles"] == 4)
    assert B.storage == "fp32"
    assert rel(B @ x, y32) < 1e-6
    B.set_storage("fp64")
    assert rel(B @ x, y) < 1e-13


@pytest.mark.parametrize("storage,mu", [("fp64", 1), ("fp32", 32), ("fp32_fp16", 4)])
def test_graph_replays_match_eager(htg, storage, mu):
    import torch
    print(f"graph {storage} mu={mu}", flush=True)
    x = geo.sphere(6000, seed=8)
    H = make(htg)
    H.set_recompression(2e-6)
    H.set_lowprec(5e-7)
    H.build(htg.Generator(geo.KERNEL_LAPLACE, 0.0, geo.laplace_diag(geo.sphere_spacing(6000))), x)
    H.set_storage(storage)
    s = torch.cuda.Stream()
    X = torch.randn(6000 * mu, dtype=torch.float64, device="cuda")
    Y = torch.zeros_like(X)
    Ye = torch.zeros_like(X)
    X2 = torch.randn_like(X)
    Y2 = torch.zeros_like(X)
    with torch.cuda.stream(s):
        H.mvprod_device(X.data_ptr(), Ye.data_ptr(), mu, s.cuda_stream)
        g = H.capture(X.data_ptr(), Y.data_ptr(), mu, s.cuda_stream)
        for r in range(5):
            Y.zero_()
            g.launch(s.cuda_stream)
            s.synchronize()
            assert torch.equal(Y, Ye), r
            # eager work between replays: another product, a torch op
            H.mvprod_device(X2.data_ptr(), Y2.data_ptr(), mu, s.cuda_stream)
            X2.mul_(1.0001)
            s.synchronize()
