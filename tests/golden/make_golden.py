"""Writes the golden fixtures of tests/golden/ from the reference Htool build.

Runs oracle/_ref/ref_driver (the reference's own headers compiled by
oracle/Makefile; needs /root/reference at build time) on small seeded cases and
stores inputs and outputs as .npz: points, permutations, the leaf list, leaf
data (all of it for the small cases, a sample of leaves for the larger ones) and
y = H x. The fixtures are data only; tests/test_oracle.py checks the oracle
against them on any machine.

    python tests/golden/make_golden.py            # every fixture
    python tests/golden/make_golden.py NAME ...   # only these
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from htool_amd import geometry as geo  # noqa: E402
from oracle import refcase  # noqa: E402


def leaf_slices(leaves):
    pos = 0
    out = []
    for oi, m, oj, n, k in leaves:
        size = m * n if k < 0 else k * (m + n)
        out.append((pos, size))
        pos += size
    return out


ONLY = set(sys.argv[1:])


def save(name, *, xt, xs, kernel, delta, diag, mcs, eta, eps, sym, x, sample=None, uplo="N", symmetry="S", **kw):
    if ONLY and name not in ONLY:
        return
    r = refcase.run(xt, xs, x=x, kernel=kernel, eta=eta, eps=eps, minclustersize=mcs, compressor=int(sym),
                    delta=delta, diag=diag, mu=x.shape[1], nreps=0, dump_data=True, uplo=uplo, symmetry=symmetry)
    order = np.lexsort((r.leaves[:, 2], r.leaves[:, 0]))
    sl = leaf_slices(r.leaves)
    leaves = r.leaves[order]
    chunks = [r.data[sl[i][0]:sl[i][0] + sl[i][1]] for i in order]
    if sample is None:
        keep = np.arange(len(leaves))
    else:
        rng = np.random.default_rng(1234)
        lr = np.flatnonzero(leaves[:, 4] >= 0)
        de = np.flatnonzero(leaves[:, 4] < 0)
        keep = np.sort(np.concatenate([rng.choice(lr, min(sample, len(lr)), replace=False),
                                       rng.choice(de, min(sample // 2, len(de)), replace=False)]))
    data = np.concatenate([chunks[i] for i in keep]) if len(keep) else np.zeros(0, dtype=r.data.dtype)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), xt=xt, xs=(np.zeros((0, 3)) if xs is None else xs),
                        same=np.int32(xs is None), kernel=np.int32(kernel), delta=delta, diag=diag,
                        mcs=np.int32(mcs), eta=eta, eps=eps, sym=np.int32(sym), perm_t=r.perm_t, perm_s=r.perm_s,
                        leaves=leaves, data_leaves=keep.astype(np.int32), data=data, x=x, y=r.y,
                        false_positive=np.int32(r.info.get("false_positive", 0)),
                        uplo=np.int32(0 if uplo == "N" else ord(uplo)), symmetry=np.int32(ord(symmetry)), **kw)
    print(name, "leaves", len(leaves), "stored leaf data", len(keep), "values", data.size)


def save_lrmat(name, *, dist, seed, reqrank, sym):
    """One leaf compression of a 500 x 100 block of two disks (the reference's
    lrmat tests, tests/functional_tests/lrmat/test_lrmat_partialACA.cpp:23-67)."""
    if ONLY and name not in ONLY:
        return
    xt = geo.disk(500, 0.0, seed=seed)
    xs = geo.disk(100, float(dist), seed=seed + 1)
    r = refcase.run(xt, xs, kernel=geo.KERNEL_LAPLACE, eps=1e-4, minclustersize=10, compressor=int(sym), mode=1,
                    reqrank=reqrank, dump_data=True)
    k = int(r.info["rank"])
    u = r.data[:500 * k].reshape(k, 500).T
    v = r.data[500 * k:].reshape(100, k).T
    np.savez_compressed(os.path.join(HERE, name + ".npz"), xt=xt, xs=xs, perm_t=r.perm_t, perm_s=r.perm_s,
                        rank=np.int32(k), U=u, V=v, reqrank=np.int32(reqrank), sym=np.int32(sym), eps=1e-4)
    print(name, "rank", k)


def save_int_wrap(name, *, n, seed, z):
    """Two disks of n points, z apart: the block tree's admissible blocks are
    n x ~n/2 with M N > 2^31 - 1, where the reference's int test q (M + N) > M N
    (partialACA.hpp:74) wraps, returns rank -1 and splits them. Stores the
    reference's leaf list, false positives and a 500-row sample of y = H x
    (x = standard normal from default_rng(seed + 2)); the points come from
    geometry.disk(n, 0, seed) and disk(n, z, seed + 1)."""
    if ONLY and name not in ONLY:
        return
    t = geo.disk(n, 0.0, seed=seed)
    s = geo.disk(n, float(z), seed=seed + 1)
    x = np.random.default_rng(seed + 2).standard_normal((n, 1))
    r = refcase.run(t, s, x=x, kernel=geo.KERNEL_LAPLACE, eta=10.0, eps=1e-6, minclustersize=10, compressor=0,
                    nreps=0, dump_data=False)
    order = np.lexsort((r.leaves[:, 2], r.leaves[:, 0]))
    rows = np.sort(np.random.default_rng(seed + 3).choice(n, 500, replace=False)).astype(np.int32)
    # the whole n x n block through one leaf compression (mode 1): rank -1 as well
    r1 = refcase.run(t, s, kernel=geo.KERNEL_LAPLACE, eps=1e-6, minclustersize=10, compressor=0, mode=1,
                     dump_data=False)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), n=np.int32(n), seed=np.int32(seed), z=float(z),
                        leaves=r.leaves[order], false_positive=np.int32(r.info.get("false_positive", 0)),
                        rows=rows, y_rows=r.y[rows, 0], whole_block_rank=np.int32(int(r1.info["rank"])))
    print(name, "leaves", r.leaves[order].tolist(), "whole-block rank", r1.info["rank"])


def main():
    if not refcase.available():
        sys.exit("reference driver not built: make -C oracle ref")
    rng = np.random.default_rng(99)
    p = geo.cylinder(300, seed=11)
    save("cyl300_partial", xt=p, xs=None, kernel=geo.KERNEL_INV_DELTA_R, delta=1e-5, diag=0.0, mcs=10, eta=10.0,
         eps=1e-6, sym=False, x=rng.uniform(-1, 1, (300, 2)))
    save("cyl300_sym", xt=p, xs=None, kernel=geo.KERNEL_INV_DELTA_R, delta=1e-5, diag=0.0, mcs=10, eta=10.0,
         eps=1e-6, sym=True, x=rng.uniform(-1, 1, (300, 1)))
    n = 2000
    s = geo.sphere(n, seed=12)
    save("sphere2000_laplace", xt=s, xs=None, kernel=geo.KERNEL_LAPLACE, delta=0.0,
         diag=geo.laplace_diag(geo.sphere_spacing(n)), mcs=32, eta=10.0, eps=1e-6, sym=False,
         x=rng.standard_normal((n, 1)), sample=20)
    t = geo.disk(800, 0.0, seed=13)
    u = geo.disk(500, 1.5, seed=14)
    save("disks800x500", xt=t, xs=u, kernel=geo.KERNEL_LAPLACE, delta=0.0, diag=0.0, mcs=10, eta=10.0, eps=1e-6,
         sym=False, x=rng.standard_normal((500, 1)), sample=20)
    e = geo.ellipsoid(1500, (1.0, 0.7, 1.6), seed=15)
    save("ellipsoid1500_smooth_sym", xt=e, xs=None, kernel=geo.KERNEL_SMOOTH, delta=1e-5, diag=0.0, mcs=10,
         eta=10.0, eps=1e-6, sym=True, x=rng.standard_normal((1500, 1)), sample=20)
    # complex families: Helmholtz e^{ikr}/(4 pi r) (wavenumber in `delta`) and (1 + i)/(4 pi r)
    n = 1000
    s = geo.sphere(n, seed=21)
    save("sphere1000_helmholtz", xt=s, xs=None, kernel=4, delta=6.0, diag=geo.laplace_diag(geo.sphere_spacing(n)),
         mcs=10, eta=10.0, eps=1e-6, sym=False, x=rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2)),
         sample=20)
    t = geo.disk(600, 0.0, seed=22)
    u = geo.disk(400, 1.2, seed=23)
    save("disks600x400_complex_sym", xt=t, xs=u, kernel=5, delta=0.0, diag=0.0, mcs=10, eta=10.0, eps=1e-6,
         sym=True, x=rng.standard_normal((400, 1)) + 1j * rng.standard_normal((400, 1)), sample=20)
    # symmetric storage HMatrix(..., 'S', UPLO): one triangle of blocks (hmatrix.hpp:343-400)
    n = 1200
    s = geo.sphere(n, seed=31)
    save("sphere1200_laplace_symL", xt=s, xs=None, kernel=geo.KERNEL_LAPLACE, delta=0.0,
         diag=geo.laplace_diag(geo.sphere_spacing(n)), mcs=10, eta=10.0, eps=1e-6, sym=True,
         x=rng.standard_normal((n, 2)), sample=20, uplo="L")
    c = geo.cylinder(900, seed=32)
    save("cyl900_partial_symU", xt=c, xs=None, kernel=geo.KERNEL_INV_DELTA_R, delta=1e-5, diag=0.0, mcs=10,
         eta=10.0, eps=1e-6, sym=False, x=rng.uniform(-1, 1, (900, 1)), sample=20, uplo="U")
    n = 800
    s = geo.sphere(n, seed=33)
    save("sphere800_helmholtz_symL", xt=s, xs=None, kernel=4, delta=5.0, diag=geo.laplace_diag(geo.sphere_spacing(n)),
         mcs=10, eta=10.0, eps=1e-6, sym=True, x=rng.standard_normal((n, 1)) + 1j * rng.standard_normal((n, 1)),
         sample=20, uplo="L")
    # Hermitian storage HMatrix(..., 'H', UPLO) with the reference's GeneratorTestComplexHermitian
    n = 900
    s = geo.sphere(n, seed=34)
    save("sphere900_hermitian_hermL", xt=s, xs=None, kernel=7, delta=0.0, diag=0.0, mcs=10, eta=10.0, eps=1e-6,
         sym=True, x=rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2)), sample=20, uplo="L",
         symmetry="H")
    c = geo.cylinder(700, seed=35)
    save("cyl700_hermitian_partial_hermU", xt=c, xs=None, kernel=7, delta=0.0, diag=0.0, mcs=10, eta=10.0,
         eps=1e-6, sym=False, x=rng.standard_normal((700, 1)) + 1j * rng.standard_normal((700, 1)), sample=20, uplo="U",
         symmetry="H")
    # the zero generator of test_hmat_zero.hpp:18-65, 92-128: disks of 500 and 400 points 10 apart,
    # eta 0.1, epsilon 1e-8 (every ACA finds a zero first row: dense leaves of zeros)
    save("disks500x400_zero", xt=geo.disk(500, 1.0, seed=36), xs=geo.disk(400, 11.0, seed=37), kernel=6, delta=0.0,
         diag=0.0, mcs=10, eta=0.1, eps=1e-8, sym=False, x=rng.uniform(0, 10000, (400, 1)))
    save_int_wrap("intwrap_disks70000", n=70000, seed=301, z=5.0)
    for dist, seed in ((15, 100), (40, 200)):
        for reqrank in (-1, 10):
            for sym in (False, True):
                save_lrmat(f"lrmat_d{dist}_r{reqrank}_{'sym' if sym else 'partial'}", dist=dist, seed=seed,
                           reqrank=reqrank, sym=sym)


if __name__ == "__main__":
    main()
