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


def save(name, *, xt, xs, kernel, delta, diag, mcs, eta, eps, sym, x, sample=None, uplo="N", **kw):
    if ONLY and name not in ONLY:
        return
    r = refcase.run(xt, xs, x=x, kernel=kernel, eta=eta, eps=eps, minclustersize=mcs, compressor=int(sym),
                    delta=delta, diag=diag, mu=x.shape[1], nreps=0, dump_data=True, uplo=uplo)
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
                        uplo=np.int32(0 if uplo == "N" else ord(uplo)), **kw)
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
    for dist, seed in ((15, 100), (40, 200)):
        for reqrank in (-1, 10):
            for sym in (False, True):
                save_lrmat(f"lrmat_d{dist}_r{reqrank}_{'sym' if sym else 'partial'}", dist=dist, seed=seed,
                           reqrank=reqrank, sym=sym)


if __name__ == "__main__":
    main()
