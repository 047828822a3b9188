"""GPU: htg_aca_batch_noalloc, the allocation-free variant of htg_aca_batch
(its device buffers in the library's static 64 MiB scratch array).

Through the C ABI (htool_amd.aca_batch, ctypes): ranks and factors bit for bit
those of htg_aca_batch on the same blocks, and equal in rank to the oracle, on
two calls in a row, the second with fewer and smaller blocks (the scratch keeps
the first call's contents); a call that needs more than the scratch is refused
with HTG_ERR_INVALID before any device work."""
import numpy as np
import pytest

from htool_amd import geometry as geo

pytestmark = pytest.mark.gpu


def blocks_of(htg, orc, specs):
    """Two-disk blocks (m target points at z = 0, n source points at z = dist);
    one point array per side holds every block's points."""
    xt_parts, xs_parts, blocks = [], [], []
    ot = os_ = 0
    for i, (m, n, dist) in enumerate(specs):
        t = geo.disk(m, 0.0, seed=400 + i) + np.array([4.0 * i, 0.0, 0.0])
        s = geo.disk(n, float(dist), seed=500 + i) + np.array([4.0 * i, 0.0, 0.0])
        T = orc.cluster_build(t, 10)
        S = orc.cluster_build(s, 10)
        blocks.append(htg.Block(T.perm + ot, S.perm + os_, tuple(T.geom[0, :3]), tuple(S.geom[0, :3]), 0, 0))
        xt_parts.append(t)
        xs_parts.append(s)
        ot += m
        os_ += n
    return blocks, np.concatenate(xt_parts), np.concatenate(xs_parts)


def check_same(htg, orc, specs, comp, cap):
    blocks, xt, xs = blocks_of(htg, orc, specs)
    gen = htg.Generator(geo.KERNEL_LAPLACE)
    kw = dict(epsilon=1e-6, compressor=comp, cap=cap)
    ref = htg.aca_batch(gen, blocks, xt, xs, **kw)
    got = htg.aca_batch(gen, blocks, xt, xs, noalloc=True, **kw)
    assert len(got) == len(ref) == len(specs)
    for (k, U, V), (k0, U0, V0), b in zip(got, ref, blocks):
        assert k == k0
        if k > 0:
            assert np.array_equal(U, U0) and np.array_equal(V, V0)
        ko, _, _ = orc.aca(xt, xs, b.rows, b.cols, kernel=geo.KERNEL_LAPLACE, eps=1e-6,
                           sym=comp == "sympartialACA", ctr_t=b.ctr_t, ctr_s=b.ctr_s)
        assert k == ko
    return [k for k, _, _ in got]


@pytest.mark.parametrize("comp", ["partialACA", "sympartialACA"])
def test_noalloc_aca_equals_allocating_entry(htg, orc, comp):
    print(f"noalloc aca {comp}: first call", flush=True)
    big = [(2000, 1500, 3.0), (900, 1200, 2.5), (500, 100, 15.0), (3000, 800, 4.0), (1200, 1200, 6.0)]
    k1 = check_same(htg, orc, big, comp, cap=64)
    assert all(k > 0 for k in k1)
    print(f"noalloc aca {comp}: second, smaller call", flush=True)
    small = [(300, 200, 3.0), (150, 400, 5.0)]
    k2 = check_same(htg, orc, small, comp, cap=32)
    assert all(k > 0 for k in k2)


def test_noalloc_aca_refuses_oversize(htg, orc):
    from htool_amd import capi
    print("noalloc aca: oversize request", flush=True)
    blocks, xt, xs = blocks_of(htg, orc, [(20000, 20000, 3.0)])
    with pytest.raises(capi.HtgError) as e:
        htg.aca_batch(htg.Generator(geo.KERNEL_LAPLACE), blocks, xt, xs, cap=512, noalloc=True)
    assert e.value.status == capi.HTG_ERR_INVALID and "scratch" in str(e.value)
    # the allocating entry point takes it
    (k, _, _), = htg.aca_batch(htg.Generator(geo.KERNEL_LAPLACE), blocks, xt, xs, cap=512)
    assert k > 0
