"""GPU: the ACA's fallback exits, forced.

* The zero generator of the reference's test_hmat_zero.hpp:18-65: every ACA
  finds a zero first row (partialACA.hpp:149-156: rank -1, a dense block, so
  the false-positive recursion runs down to the leaf clusters), and
  ||H x|| < 1e-15 (test_hmat_zero.hpp:163-166). Checked against the fixture the
  reference build wrote and against the oracle, for both compressors.
* Under-prediction of the learned rank predictor: a network whose output is 0
  presizes every block's workspace to 2 terms, so nearly every block outgrows it
  and runs again with the full workspace; the leaves must be bitwise those of
  the plain build and n_aca_retry the number of runs that needed more than
  their first workspace, on the one-workgroup and the multi-workgroup ACA."""
import os

import numpy as np
import pytest

from htool_amd import geometry as geo
from htool_amd.predictor import NFEAT, NPARAMS, RankPredictor

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("comp", ["partialACA", "sympartialACA"])
def test_zero_generator(htg, orc, comp):
    print(f"zero generator {comp}", flush=True)
    g = dict(np.load(os.path.join(GOLDEN, "disks500x400_zero.npz")))
    H = htg.HMatrix(epsilon=float(g["eps"]), eta=float(g["eta"]), minclustersize=int(g["mcs"]))
    H.set_compression(comp)
    H.build(htg.Generator(geo.KERNEL_ZERO), g["xt"], g["xs"])
    O = orc.hmat_build(g["xt"], g["xs"], kernel=geo.KERNEL_ZERO, minclustersize=int(g["mcs"]), eta=float(g["eta"]),
                       eps=float(g["eps"]), sym=comp == "sympartialACA")
    leaves = H.get_leaves()
    assert np.array_equal(leaves, O.leaves)
    assert H.get_info().n_false_positive == O.false_positive > 0
    assert (leaves[:, 4] == -1).all()
    if comp == "partialACA":
        assert np.array_equal(leaves, g["leaves"]) and O.false_positive == int(g["false_positive"])
    y = H @ g["x"][:, 0]
    assert np.linalg.norm(y) < 1e-15
    y = H.mvprod_global_to_global(np.random.default_rng(4).uniform(0, 10000, (g["xs"].shape[0], 32)))
    assert np.linalg.norm(y) < 1e-15


@pytest.mark.parametrize("wide_min", [0, 1024])
def test_forced_underprediction_reruns(htg, wide_min):
    print(f"forced under-prediction aca_wide_min={wide_min}", flush=True)
    n = 20000
    x = geo.sphere(n, seed=91)
    gen = htg.Generator(geo.KERNEL_LAPLACE, 0.0, geo.laplace_diag(geo.sphere_spacing(n)))
    P = RankPredictor(seed=0)
    P.set_params(np.zeros(NPARAMS, np.float32), np.zeros(NFEAT, np.float32), np.ones(NFEAT, np.float32))
    out = []
    for pred in (None, P):
        H = htg.HMatrix(epsilon=1e-6, eta=10.0, minclustersize=32)
        H.set_compression("partialACA")
        H.set_aca_wide_min(wide_min)
        if pred is not None:
            H.set_predictor(pred, margin=0.25)
        H.build(gen, x)
        out.append(H)
    H0, H1 = out
    l0, d0 = H0.get_leaves(with_data=True)
    l1, d1 = H1.get_leaves(with_data=True)
    assert np.array_equal(l0, l1) and np.array_equal(d0, d1)
    f, terms, cap = H1.aca_records()
    assert (cap <= 2).all()
    i1 = H1.get_info()
    assert i1.n_aca_retry == int(np.sum(terms > cap)) > 0.9 * i1.n_aca
    if wide_min:
        assert i1.n_aca_wide > 0
    assert H0.get_info().n_aca_retry == 0
    xv = np.random.default_rng(5).standard_normal(n)
    assert np.array_equal(H0 @ xv, H1 @ xv)
