"""htool_amd::HMatrixGpu<T> (include/htool_amd/htool_hmatrix.hpp): the
reference's htool::VirtualHMatrix<T> interface implemented over the C ABI.

tests/cpp/virtual_hmatrix_check is compiled against the reference's own
headers (tests/cpp/Makefile, where the reference is present at build time) and
drives the class only through VirtualHMatrix<T>, beside the reference's
HMatrix<T> on the same clusters and with the reference's own test generators
(GeneratorTestDouble, GeneratorTestDoubleSymmetric with 'S' storage,
GeneratorTestComplex, and the device family behind KernelGenerator). The
products, leaf counts, false positives and block sizes must agree with the
reference's, and the product with the oracle's, at 1e-12."""
import os
import subprocess

import numpy as np
import pytest

from htool_amd import geometry as geo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "virtual_hmatrix_check")


@pytest.fixture(scope="module")
def prog():
    if not os.path.isfile(BIN):
        pytest.skip("tests/cpp/_build/virtual_hmatrix_check not built (needs the reference's headers at build time)")
    return BIN


def run(prog, tmp_path, case, t, s):
    pt, ps = str(tmp_path / "t.f64"), str(tmp_path / "s.f64")
    t.astype(np.float64).tofile(pt)
    s.astype(np.float64).tofile(ps)
    out = str(tmp_path / f"case{case}")
    r = subprocess.run([prog, str(case), pt, ps, str(t.shape[0]), str(s.shape[0]), out], capture_output=True,
                       text=True, timeout=300)
    vals = dict(line.split("=", 1) for line in r.stdout.splitlines() if "=" in line)
    return r, vals, out + "_y.f64"


def test_adapter_reports_missing_device(prog, tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r, _, _ = run(prog, tmp_path, 0, geo.disk(300, 0.0, seed=1), geo.disk(200, 1.0, seed=2))
    assert r.returncode == 3 and "[Htool error]" in r.stderr


def xcols(n, cplx):
    j = np.arange(n)[:, None]
    c = np.arange(3)[None, :]
    x = np.sin(1.0 + j + 0.37 * c)
    return x + 1j * np.cos(2.0 * j - c) if cplx else x


@pytest.mark.gpu
@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_virtual_hmatrix_matches_reference_and_oracle(prog, tmp_path, orc, case):
    print(f"VirtualHMatrix case {case}", flush=True)
    if case == 1:
        t = s = geo.sphere(2500, seed=71)
    else:
        t, s = geo.disk(1800, 0.0, seed=72), geo.disk(1300, 1.0, seed=73)
    r, v, ypath = run(prog, tmp_path, case, t, s)
    assert r.returncode == 0, r.stderr
    for key in ("global_to_global", "local_to_local", "mymvprod_local_to_local"):
        assert float(v[key]) < 1e-12, (key, v[key])
    if case == 1:
        assert float(v["diagonal"]) == 0.0
    else:
        assert float(v["transp_global_to_global"]) < 1e-12
    for key in ("nlrmat", "ndmat", "false_positive", "max_size_blocks"):
        a, b = v[key].split()
        assert a == b, (key, v[key])
    assert int(v["nb_mat_vec_prod"]) == (3 if case == 1 else 4)  # the products above, transposed included
    if case == 3:
        assert float(v["copy_submatrix"]) == 0.0
    cplx = case == 2
    y = np.fromfile(ypath, dtype=np.complex128 if cplx else np.float64).reshape(3, t.shape[0]).T
    if case == 1:
        O = orc.hmat_build(t, kernel=geo.KERNEL_INV_DELTA_4PIR, delta=1e-5, minclustersize=10, eta=10.0, eps=1e-6,
                           sym=True, uplo="L")
    else:
        O = orc.hmat_build(t, s, kernel=geo.KERNEL_COMPLEX_TEST if cplx else geo.KERNEL_LAPLACE, minclustersize=10,
                           eta=10.0, eps=1e-6)
    yo = O.mvprod(xcols(s.shape[0], cplx))
    assert np.linalg.norm(y - yo) / np.linalg.norm(yo) < 1e-12
