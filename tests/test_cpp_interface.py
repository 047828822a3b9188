"""The C++ interface (include/htool_amd/gpu_hmatrix.hpp) compiles against the
C ABI and is usable like the reference's HMatrix: on the CPU it reports the
missing device as an exception; on the GPU its product matches the oracle."""
import os
import subprocess

import numpy as np
import pytest

from htool_amd import geometry as geo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "adapter_check.cpp")


@pytest.fixture(scope="module")
def adapter(tmp_path_factory, htg):
    out = str(tmp_path_factory.mktemp("cpp") / "adapter_check")
    lib_dir = os.path.join(ROOT, "htool_amd")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), SRC, "-o", out,
                    "-L", lib_dir, "-lhtool_amd", f"-Wl,-rpath,{lib_dir}"], check=True)
    return out


def run(adapter, tmp_path, n, pts, x, kernel, delta, diag, mcs, sym):
    pf, xf, yf = (str(tmp_path / f) for f in ("p.f64", "x.f64", "y.f64"))
    pts.astype(np.float64).tofile(pf)
    x.astype(np.float64).tofile(xf)
    r = subprocess.run([adapter, pf, xf, yf, str(n), str(kernel), repr(delta), repr(diag), str(mcs), str(sym)],
                       capture_output=True, text=True, timeout=300)
    return r, (np.fromfile(yf) if r.returncode == 0 else None)


def test_cpp_interface_reports_missing_device(adapter, tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    n = 500
    r, _ = run(adapter, tmp_path, n, geo.cylinder(n, seed=1), np.ones(n), 1, 1e-5, 0.0, 10, 0)
    assert r.returncode == 3 and "[Htool error]" in r.stderr


@pytest.mark.gpu
def test_cpp_interface_product_matches_oracle(adapter, tmp_path, orc):
    n = 4000
    pts = geo.cylinder(n, seed=1)
    x = np.random.default_rng(3).uniform(-1, 1, n)
    r, y = run(adapter, tmp_path, n, pts, x, geo.KERNEL_INV_DELTA_R, 1e-5, 0.0, 10, 0)
    assert r.returncode == 0, r.stderr
    O = orc.hmat_build(pts, kernel=geo.KERNEL_INV_DELTA_R, delta=1e-5, minclustersize=10, eta=10.0, eps=1e-6)
    yo = O.mvprod(x)
    assert np.linalg.norm(y - yo) / np.linalg.norm(yo) < 1e-12
