"""Runs the reference Htool driver (oracle/_ref/ref_driver) on one case.

Test infrastructure only: used by tests/ and by bench.py's CPU-baseline leg.
The driver is built from the reference's own headers by oracle/Makefile; this
module writes its input file, runs it as a child process and reads back what it
dumps (see oracle/ref_driver.cpp for the file layouts).
"""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")


def available() -> bool:
    return os.path.isfile(REF_DRIVER) and os.access(REF_DRIVER, os.X_OK)


@dataclass
class RefResult:
    perm_t: np.ndarray
    perm_s: np.ndarray
    leaves: np.ndarray            # (L, 5) int: off_i, m, off_j, n, rank (-1 dense)
    data: np.ndarray | None       # concatenated leaf data in leaf order
    y: np.ndarray | None          # (nt, mu) column-major product, original ordering
    info: dict = field(default_factory=dict)

    def leaf_arrays(self):
        """Yields (off_i, m, off_j, n, rank, A or (U, V)) with U (m,k), V (k,n)."""
        pos = 0
        for oi, m, oj, n, k in self.leaves:
            if k < 0:
                a = self.data[pos:pos + m * n].reshape(n, m).T
                pos += m * n
                yield oi, m, oj, n, k, a
            else:
                u = self.data[pos:pos + m * k].reshape(k, m).T
                pos += m * k
                v = self.data[pos:pos + k * n].reshape(n, k).T
                pos += k * n
                yield oi, m, oj, n, k, (u, v)


def run(xt, xs=None, x=None, *, kernel=0, eta=10.0, eps=1e-6, minclustersize=10, compressor=0,
        delta=0.0, diag=0.0, maxblocksize=1e6, mu=1, nreps=0, dump_data=True, reqrank=-1, mode=0,
        threads=None, timeout=600, wavenumber=0.0, uplo="N", symmetry="S", mintargetdepth=0,
        minsourcedepth=0) -> RefResult:
    """kernel 4 (Helmholtz, wavenumber travels in the case file's delta) and 5
    run the reference's HMatrix<std::complex<double>>. uplo 'L' / 'U': the
    reference's symmetric storage HMatrix(..., 'S', uplo) (one point set)."""
    if kernel == 4 and wavenumber:
        delta = wavenumber
    if not available():
        raise RuntimeError("reference driver not built: run `make -C oracle`")
    xt = np.ascontiguousarray(xt, dtype=np.float64)
    same = xs is None
    nt = xt.shape[0]
    ns = nt if same else xs.shape[0]
    cplx = kernel in (4, 5, 7)  # complex families: x, y and the leaf data are complex128
    if x is None:
        x = np.zeros((ns, mu))
    x = np.asarray(x, dtype=np.complex128 if cplx else np.float64).reshape(ns, mu)
    if uplo != "N" and not same:
        raise ValueError("symmetric storage needs one point set")
    hdr = struct.pack("<16i", 0x48545246, 2, mode, nt, ns, int(same), kernel, minclustersize, compressor,
                      mu, nreps, int(dump_data), reqrank,
                      0 if uplo == "N" else ord(uplo) | (ord("H") << 8 if symmetry == "H" else 0),
                      int(mintargetdepth), int(minsourcedepth))
    prm = struct.pack("<5d", eta, eps, delta, diag, maxblocksize)
    with tempfile.TemporaryDirectory() as td:
        case = os.path.join(td, "case.bin")
        with open(case, "wb") as f:
            f.write(hdr)
            f.write(prm)
            f.write(xt.tobytes())
            if not same:
                f.write(np.ascontiguousarray(xs, dtype=np.float64).tobytes())
            f.write(np.asfortranarray(x).tobytes(order="F"))
        env = dict(os.environ)
        env["MKL_NUM_THREADS"] = "1"
        if threads is not None:
            env["OMP_NUM_THREADS"] = str(threads)
        subprocess.run([REF_DRIVER, case, td], check=True, env=env, timeout=timeout,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE)
        info = {}
        with open(os.path.join(td, "info.txt")) as f:
            for line in f:
                k, _, v = line.strip().partition("=")
                info[k] = v
        perm_t = np.fromfile(os.path.join(td, "perm_t.i32"), dtype=np.int32)
        perm_s = np.fromfile(os.path.join(td, "perm_s.i32"), dtype=np.int32)
        leaves = np.fromfile(os.path.join(td, "leaves.i32"), dtype=np.int32).reshape(-1, 5)
        dpath = os.path.join(td, "data.f64")
        dt = np.complex128 if cplx else np.float64
        data = np.fromfile(dpath, dtype=dt) if os.path.exists(dpath) else None
        ypath = os.path.join(td, "y.f64")
        y = np.fromfile(ypath, dtype=dt).reshape(mu, nt).T if os.path.exists(ypath) else None
    return RefResult(perm_t, perm_s, leaves, data, y, info)
