"""CPU checks of the C ABI: the library loads, exports every entry point the
header declares, and rejects bad arguments before touching a device."""
import ctypes as C

import numpy as np
import pytest


def test_library_exports_header_symbols(htg):
    from htool_amd import capi
    names = capi.header_symbols()
    assert len(names) >= 14
    L = capi.lib()
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/htool_amd.h but not exported"
    # every symbol the binding declares is in the header too
    assert set(capi._SIGNATURES) <= set(names)


def test_status_strings(htg):
    from htool_amd import capi
    L = capi.lib()
    assert L.htg_status_string(0) == b"ok"
    assert L.htg_status_string(capi.HTG_ERR_RANK_CAP) == b"ACA rank cap exceeded"


def test_invalid_arguments_rejected_without_device(htg):
    from htool_amd import capi
    L = capi.lib()
    prm = capi.htg_params(1e-6, 10.0, 1e6, 0, 0, -1, 0, 0.0, 0.0, None, 0.25)  # minclustersize 0
    gen = capi.htg_generator(0, 1, 0.0, 0.0)
    pts = np.zeros((4, 3))
    h = C.c_void_p()
    st = L.htg_hmatrix_build(C.byref(prm), C.byref(gen), 4, capi.ptr(pts), 4, None, C.byref(h))
    assert st == capi.HTG_ERR_INVALID
    assert b"MinClusterSize" in L.htg_last_error()
    prm.minclustersize = 10
    prm.compressor = 7
    assert L.htg_hmatrix_build(C.byref(prm), C.byref(gen), 4, capi.ptr(pts), 4, None, C.byref(h)) == \
        capi.HTG_ERR_UNSUPPORTED
    # symmetry / UPLO as HMatrix's constructor checks them (hmatrix.hpp:116-121)
    prm.compressor = 0
    for sy, up, status in ((ord("S"), ord("N"), capi.HTG_ERR_INVALID), (ord("N"), ord("L"), capi.HTG_ERR_INVALID),
                           (ord("X"), ord("L"), capi.HTG_ERR_INVALID)):
        prm.symmetry, prm.uplo = sy, up
        assert L.htg_hmatrix_build(C.byref(prm), C.byref(gen), 4, capi.ptr(pts), 4, None, C.byref(h)) == status
    # Hermitian storage needs a Hermitian family (refused before any device is touched)
    prm.symmetry, prm.uplo = ord("H"), ord("L")
    helm = capi.htg_generator(4, 1, 0.0, 1.0, 3.0)
    assert L.htg_hmatrix_build(C.byref(prm), C.byref(helm), 4, capi.ptr(pts), 4, None, C.byref(h)) == \
        capi.HTG_ERR_UNSUPPORTED
    assert b"Hermitian" in L.htg_last_error()
    prm.symmetry, prm.uplo = ord("S"), ord("L")
    gen2 = capi.htg_generator(0, 0, 0.0, 0.0)
    assert L.htg_hmatrix_build(C.byref(prm), C.byref(gen2), 4, capi.ptr(pts), 4, capi.ptr(pts), C.byref(h)) == \
        capi.HTG_ERR_INVALID  # symmetric storage needs one point set
    assert b"single point set" in L.htg_last_error()
    prm.symmetry, prm.uplo = 0, 0
    # from_leaves: leaf out of range
    perm = np.arange(10, dtype=np.int32)
    leaves = np.array([[5, 10, 0, 10, -1]], np.int32)
    data = np.zeros(100)
    st = L.htg_hmatrix_from_leaves(10, 10, capi.ptr(perm), capi.ptr(perm), 1, capi.ptr(leaves), capi.ptr(data),
                                   C.byref(h))
    assert st == capi.HTG_ERR_INVALID
    assert L.htg_hmatrix_destroy(None) == capi.HTG_OK


def test_python_mirror_setters(htg):
    H = htg.HMatrix()
    with pytest.raises(ValueError):
        H.set_maxblocksize(0)
    with pytest.raises(ValueError):
        H.set_minclustersize(0)
    H.set_compression("partialACA")
    with pytest.raises(RuntimeError):
        H.get_info()
