"""ctypes binding of libhtool_amd.so (the C ABI of include/htool_amd.h).

The library is the product: it runs only on the GPU and there is no CPU
fallback. If the shared library is missing, importing the binding raises with
the command that builds it.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhtool_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "htool_amd.h")

HTG_OK = 0
HTG_ERR_INVALID = 1
HTG_ERR_HIP = 2
HTG_ERR_OOM = 3
HTG_ERR_UNSUPPORTED = 4
HTG_ERR_RANK_CAP = 5

STORAGE_FP64 = 0
STORAGE_FP32 = 1
STORAGE_FP32_FP16 = 2
STORAGE_FP32_BF16 = 3

PARTIAL_ACA = 0
SYM_PARTIAL_ACA = 1


class htg_generator(C.Structure):
    _fields_ = [("kernel", C.c_int), ("same_points", C.c_int), ("delta", C.c_double), ("diag", C.c_double),
                ("wavenumber", C.c_double)]


class htg_params(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("eta", C.c_double), ("maxblocksize", C.c_double),
                ("minclustersize", C.c_int), ("compressor", C.c_int), ("reqrank", C.c_int),
                ("lowprec_all", C.c_int), ("recompress_tol", C.c_double), ("lowprec_tol", C.c_double),
                ("predictor", C.c_void_p), ("predictor_margin", C.c_double), ("aca_wide_min", C.c_int),
                ("symmetry", C.c_int), ("uplo", C.c_int), ("delay_dense", C.c_int), ("mintargetdepth", C.c_int),
                ("minsourcedepth", C.c_int), ("no_permutation", C.c_int), ("ref_int_overflow", C.c_int)]


COPY_SUBMATRIX_FN = C.CFUNCTYPE(None, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.c_double), C.c_void_p)


class htg_host_generator(C.Structure):
    _fields_ = [("copy_submatrix", COPY_SUBMATRIX_FN), ("ctx", C.c_void_p), ("is_complex", C.c_int)]


class htg_cluster_tree(C.Structure):
    _fields_ = [("n", C.c_int), ("perm", C.c_void_p), ("nnodes", C.c_int), ("nodes_int", C.c_void_p),
                ("nodes_dbl", C.c_void_p)]


class htg_layout(C.Structure):
    _fields_ = [("rb_max", C.c_int), ("vchunk_el", C.c_int), ("vgroup_el", C.c_int), ("vgroup_tiles", C.c_int),
                ("vrows", C.c_int)]


class htg_info(C.Structure):
    _fields_ = [("nr", C.c_int), ("nc", C.c_int), ("n_dense", C.c_int), ("n_lowrank", C.c_int),
                ("n_false_positive", C.c_int), ("rank_min", C.c_int), ("rank_max", C.c_int),
                ("rank_mean", C.c_double), ("compression_ratio", C.c_double), ("dense_entries", C.c_int64),
                ("lowrank_entries", C.c_int64), ("lowprec_entries", C.c_int64), ("lowrank_u_entries", C.c_int64),
                ("lowprec_u_entries", C.c_int64), ("n_row_blocks", C.c_int),
                ("n_columns", C.c_int),
                ("n_vtiles", C.c_int), ("build_seconds", C.c_double), ("n_aca", C.c_int),
                ("n_aca_retry", C.c_int), ("aca_ws_entries", C.c_int64), ("aca_need_entries", C.c_int64),
                ("aca_seconds", C.c_double), ("is_complex", C.c_int), ("n_aca_wide", C.c_int),
                ("symmetry", C.c_int), ("uplo", C.c_int), ("mirrored_entries", C.c_int64),
                ("mirrored_lowprec_entries", C.c_int64), ("nb_mat_vec_prod", C.c_int64),
                ("total_time_mat_vec_prod", C.c_double), ("n_aca_int_wrap", C.c_int)]


class HtgError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"htool_amd error {status}: {message}")
        self.status = status


_VP = C.c_void_p
_I = C.c_int
_D = C.c_double

_SIGNATURES = {
    "htg_last_error": (C.c_char_p, []),
    "htg_status_string": (C.c_char_p, [_I]),
    "htg_hmatrix_build": (_I, [C.POINTER(htg_params), C.POINTER(htg_generator), _I, _VP, _I, _VP, C.POINTER(_VP)]),
    "htg_hmatrix_from_leaves": (_I, [_I, _I, _VP, _VP, _I, _VP, _VP, C.POINTER(_VP)]),
    "htg_hmatrix_destroy": (_I, [_VP]),
    "htg_hmatrix_set_storage": (_I, [_VP, _I]),
    "htg_hmatrix_get_info": (_I, [_VP, C.POINTER(htg_info)]),
    "htg_hmatrix_nleaves": (_I, [_VP, C.POINTER(_I), C.POINTER(C.c_int64)]),
    "htg_hmatrix_get_leaves": (_I, [_VP, _VP, _VP]),
    "htg_hmatrix_get_perms": (_I, [_VP, _VP, _VP]),
    "htg_cluster_build": (_I, [_I, _VP, _I, _VP, _VP, _VP, _VP]),
    "htg_hmatrix_set_dense_blocks": (_I, [_VP, _VP]),
    "htg_hmatrix_mvprod": (_I, [_VP, _VP, _VP, _I]),
    "htg_hmatrix_mvprod_device": (_I, [_VP, _VP, _VP, _I, _VP]),
    "htg_hmatrix_profile": (_I, [_VP, _VP, _VP, _I, _I, _VP, _VP]),
    "htg_aca_batch": (_I, [C.POINTER(htg_generator), _I, _D, _I, _I, _VP, _I, _VP, _I, _VP, _VP, _VP, _VP, _VP, _VP,
                           _VP, _VP, _VP, _VP, _I, _VP, _VP, _VP, _VP, _VP]),
    "htg_aca_batch_noalloc": (_I, [C.POINTER(htg_generator), _I, _D, _I, _I, _VP, _I, _VP, _I, _VP, _VP, _VP, _VP, _VP,
                                   _VP, _VP, _VP, _VP, _VP, _I, _VP, _VP, _VP, _VP, _VP]),
    "htg_release_workspace": (_I, []),
    "htg_hmatrix_build_ex": (_I, [C.POINTER(htg_params), C.POINTER(htg_generator), C.POINTER(htg_host_generator), _I,
                                  _VP, _I, _VP, C.POINTER(htg_cluster_tree), C.POINTER(htg_cluster_tree),
                                  C.POINTER(_VP)]),
    "htg_hmatrix_mymvprod_local_to_local_device": (_I, [_VP, _VP, _VP, _I, _VP, _VP]),
    "htg_hmatrix_mymvprod_local_to_local": (_I, [_VP, _VP, _VP, _I]),
    "htg_hmatrix_mvprod_subrhs": (_I, [_VP, _VP, _VP, _I, _I, _I, _I]),
    "htg_hmatrix_get_dense": (_I, [_VP, _VP, _I]),
    "htg_hmatrix_get_diagonal": (_I, [_VP, _VP, _I]),
    "htg_generator_block": (_I, [C.POINTER(htg_generator), _I, _VP, _I, _VP, _I, _I, _VP, _VP, _VP]),
    "htg_stream_probe": (_I, [C.c_uint64, _I, C.POINTER(_D), C.POINTER(_D), C.POINTER(_I)]),
    "htg_stream_copy": (_I, [_VP, _VP, C.c_uint64]),
    "htg_hmatrix_aca_records": (_I, [_VP, C.POINTER(_I), _VP, _VP, _VP]),
    "htg_hmatrix_build_batch": (_I, [C.POINTER(htg_params), _I, _VP, _VP, _VP, C.POINTER(_VP)]),
    "htg_hmatrix_transpose": (_I, [_VP, C.POINTER(_VP)]),
    "htg_hmatrix_get_layout": (_I, [_VP, C.POINTER(htg_layout)]),
    "htg_hmatrix_set_layout": (_I, [_VP, C.POINTER(htg_layout)]),
    "htg_hmatrix_autotune": (_I, [_VP, _I, _I, C.POINTER(htg_layout), C.POINTER(_D), C.POINTER(_I)]),
    "htg_graph_capture": (_I, [_VP, _VP, _VP, _I, _VP, C.POINTER(_VP)]),
    "htg_graph_launch": (_I, [_VP, _VP]),
    "htg_graph_destroy": (_I, [_VP]),
    "htg_predictor_create": (_I, [C.c_ulonglong, C.POINTER(_VP)]),
    "htg_predictor_destroy": (_I, [_VP]),
    "htg_predictor_set_params": (_I, [_VP, _VP, _VP, _VP]),
    "htg_predictor_get_params": (_I, [_VP, _VP, _VP, _VP]),
    "htg_predictor_loss_grad": (_I, [_VP, _VP, _VP, _I, C.POINTER(_D), _VP]),
    "htg_predictor_train": (_I, [_VP, _VP, _VP, _I, _I, _I, C.c_float, C.c_ulonglong, _I, _VP]),
    "htg_predictor_predict": (_I, [_VP, _VP, _I, _VP]),
}

MLP_NFEAT = 8
MLP_NPARAMS = 4801

_lib = None


def header_symbols() -> list[str]:
    """Entry points declared in include/htool_amd.h."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s*(htg_\w+)\s*\(", text, re.M)))


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        # PyTorch ships its own libamdhip64.so.7. Loading it first makes this
        # library bind to the same HIP runtime (same SONAME), so device pointers
        # and streams from torch are valid here and torch still sees the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.isfile(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `make -C htool_amd/csrc` "
                               "(or __graft_entry__.build()); htool_amd has no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != HTG_OK:
        msg = lib().htg_last_error().decode(errors="replace")
        raise HtgError(status, msg)


def ptr(a) -> C.c_void_p:
    """Address of a numpy array's data (None stays NULL)."""
    if a is None:
        return None
    return C.c_void_p(a.ctypes.data)
