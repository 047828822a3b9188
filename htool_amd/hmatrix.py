"""Python mirror of the reference's HMatrix product surface over the C ABI.

Names and argument meaning follow Htool 0.8.1 (reference files cited per
method); the work is done by libhtool_amd.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import capi
from . import geometry as _geo


@dataclass
class Generator:
    """Device kernel descriptor standing in for a VirtualGenerator
    (include/htool/types/virtual_generator.hpp:11-29) of a built-in kernel family."""
    kernel: int = _geo.KERNEL_LAPLACE
    delta: float = 0.0
    diag: float = 0.0
    wavenumber: float = 0.0  # KERNEL_HELMHOLTZ

    def _c(self, same: bool) -> capi.htg_generator:
        return capi.htg_generator(self.kernel, 1 if same else 0, float(self.delta), float(self.diag),
                                  float(self.wavenumber))

    @property
    def is_complex(self) -> bool:
        return self.kernel in _geo.COMPLEX_KERNELS

    def copy_submatrix(self, xt, xs, rows, cols) -> np.ndarray:
        """VirtualGenerator::copy_submatrix (virtual_generator.hpp:22): the (M, N)
        block A[rows, cols] of this family on the points xt (rows) and xs (columns;
        None: the same points), computed on the device."""
        xt = np.ascontiguousarray(xt, dtype=np.float64)
        same = xs is None
        xs_ = None if same else np.ascontiguousarray(xs, dtype=np.float64)
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        M, N = rows.shape[0], cols.shape[0]
        dt = np.complex128 if self.is_complex else np.float64
        out = np.zeros((M, N), dtype=dt, order="F")
        g = self._c(same)
        capi.check(capi.lib().htg_generator_block(C.byref(g), xt.shape[0], capi.ptr(xt),
                                                  xt.shape[0] if same else xs_.shape[0], capi.ptr(xs_), M, N,
                                                  capi.ptr(rows), capi.ptr(cols), capi.ptr(out.reshape(-1, order="F")
                                                                                        .view(np.float64))))
        return out


_COMPRESSORS = {"partialACA": capi.PARTIAL_ACA, "sympartialACA": capi.SYM_PARTIAL_ACA}
_STORAGE = {"fp64": capi.STORAGE_FP64, "fp32": capi.STORAGE_FP32, "fp32_fp16": capi.STORAGE_FP32_FP16,
            "fp32_bf16": capi.STORAGE_FP32_BF16}
# bytes per stored entry of (class 0, class 1)
_BYTES = {"fp64": (8, 8), "fp32": (4, 4), "fp32_fp16": (4, 2), "fp32_bf16": (4, 2)}


class HMatrix:
    """H-matrix on the GPU.

    HMatrix(epsilon, eta, reqrank, symmetry=..., UPLO=...) mirrors HMatrix<T>(t, s, epsilon, eta,
    symmetry, UPLO, reqrank) (include/htool/types/hmatrix.hpp:115-122): symmetry 'S' with UPLO
    'L' or 'U' keeps one triangle of blocks (single point set; the product applies each kept
    off-diagonal block and its transpose); the cluster trees are built inside
    `build` from the points and `minclustersize` (cluster.hpp:57-92, 334).
    """

    def __init__(self, epsilon: float = 1e-6, eta: float = 10.0, reqrank: int = -1, minclustersize: int = 10,
                 symmetry: str = "N", UPLO: str = "N"):
        if not (symmetry in "NSH" and UPLO in "NLU" and len(symmetry) == 1 and len(UPLO) == 1
                and ((symmetry == "N") == (UPLO == "N"))):
            raise ValueError("[Htool error] Invalid arguments to create HMatrix")
        self.symmetry = symmetry
        self.UPLO = UPLO
        self.epsilon = float(epsilon)
        self.eta = float(eta)
        self.reqrank = int(reqrank)
        self.minclustersize = int(minclustersize)
        self.maxblocksize = 1e6
        self.compressor = capi.SYM_PARTIAL_ACA  # the reference's default (hmatrix.hpp:297-299)
        self.recompress_tol = 0.0
        self.lowprec_tol = 0.0
        self.lowprec_all = False
        self.predictor = None
        self.predictor_margin = 0.25
        self.aca_wide_min = 0
        self.delay_dense = False
        self.ref_int_overflow = False
        self.mintargetdepth = 0
        self.minsourcedepth = 0
        self.use_permutation = True
        self._h = None
        self._storage = "fp64"

    # setters of include/htool/types/hmatrix.hpp:213-220
    def set_epsilon(self, e): self.epsilon = float(e)
    def set_eta(self, e): self.eta = float(e)
    def set_maxblocksize(self, m):
        if m == 0:
            raise ValueError("[Htool error] MaxBlockSize parameter cannot be zero")
        self.maxblocksize = float(m)
    def set_minclustersize(self, m):
        if m == 0:
            raise ValueError("[Htool error] MinClusterSize parameter cannot be zero")
        self.minclustersize = int(m)

    def set_compression(self, name: str):
        self.compressor = _COMPRESSORS[name]

    def set_mintargetdepth(self, d: int):
        """hmatrix.hpp:216 (blocks.hpp:51): blocks only from target clusters at least d deep."""
        self.mintargetdepth = int(d)

    def set_minsourcedepth(self, d: int):
        """hmatrix.hpp:215: blocks only from source clusters at least d deep."""
        self.minsourcedepth = int(d)

    def get_mintargetdepth(self) -> int:
        return self.mintargetdepth

    def get_minsourcedepth(self) -> int:
        return self.minsourcedepth

    def set_use_permutation(self, choice: bool):
        """hmatrix.hpp:218, 306-311: False means the generator, the points and the product
        vectors are in the cluster ordering of the trees passed to `build`."""
        self.use_permutation = bool(choice)

    # build-time options of the mixed-precision product (no reference counterpart)
    def set_recompression(self, tol: float):
        """FP64 QR+SVD recompression of every low-rank leaf, dropping a Frobenius tail <= tol (relative)."""
        self.recompress_tol = float(tol)

    def set_lowprec(self, tol: float = 0.0, all_terms: bool = False):
        """Assign trailing low-rank terms to 16-bit storage within an error budget of tol (relative,
        per leaf; needs recompression), or every low-rank term when all_terms."""
        self.lowprec_tol = float(tol)
        self.lowprec_all = bool(all_terms)

    def set_predictor(self, predictor, margin: float = 0.25):
        """Presize ACA workspaces from a trained RankPredictor (None: fixed first
        workspace of 64 terms). Ranks are the same either way."""
        self.predictor = predictor
        self.predictor_margin = float(margin)

    def set_delay_dense_computation(self, delay: bool = True):
        """set_delay_dense_computation (hmatrix.hpp:213-220): the build leaves the
        dense leaves zero; `build_dense_blocks` fills them afterwards."""
        self.delay_dense = bool(delay)

    def build_dense_blocks(self, generator) -> None:
        """HMatrix::build_dense_blocks (hmatrix.hpp:404-418): the generator's
        copy_dense_blocks(M, N, rows, cols, ptr) fills ptr[b], an M[b] x N[b]
        array, with the entries of dense leaf b for original row indices
        rows[b] and column indices cols[b] (VirtualDenseBlocksGenerator,
        virtual_dense_blocks_generator.hpp:11-17)."""
        leaves = self.get_leaves()
        pt, ps = self.get_perms()
        dense = leaves[leaves[:, 4] < 0]
        dt = np.complex128 if self.is_complex else np.float64
        M = [int(m) for m in dense[:, 1]]
        N = [int(n) for n in dense[:, 3]]
        rows = [pt[oi:oi + m] for oi, m in zip(dense[:, 0], M)]
        cols = [ps[oj:oj + n] for oj, n in zip(dense[:, 2], N)]
        ptr = [np.zeros((m, n), dtype=dt, order="F") for m, n in zip(M, N)]
        generator.copy_dense_blocks(M, N, rows, cols, ptr)
        data = np.concatenate([np.asarray(p, dtype=dt).ravel(order="F") for p in ptr]) if ptr else np.zeros(0, dt)
        data = np.ascontiguousarray(data)
        capi.check(capi.lib().htg_hmatrix_set_dense_blocks(self._handle(), capi.ptr(data.view(np.float64))))
        self._transposed = None  # made from the previous dense blocks

    def save_plot(self, outputname: str) -> str:
        """HMatrix::save_plot (hmatrix.hpp:1695-1712): the leaf list CSV `<outputname>_0.csv`."""
        from .io import save_plot
        return save_plot(self, outputname)

    def set_ref_int_overflow(self, on: bool = True):
        """Evaluate the ACA's q (M + N) > M N exit in 32-bit int, as the reference
        does (partialACA.hpp:74, sympartialACA.hpp:98): blocks with M N > 2^31 - 1
        then return rank -1 and are split, giving the reference's leaf list.
        Default off: the exact 64-bit test (get_info().n_aca_int_wrap counts
        the blocks where the two differ)."""
        self.ref_int_overflow = bool(on)

    def set_aca_wide_min(self, m: int):
        """Blocks with M + N >= m run the multi-workgroup ACA (0: the library
        default, 16384 for real and 4096 for complex entries; negative: never).
        Factors and ranks do not change."""
        self.aca_wide_min = int(m)

    # -- assembly ---------------------------------------------------------------
    def build(self, gen, xt: np.ndarray, xs: np.ndarray | None = None, cluster_t=None, cluster_s=None) -> "HMatrix":
        """HMatrix::build(A, xt, xs) (include/htool/types/hmatrix.hpp:285-341).

        gen: a Generator (a device kernel family), or any object with
        copy_submatrix(M, N, rows, cols, ptr) filling the (M, N) array ptr with
        A[rows[i], cols[j]] (a VirtualGenerator, virtual_generator.hpp:22; attribute
        is_complex for complex entries), evaluated on the host and fed to the GPU
        ACA row by row and column by column. cluster_t / cluster_s: the caller's
        cluster trees (io.ClusterTree, e.g. from io.read_cluster), else built from
        the points."""
        L = capi.lib()
        xt = np.ascontiguousarray(xt, dtype=np.float64)
        same = xs is None
        if not same:
            xs = np.ascontiguousarray(xs, dtype=np.float64)
        prm = self._params()
        h = C.c_void_p()
        host = not isinstance(gen, Generator)
        if not host and cluster_t is None and cluster_s is None:
            g = gen._c(same)
            capi.check(L.htg_hmatrix_build(C.byref(prm), C.byref(g), xt.shape[0], capi.ptr(xt),
                                           xt.shape[0] if same else xs.shape[0], capi.ptr(xs), C.byref(h)))
        else:
            keep = []
            tt = _tree_c(cluster_t, keep)
            ts = _tree_c(cluster_s, keep)
            g = None if host else gen._c(same)
            hg, errors = (_host_gen_c(gen, keep) if host else (None, []))
            capi.check(L.htg_hmatrix_build_ex(C.byref(prm), None if g is None else C.byref(g),
                                              None if hg is None else C.byref(hg), xt.shape[0], capi.ptr(xt),
                                              xt.shape[0] if same else xs.shape[0], capi.ptr(xs),
                                              None if tt is None else C.byref(tt),
                                              None if ts is None else C.byref(ts), C.byref(h)))
            if errors:
                capi.lib().htg_hmatrix_destroy(h)
                raise errors[0]
        self._close()
        self._h = h
        self._storage = "fp64"
        return self

    def _params(self) -> capi.htg_params:
        return capi.htg_params(self.epsilon, self.eta, self.maxblocksize, self.minclustersize, self.compressor,
                               self.reqrank, int(self.lowprec_all), self.recompress_tol, self.lowprec_tol,
                               self.predictor.handle if self.predictor is not None else None,
                               self.predictor_margin, self.aca_wide_min, ord(self.symmetry), ord(self.UPLO),
                               int(self.delay_dense), self.mintargetdepth, self.minsourcedepth,
                               int(not self.use_permutation), int(self.ref_int_overflow))

    def build_batch(self, gens, point_sets) -> "HMatrix":
        """Many independent square H-matrices as one block-diagonal H-matrix
        diag(H_0, ..., H_{K-1}); x and y stack the instances' vectors."""
        L = capi.lib()
        pts = [np.ascontiguousarray(p, dtype=np.float64) for p in point_sets]
        if len(gens) != len(pts) or not pts:
            raise ValueError("one generator per point set")
        n = np.array([p.shape[0] for p in pts], np.int32)
        garr = (capi.htg_generator * len(pts))(*[g._c(True) for g in gens])
        parr = (C.c_void_p * len(pts))(*[p.ctypes.data for p in pts])
        prm = self._params()
        h = C.c_void_p()
        capi.check(L.htg_hmatrix_build_batch(C.byref(prm), len(pts), garr, capi.ptr(n), parr, C.byref(h)))
        self._close()
        self._h = h
        self._storage = "fp64"
        self.instance_sizes = n
        return self

    # -- work decomposition -------------------------------------------------------
    def get_layout(self) -> dict:
        lay = capi.htg_layout()
        capi.check(capi.lib().htg_hmatrix_get_layout(self._handle(), C.byref(lay)))
        return {f: getattr(lay, f) for f, _ in capi.htg_layout._fields_}

    def set_layout(self, **kw) -> None:
        cur = self.get_layout()
        cur.update(kw)
        lay = capi.htg_layout(**cur)
        capi.check(capi.lib().htg_hmatrix_set_layout(self._handle(), C.byref(lay)))

    def autotune(self, mu: int = 1, reps: int = 10):
        """Times each candidate layout in the current storage, times the three
        fastest again (2 * reps products) and keeps the fastest of that second
        timing; returns (layout dict, its second time in ms per product,
        candidates tried)."""
        lay = capi.htg_layout()
        ms = C.c_double()
        cnt = C.c_int()
        capi.check(capi.lib().htg_hmatrix_autotune(self._handle(), mu, reps, C.byref(lay), C.byref(ms),
                                                   C.byref(cnt)))
        return {f: getattr(lay, f) for f, _ in capi.htg_layout._fields_}, ms.value, cnt.value

    def capture(self, d_x: int, d_y: int, mu: int, stream: int) -> "ProductGraph":
        """One product of device buffers captured in a HIP graph on `stream`."""
        return ProductGraph(self, d_x, d_y, mu, stream)

    @classmethod
    def from_leaves(cls, nr: int, nc: int, perm_t, perm_s, leaves, data) -> "HMatrix":
        """Product structures from externally built leaves (dense SubMatrix data and
        LowRankMatrix U, V factors; layout of include/htool_amd.h)."""
        L = capi.lib()
        perm_t = np.ascontiguousarray(perm_t, dtype=np.int32)
        perm_s = np.ascontiguousarray(perm_s, dtype=np.int32)
        leaves = np.ascontiguousarray(leaves, dtype=np.int32).reshape(-1, 5)
        data = np.ascontiguousarray(data, dtype=np.float64)
        h = C.c_void_p()
        capi.check(L.htg_hmatrix_from_leaves(nr, nc, capi.ptr(perm_t), capi.ptr(perm_s), leaves.shape[0],
                                             capi.ptr(leaves), capi.ptr(data), C.byref(h)))
        obj = cls()
        obj._h = h
        return obj

    def _close(self):
        self._transposed = None
        if self._h is not None and self._h.value:
            capi.lib().htg_hmatrix_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self._close()
        except Exception:
            pass

    def _handle(self):
        if self._h is None or not self._h.value:
            raise RuntimeError("HMatrix is not built")
        return self._h

    # -- storage ----------------------------------------------------------------
    def set_storage(self, storage: str) -> None:
        capi.check(capi.lib().htg_hmatrix_set_storage(self._handle(), _STORAGE[storage]))
        self._storage = storage

    @property
    def storage(self) -> str:
        return self._storage

    # -- products ---------------------------------------------------------------
    @property
    def is_complex(self) -> bool:
        return bool(self.get_info().is_complex)

    def mvprod_global_to_global(self, x: np.ndarray, mu: int = 1) -> np.ndarray:
        """y = H x with x (nc, mu) and y (nr, mu) in original ordering
        (include/htool/types/hmatrix.hpp:1073-1182). Host arrays, copied over PCIe;
        complex128 for the complex kernel families."""
        info = self.get_info()
        dt = np.complex128 if info.is_complex else np.float64
        x = np.asarray(x, dtype=dt)
        vec = x.ndim == 1
        xm = x.reshape(info.nc, -1)
        mu = xm.shape[1]
        xf = np.asfortranarray(xm)
        y = np.zeros((info.nr, mu), dtype=dt, order="F")
        capi.check(capi.lib().htg_hmatrix_mvprod(self._handle(), capi.ptr(xf), capi.ptr(y), mu))
        return y[:, 0].copy() if vec else np.ascontiguousarray(y)

    def transpose(self) -> "HMatrix":
        """The H-matrix of A^T (htg_hmatrix_transpose), in FP64 storage."""
        h = C.c_void_p()
        capi.check(capi.lib().htg_hmatrix_transpose(self._handle(), C.byref(h)))
        T = HMatrix(self.epsilon, self.eta, self.reqrank, self.minclustersize)
        T._h = h
        return T

    def mvprod_transp_global_to_global(self, x: np.ndarray) -> np.ndarray:
        """y = A^T x (include/htool/types/virtual_hmatrix.hpp:58), through the
        transposed H-matrix, built on first use and kept in this storage mode."""
        T = getattr(self, "_transposed", None)
        if T is None or T.storage != self._storage:
            T = self.transpose()
            if self._storage != "fp64":
                T.set_storage(self._storage)
            self._transposed = T
        return T.mvprod_global_to_global(x)

    def __matmul__(self, x):
        """operator* (include/htool/types/hmatrix.hpp:1596-1602)."""
        return self.mvprod_global_to_global(x)

    def mvprod_local_to_local(self, x: np.ndarray, mu: int = 1) -> np.ndarray:
        """hmatrix.hpp:1184-1274: with one process the local vectors are the global
        ones (original ordering), so this is mvprod_global_to_global."""
        return self.mvprod_global_to_global(x, mu)

    def mymvprod_local_to_local(self, x: np.ndarray, mu: int = 1) -> np.ndarray:
        """hmatrix.hpp:979-996: x (nc * mu) in the cluster ordering with interleaved
        right-hand sides (entry (j, s) at j * mu + s); returns y (nr * mu) likewise."""
        info = self.get_info()
        dt = np.complex128 if info.is_complex else np.float64
        x = np.ascontiguousarray(np.asarray(x, dtype=dt).reshape(-1))
        if x.size != info.nc * mu:
            raise ValueError("x must hold nc * mu entries")
        y = np.zeros(info.nr * mu, dtype=dt)
        capi.check(capi.lib().htg_hmatrix_mymvprod_local_to_local(self._handle(), capi.ptr(x), capi.ptr(y), mu))
        return y

    def mymvprod_local_to_local_device(self, d_in: int, d_out: int, mu: int, d_work: int = 0,
                                       stream: int = 0) -> None:
        """The same on device pointers (cluster ordering, interleaved), enqueued on
        `stream`; d_work: (nr + nc) * mu entries of device workspace, or 0."""
        capi.check(capi.lib().htg_hmatrix_mymvprod_local_to_local_device(
            self._handle(), C.c_void_p(d_in), C.c_void_p(d_out), mu, C.c_void_p(d_work or None),
            C.c_void_p(stream)))

    def mvprod_subrhs(self, x: np.ndarray, mu: int, offset: int, size: int, margin: int) -> np.ndarray:
        """hmatrix.hpp:1457-1554: x holds the cluster-ordered source entries
        [offset - margin, offset + size + margin) (interleaved); returns nr * mu."""
        info = self.get_info()
        dt = np.complex128 if info.is_complex else np.float64
        x = np.ascontiguousarray(np.asarray(x, dtype=dt).reshape(-1))
        if x.size < (size + 2 * margin) * mu:
            raise ValueError("x must hold (size + 2 margin) * mu entries")
        y = np.zeros(info.nr * mu, dtype=dt)
        capi.check(capi.lib().htg_hmatrix_mvprod_subrhs(self._handle(), capi.ptr(x), capi.ptr(y), mu, offset, size,
                                                        margin))
        return y

    def _dense(self, original: bool) -> np.ndarray:
        info = self.get_info()
        dt = np.complex128 if info.is_complex else np.float64
        A = np.zeros((info.nr, info.nc), dtype=dt, order="F")
        capi.check(capi.lib().htg_hmatrix_get_dense(self._handle(), capi.ptr(A), int(original)))
        return A

    def get_local_dense(self) -> np.ndarray:
        """hmatrix.hpp:1728-1757: the dense matrix in the cluster ordering."""
        return self._dense(False)

    def get_local_dense_perm(self) -> np.ndarray:
        """copy_local_dense_perm: the dense matrix in the original ordering."""
        return self._dense(True)

    def get_local_diagonal(self, permutation: bool = True) -> np.ndarray:
        """get_local_diagonal(permutation) (hmatrix.hpp:1842-1874): the diagonal entries
        in the original ordering (permutation, the default) or the cluster one."""
        info = self.get_info()
        d = np.zeros(info.nr, dtype=np.complex128 if info.is_complex else np.float64)
        capi.check(capi.lib().htg_hmatrix_get_diagonal(self._handle(), capi.ptr(d), int(bool(permutation))))
        return d

    def mvprod_device(self, d_x: int, d_y: int, mu: int, stream: int = 0) -> None:
        """Same product on device pointers (e.g. torch tensors' data_ptr()), column-major
        nc x mu / nr x mu, enqueued on `stream`."""
        capi.check(capi.lib().htg_hmatrix_mvprod_device(self._handle(), C.c_void_p(d_x), C.c_void_p(d_y), mu,
                                                        C.c_void_p(stream)))

    def profile(self, d_x: int, d_y: int, mu: int, reps: int, stream: int = 0) -> dict:
        ms = (C.c_double * 4)()
        capi.check(capi.lib().htg_hmatrix_profile(self._handle(), C.c_void_p(d_x), C.c_void_p(d_y), mu, reps,
                                                  C.c_void_p(stream), ms))
        return {"gather_ms": ms[0], "vx_ms": ms[1], "row_apply_ms": ms[2], "total_ms": ms[3]}

    # -- introspection ------------------------------------------------------------
    def get_info(self) -> capi.htg_info:
        info = capi.htg_info()
        capi.check(capi.lib().htg_hmatrix_get_info(self._handle(), C.byref(info)))
        return info

    def get_infos(self) -> dict:
        """Subset of the reference's infos map (include/htool/types/hmatrix.hpp:794-840)."""
        i = self.get_info()
        return {"Number_of_lrmat": i.n_lowrank, "Number_of_dmat": i.n_dense,
                "Number_of_false_positive": i.n_false_positive, "Rank_min": i.rank_min, "Rank_max": i.rank_max,
                "Rank_mean": i.rank_mean, "Compression_ratio": i.compression_ratio,
                "Target_size": i.nr, "Source_size": i.nc, "Blocks_max": i.build_seconds,
                "nb_mat_vec_prod": i.nb_mat_vec_prod, "total_time_mat_vec_prod": i.total_time_mat_vec_prod}

    def nb_rows(self) -> int:
        return self.get_info().nr

    def nb_cols(self) -> int:
        return self.get_info().nc

    def aca_records(self):
        """(features n x 8, terms needed, first workspace terms) of the build's ACA runs."""
        n = C.c_int(0)
        L = capi.lib()
        capi.check(L.htg_hmatrix_aca_records(self._handle(), C.byref(n), None, None, None))
        feat = np.zeros((n.value, 8), np.float32)
        terms = np.zeros(n.value, np.int32)
        cap = np.zeros(n.value, np.int32)
        capi.check(L.htg_hmatrix_aca_records(self._handle(), C.byref(n), capi.ptr(feat), capi.ptr(terms),
                                             capi.ptr(cap)))
        return feat, terms, cap

    def get_perms(self):
        i = self.get_info()
        pt = np.zeros(i.nr, np.int32)
        ps = np.zeros(i.nc, np.int32)
        capi.check(capi.lib().htg_hmatrix_get_perms(self._handle(), capi.ptr(pt), capi.ptr(ps)))
        return pt, ps

    def get_leaves(self, with_data: bool = False):
        """Leaf list (save_plot layout, sorted by (off_i, off_j)) and, optionally,
        the FP64 leaf data."""
        n = C.c_int()
        ds = C.c_int64()
        capi.check(capi.lib().htg_hmatrix_nleaves(self._handle(), C.byref(n), C.byref(ds)))
        leaves = np.zeros((n.value, 5), np.int32)
        cplx = self.is_complex
        data = np.zeros(ds.value * (2 if cplx else 1)) if with_data else None
        capi.check(capi.lib().htg_hmatrix_get_leaves(self._handle(), capi.ptr(leaves), capi.ptr(data)))
        if with_data and cplx:
            data = data.view(np.complex128)
        return (leaves, data) if with_data else leaves

    def algorithmic_bytes(self, mu: int = 1, storage: str | None = None) -> int:
        """Bytes one product must move: every stored entry once (at its storage
        width), x read and y written once, the two permutations."""
        i = self.get_info()
        s0, s1 = _BYTES[storage or self._storage]
        e = 2 if i.is_complex else 1  # a complex entry is two values
        # symmetric storage: the product also streams the mirrored copies
        hi = i.dense_entries + i.lowrank_entries - i.lowprec_entries + i.mirrored_entries - i.mirrored_lowprec_entries
        stored = e * (s0 * hi + s1 * (i.lowprec_entries + i.mirrored_lowprec_entries))
        return int(stored + e * mu * (i.nc + i.nr) * 8 + 4 * (i.nr + i.nc))

    def flops(self, mu: int = 1) -> int:
        i = self.get_info()
        return int((8 if i.is_complex else 2) * mu * (i.dense_entries + i.lowrank_entries + i.mirrored_entries))


def _tree_c(tree, keep):
    """An io.ClusterTree as the C ABI's htg_cluster_tree (arrays kept alive in `keep`)."""
    if tree is None:
        return None
    perm = np.ascontiguousarray(tree.perm, dtype=np.int32)
    ni = np.ascontiguousarray(np.stack([tree.offset, tree.size, tree.depth, tree.rank, tree.nb_sons,
                                        tree.first_son], axis=1), dtype=np.int32)
    nd = np.ascontiguousarray(np.concatenate([np.asarray(tree.ctr, np.float64).reshape(-1, 3),
                                              np.asarray(tree.rad, np.float64).reshape(-1, 1)], axis=1))
    keep.extend([perm, ni, nd])
    return capi.htg_cluster_tree(perm.shape[0], perm.ctypes.data, ni.shape[0], ni.ctypes.data, nd.ctypes.data)


def _host_gen_c(gen, keep):
    """A Python generator (copy_submatrix(M, N, rows, cols, ptr) on an (M, N) array)
    behind the C ABI's callback; exceptions are collected and raised after the build."""
    cplx = bool(getattr(gen, "is_complex", False))
    dt = np.complex128 if cplx else np.float64
    errors = []

    def cb(M, N, rows, cols, ptr, ctx):
        if errors:
            return
        try:
            r = np.ctypeslib.as_array(rows, shape=(M,)).copy()
            c = np.ctypeslib.as_array(cols, shape=(N,)).copy()
            out = np.zeros((M, N), dtype=dt, order="F")
            gen.copy_submatrix(M, N, r, c, out)
            flat = np.ctypeslib.as_array(ptr, shape=(M * N * (2 if cplx else 1),))
            flat[:] = np.asfortranarray(out, dtype=dt).reshape(-1, order="F").view(np.float64)
        except Exception as e:  # noqa: BLE001 - re-raised by the caller after the build
            errors.append(e)

    fn = capi.COPY_SUBMATRIX_FN(cb)
    keep.append(fn)
    return capi.htg_host_generator(fn, None, int(cplx)), errors


class ProductGraph:
    """A captured product (htg_graph_capture); `launch(stream)` replays it. Keeps
    the H-matrix alive; capture again after a storage or layout change."""

    def __init__(self, H: HMatrix, d_x: int, d_y: int, mu: int, stream: int):
        self._H = H
        g = C.c_void_p()
        capi.check(capi.lib().htg_graph_capture(H._handle(), C.c_void_p(d_x), C.c_void_p(d_y), mu,
                                                C.c_void_p(stream), C.byref(g)))
        self._g = g

    def launch(self, stream: int) -> None:
        capi.check(capi.lib().htg_graph_launch(self._g, C.c_void_p(stream)))

    def __del__(self):
        try:
            if self._g is not None and self._g.value:
                capi.lib().htg_graph_destroy(self._g)
        except Exception:
            pass
        self._g = None
