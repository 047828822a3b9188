// Device data structures and kernel launchers of the H-matrix hot path.
#pragma once

#include "../../include/htool_amd.h"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace htg {

// Kernel family of the generator, evaluated on the device
// (replaces VirtualGenerator::copy_submatrix for the built-in kernels,
// reference: include/htool/types/virtual_generator.hpp:22).
struct KernelDesc {
    int kind;          // HTG_KERNEL_*
    int same;          // target and source point sets are the same set
    double delta;      // regularisation
    double diag;       // value of the singular kernels on their diagonal
    double wavenumber; // Helmholtz
};

inline bool kernel_is_complex(int kind) {
    return kind == HTG_KERNEL_HELMHOLTZ || kind == HTG_KERNEL_COMPLEX_TEST || kind == HTG_KERNEL_HERMITIAN_TEST;
}
// Kernel families whose matrix is Hermitian (A(j, i) = conj A(i, j), bitwise
// as evaluated) on one point set; every real family is symmetric.
inline bool kernel_is_hermitian(int kind) { return !kernel_is_complex(kind) || kind == HTG_KERNEL_HERMITIAN_TEST; }

// complex<double> in device memory and registers: re then im.
struct cdbl {
    double re, im;
};

// Cluster-ordered points (structure of arrays) and the cluster -> original map.
struct DevPoints {
    const double *x, *y, *z;
    const int *perm;
};

// One dense-generation tile: rows [row0, row0+R) x cols [col0, col0+ncols) of the
// matrix in cluster order, written row-major with row stride ld.
struct FillTile {
    uint64_t off;
    int row0, R, col0, ncols, ld;
    int tr; // 1: a tile of the transpose, entry (i, j) = G(target col0 + j, source row0 + i)
};

// One block handed to the batched ACA kernel.
struct AcaBlock {
    int off_t, M, off_s, N; // rows [off_t, off_t+M), cols [off_s, off_s+N) in cluster order
    int swap;               // sympartialACA with t.offset < s.offset: rows and columns exchange roles
    int cap;                // room for this many cross terms in the workspace
    uint64_t ws1, ws2;      // workspace offsets: side-1 vectors (n1*cap), side-2 vectors (n2*cap)
    double ctr[3];          // centre of the cluster whose points choose the first pivot
    int64_t mn;             // right side of the q (M + N) > MN exit: M*N, or the reference's 32-bit
                            // wrap of it (aca_size_bound)
};

// Right side of partialACA's "next rank not advantageous" test q (M + N) > M N
// (partialACA.hpp:74, sympartialACA.hpp:98). The reference evaluates it in
// int, so for M N > 2^31 - 1 its product wraps modulo 2^32 (to a negative
// value below 2^32: the test then fires at q = 1 and the block is split);
// int32_wrap reproduces that, otherwise the exact 64-bit product is used.
inline int64_t aca_size_bound(int M, int N, bool int32_wrap) {
    const int64_t mn = (int64_t)M * N;
    return int32_wrap ? (int64_t)(int32_t)(uint32_t)(uint64_t)mn : mn;
}
// Whether the reference's int test and the 64-bit one part ways for a block
// whose ACA ran k >= 1 cross terms: the 64-bit test let every q <= k pass,
// the 32-bit one fails at the latest at q = k.
inline bool aca_int32_wrap_differs(int M, int N, int k) {
    return k >= 1 && (int64_t)k * (M + N) > aca_size_bound(M, N, true) && (int64_t)k * (M + N) <= (int64_t)M * N;
}

// Column chunk of a low-rank leaf's V (k x n, row-major, row stride ld). Each
// stored row covers the global columns [oj & ~(a-1), round_a(oj + n)), zero
// outside the leaf (a = 4 for class 0, 8 for class 1), so stored column c
// multiplies z entry (oj & ~(a-1)) + c and both are 16-byte aligned; the row
// stride is that extent rounded up to 32 (class 0) or 64 (class 1) entries, so
// every stored row starts on a 128-byte line. A chunk is k sums over ncols (a
// multiple of a) stored columns starting at z entry x_off; chunks are grouped
// per wavefront.
struct VTile {
    uint64_t v_off; // first entry of the chunk in the V arena
    int k, ncols, x_off, ld;
    int t_index;   // z position (in RHS vectors) of the leaf's first t entry
    int part_off;  // partial sums of this chunk (in RHS vectors), when nparts > 1
    int part_base; // partial sums of chunk 0 of this leaf
    int nparts;    // number of column chunks of this leaf
    int cls;       // precision class of the rows (which V arena)
    int pad;
};

// One low-rank leaf handed to the recompression kernel: U (M x k column-major
// at u_off) and V (k x N row-major at v_off) in the compressed store.
struct RcLeaf {
    uint64_t u_off, v_off;
    int M, N, k, pad;
};
constexpr int RC_KMAX = 48;   // largest rank the recompression kernel handles
constexpr int RC_KMAX_C = 36; // the same for complex leaves (twice the LDS per entry)

// Copy of one compressed leaf: U (M x k column-major at src_u) into U tiles, V
// (k rows of N entries, row strides src_ldv / dst_ldv) from src_v to dst_v.
struct LrPack {
    uint64_t src_u, src_v, dst_v;
    int M, N, k, src_ldv, dst_ldv;
    int piece_first; // first U tile in the U-tile list
    int npieces;
    int conj;        // complex entries: store the conjugates (mirrored copies of Hermitian storage)
};

// U rows [row0, row0+R) of a leaf: entry (i, r) goes to off + i*si + r*sr.
struct UTileDst {
    uint64_t off;
    int row0, R, si, sr;
};

// ---- launchers (kernels.hip) -------------------------------------------------
void launch_dense_fill(const KernelDesc &k, DevPoints T, DevPoints S, const FillTile *tiles, int ntiles,
                       double *arena, hipStream_t st);
void launch_aca(const KernelDesc &k, DevPoints T, DevPoints S, const AcaBlock *blocks, int nblocks, double eps,
                int reqrank, double *ws, unsigned char *vis_ws, const uint64_t *vis_off, int *ranks,
                hipStream_t st);
void launch_lr_pack(const LrPack *packs, int npacks, const UTileDst *utiles, const double *ws, double *arena,
                    double *varena, hipStream_t st);
// In place: U <- U A, V <- B V with rank r <= k; rank[b] = r, ok[b] = 1 when
// recompressed (0: kept as is), sigma[b*RC_KMAX + j] = the kept singular values.
void launch_recompress(const RcLeaf *leaves, int nleaves, double *store, double tol, int *rank, int *ok,
                       double *sigma, hipStream_t st);
void launch_recompress_c(const RcLeaf *leaves, int nleaves, cdbl *store, double tol, int *rank, int *ok,
                         double *sigma, hipStream_t st);

// Product phases. TS = storage type, TC = compute type inside a tile. z holds
// the cluster-ordered x (nc RHS vectors) followed by t (one RHS vector per
// low-rank term), both in TC.
template <typename TC>
void launch_gather_x(const double *x, const int *perm_s, int nc, int tpad, int mu_base, int MU, TC *z,
                     hipStream_t st);
// 32-RHS gather into the blocked layout of the matrix-core passes; iperm_s is
// the inverse of perm_s, stage holds nc x 32 floats.
void launch_gather_xc32(const double *x, const int *iperm_s, int nc, int tpad, int mu_base, float *stage, float *z,
                        hipStream_t st);
void launch_gather_xc64(const double *x, const int *iperm_s, int nc, int tpad, int mu_base, int MU, double *z,
                        hipStream_t st);
void launch_gather_x64(const double *x, const int *iperm_s, int nc, int tpad, int mu_base, double *z, hipStream_t st);
// xmax (tpad / 32 + 1 floats): the largest |x| of each 32-entry block of z.
void launch_gather_x32(const double *x, const int *iperm_s, int nc, int tpad, int mu_base, float *stage, float *z,
                       float *xmax, unsigned *tmax, hipStream_t st);
// Both precision classes in one launch: chunk d reads varena0 or varena1 by d.cls.
template <typename TS0, typename TS1, typename TC>
void launch_vx(const int *vg_off, int ngroups, const VTile *tiles, const VTile *multi_tiles, int nmulti,
               const TS0 *varena0, const TS1 *varena1, TC *z, double *part, const double *tscale, int MU,
               hipStream_t st);

// 32-RHS V phase by row ranges (one workgroup each: at most RR_ITEM_ROWS rows of
// a panel's class over a range of its columns, a multiple of RR_ITEM_CHUNK32
// (FP32 arithmetic) or RR_ITEM_CHUNK64 (FP64) columns long except at the
// panel's end), then the partial sums of ranges that split a panel's columns.
// xmax: the largest |x| of each 32-entry block of z (launch_gather_x32).
// FP32 arithmetic: x is read from the blocked z, the t entries are written to
// zil (RHS-interleaved z, FP32, (tpad + t) x 32) and the largest class-1 |t|
// to tmax (TMAX_SLOTS slots, cleared by launch_gather_x32).
template <typename TS0, typename TS1, typename TC>
void launch_vx_rr(const VTile *items, int nitems, const VTile *multi_tiles, int nmulti, const TS0 *varena0,
                  const TS1 *varena1, TC *z, float *zil, const float *xmax, unsigned *tmax, double *part,
                  const double *tscale, hipStream_t st);
constexpr int RR_ITEM_ROWS    = 64;
// The largest class-1 |t| is kept in this many slots (atomic maxima spread
// over addresses; the row phase takes the largest): tmax holds TMAX_SLOTS.
constexpr int TMAX_SLOTS = 256;
constexpr int RR_ITEM_CHUNK32 = 128;
constexpr int RR_ITEM_CHUNK64 = 64;

// One precision class of the row blocks: per row block its first column in
// colidx and the offset of its row-major block (row stride = columns rounded
// up to 4) in the class's arena.
struct RowClass {
    const int *rb_col0;
    const uint64_t *rb_aoff;
    const int *colidx;
    int align;  // columns read: the row's columns rounded up to a multiple of align (4, or 8 for class 1)
    int salign; // row stride: the columns rounded up to a multiple of salign (rows start on 128-byte lines)
};
template <typename TS0, typename TS1, typename TC>
void launch_rows(const int *rb_order, const int *rb_row0, int nrb, RowClass c0, const TS0 *arena0, RowClass c1,
                 const TS1 *arena1, const TC *z, const int *perm_t, int nr, int mu_base, int MU, double *y,
                 hipStream_t st);

// 32-RHS row phase by row-block parts of at most RR_ITEM_ROWS rows: items[b] =
// row block | part << 24. FP32 arithmetic reads zil, FP64 the interleaved z.
template <typename TS0, typename TS1, typename TC>
void launch_rows_rr(const int *items, int nitems, const int *rb_row0, RowClass c0, const TS0 *arena0,
                    RowClass c1, const TS1 *arena1, const TC *z, const float *zil, const unsigned *tmax,
                    const int *perm_t, int nr, int mu_base, double *y, hipStream_t st);

// Complex FP64 product, 32 right-hand sides (z interleaved, (re, im) pairs): the
// V phase by row ranges (items of the FP64 cut, multi = their partial sums),
// an event recorded when it is done (if given), then the row phase by
// row-block parts.
void launch_rrc64(const VTile *items, int nitems, const VTile *multi_tiles, int nmulti, const int *rb_items,
                  int nrb_items, const int *rb_row0, RowClass c0, const double *arena0, RowClass c1,
                  const double *arena1, const double *varena0, const double *varena1, double *z, double *part,
                  const double *tscale, const int *perm_t, int nr, int mu_base, double *y, hipStream_t st,
                  hipEvent_t ev_vx_done);

// Complex product in FP32 arithmetic, muc = 32 or 64 right-hand sides: the
// gather into zc (RHS-interleaved (re, im) FP32 pairs, (tpad + t) x muc), the
// per-block |x| maxima, the V phase by row ranges (items of the FP64 cut), its
// partial sums (part: muc complex entries per row), an event after each of the
// gather and the V phase (if given), then the row phase by row-block parts.
// tmax holds TMAX_SLOTS. With two storage types, items [0, nitems0) are the
// class-0 items and the rest class 1 (one V launch each).
template <typename TS0, typename TS1>
void launch_rrc32(int muc, const double *x, const int *iperm_s, int nc, int tpad, int mu_base, float *zc, float *xmax,
                  unsigned *tmax, const VTile *items, int nitems, int nitems0, const VTile *multi_tiles, int nmulti,
                  const int *rb_items, int nrb_items, const int *rb_row0, RowClass c0, const TS0 *arena0, RowClass c1,
                  const TS1 *arena1, const TS0 *varena0, const TS1 *varena1, double *part, const double *tscale,
                  const int *perm_t, int nr, double *y, hipStream_t st, hipEvent_t ev_gather, hipEvent_t ev_vx);

// FP64 master copy -> storage type (round to nearest).
template <typename TS>
void launch_convert(const double *src, TS *dst, uint64_t n, hipStream_t st);
// Cluster-ordered, RHS-interleaved (j * mu + s) <-> original, column-major (perm[j] + n * s).
void launch_to_original(const double *in, const int *perm, int n, int mu, int e, double *out, hipStream_t st);
void launch_to_cluster(const double *in, const int *perm, int n, int mu, int e, double *out, hipStream_t st);

// Multi-workgroup partial ACA of very large blocks: per-block state between
// the launches of one cross term (cluster-order row/column indices local to the
// block, side 1 = the side that picks pivot rows, as in AcaBlock).
constexpr int ACAW_CHUNK = 2048; // residual entries per workgroup
struct AcaWide {
    int I1, I2;       // current pivot row (side 1) and column (side 2)
    int q;            // cross terms accepted so far
    int active;       // 1 while the block needs another cross term
    int rank;         // result once inactive: rank, -1 (not compressible) or -3 (workspace full)
    int nch1, nch2;   // chunks of side 1 / side 2
    int task1, task2; // first chunk task of each side in the partial-result arrays
    int pad;
    double frob, aux; // Frobenius estimate, |u|^2 |v|^2 of the last cross term
    double gre, gim;  // reciprocal of the last pivot
};
struct AcawTask {
    int block, chunk;
};
// One launch sequence per stage; `E` selects real (double) or complex (cdbl).
void launch_acaw_init(const KernelDesc &k, DevPoints T, DevPoints S, const AcaBlock *blocks, AcaWide *st,
                      int nblocks, int reqrank, double eps, unsigned char *vis_ws, const uint64_t *vis_off,
                      hipStream_t s);
void launch_acaw_term(const KernelDesc &k, DevPoints T, DevPoints S, const AcaBlock *blocks, AcaWide *st,
                      int nblocks, const AcawTask *tasks1, int ntasks1, const AcawTask *tasks2, int ntasks2,
                      int reqrank, double eps, void *ws, bool cplx, unsigned char *vis_ws, const uint64_t *vis_off,
                      double *pv, int *pi, double *dd, void *xs, hipStream_t s);
// Half a cross term (row: residual row + pivot column; else residual column +
// next pivot row and stopping test) with the generator values of every active
// block supplied by the host in gv (block k's vector at gv_off[k]).
void launch_acaw_half(bool row, DevPoints T, DevPoints S, const AcaBlock *blocks, AcaWide *st, int nblocks,
                      const AcawTask *tasks, int ntasks, int slot0, int reqrank, double eps, void *ws, bool cplx,
                      unsigned char *vis_ws, const uint64_t *vis_off, double *pv, int *pi, double *dd, void *xs,
                      const void *gv, const uint64_t *gv_off, hipStream_t s);

// Complex kernel families.
void launch_dense_fill_c(const KernelDesc &k, DevPoints T, DevPoints S, const FillTile *tiles, int ntiles, cdbl *arena,
                         hipStream_t st);
void launch_aca_c(const KernelDesc &k, DevPoints T, DevPoints S, const AcaBlock *blocks, int nblocks, double eps,
                  int reqrank, cdbl *ws, unsigned char *vis_ws, const uint64_t *vis_off, int *ranks, hipStream_t st);
void launch_lr_pack_c(const LrPack *packs, int npacks, const UTileDst *utiles, const cdbl *ws, cdbl *arena,
                      cdbl *varena, hipStream_t st);

// Per-term scaling of low-rank leaves in a store (real: E = double, complex:
// E = cdbl): term j of leaf l becomes (U[:,j] / |U[:,j]|) sig_j (V[j,:] / |V[j,:]|),
// sig_j = |U[:,j]| |V[j,:]| written at sig[sig_off[l] + j].
struct ScaleLeaf {
    uint64_t src_u, src_v; // U (M x k col-major) and V (k x N row-major) in the store
    int M, N, k;
    int sig_off;
};
void launch_term_scale(const ScaleLeaf *leaves, int nleaves, double *store, double *sig, hipStream_t st);
void launch_term_scale_c(const ScaleLeaf *leaves, int nleaves, cdbl *store, double *sig, hipStream_t st);

// Complex products: entries stored as (re, im) pairs of TS in the real
// product's layout; R = float or double arithmetic; z holds R pairs.
template <typename R>
void launch_gather_xc(const double *x, const int *perm_s, int nc, int tpad, int mu_base, int MU, R *z,
                      hipStream_t st);
template <typename TS0, typename TS1, typename R>
void launch_vx_c(const int *vg_off, int ngroups, const VTile *tiles, const VTile *multi_tiles, int nmulti,
                 const TS0 *varena0, const TS1 *varena1, R *z, double *part, const double *tscale, int MU,
                 hipStream_t st);
template <typename TS0, typename TS1, typename R>
void launch_rows_c(const int *rb_order, const int *rb_row0, int nrb, RowClass c0, const TS0 *arena0, RowClass c1,
                   const TS1 *arena1, const R *z, const int *perm_t, int nr, int mu_base, int MU, double *y,
                   hipStream_t st);

} // namespace htg
