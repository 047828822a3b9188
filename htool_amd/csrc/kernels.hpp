// Device data structures and kernel launchers of the H-matrix hot path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace htg {

// Kernel family of the generator, evaluated on the device
// (replaces VirtualGenerator::copy_submatrix for the built-in kernels,
// reference: include/htool/types/virtual_generator.hpp:22).
struct KernelDesc {
    int kind;     // HTG_KERNEL_*
    int same;     // target and source point sets are the same set
    double delta; // regularisation
    double diag;  // value of the singular Laplace kernel on its diagonal
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
};

// One block handed to the batched ACA kernel.
struct AcaBlock {
    int off_t, M, off_s, N; // rows [off_t, off_t+M), cols [off_s, off_s+N) in cluster order
    int swap;               // sympartialACA with t.offset < s.offset: rows and columns exchange roles
    int cap;                // room for this many cross terms in the workspace
    uint64_t ws1, ws2;      // workspace offsets: side-1 vectors (n1*cap), side-2 vectors (n2*cap)
    double ctr[3];          // centre of the cluster whose points choose the first pivot
};

// Column chunk of a low-rank leaf's V (k x n, row-major, row stride ld). Each
// stored row covers the global columns [oj & ~3, round4(oj + n)), zero outside
// the leaf, so stored column c multiplies z entry (oj & ~3) + c and both are
// 16-byte aligned. A chunk is k sums over ncols (a multiple of 4) stored
// columns starting at z entry x_off; chunks are grouped per wavefront.
struct VTile {
    uint64_t v_off; // first entry of the chunk in the V arena
    int k, ncols, x_off, ld;
    int t_index;   // z position (in RHS vectors) of the leaf's first t entry
    int part_off;  // partial sums of this chunk (in RHS vectors), when nparts > 1
    int part_base; // partial sums of chunk 0 of this leaf
    int nparts;    // number of column chunks of this leaf
};

// Copy of one compressed leaf: U (M x k column-major at src_u) into U tiles, V
// (k rows of N entries, row strides src_ldv / dst_ldv) from src_v to dst_v.
struct LrPack {
    uint64_t src_u, src_v, dst_v;
    int M, N, k, src_ldv, dst_ldv;
    int piece_first; // first U tile in the U-tile list
    int npieces;
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
void launch_convert_f32(const double *src, float *dst, uint64_t n, hipStream_t st);

// Product phases. TS = storage type, TC = compute type inside a tile. z holds
// the cluster-ordered x (nc RHS vectors) followed by t (one RHS vector per
// low-rank term), both in TC.
template <typename TC>
void launch_gather_x(const double *x, const int *perm_s, int nc, int mu_base, int MU, TC *z, hipStream_t st);
template <typename TS, typename TC>
void launch_vx(const int *vg_off, int ngroups, const VTile *tiles, const int *multi_first, int nmulti,
               const TS *varena, TC *z, double *part, int MU, hipStream_t st);
template <typename TS, typename TC>
void launch_rows(const int *rb_order, const int *rb_row0, const int *rb_col0, const uint64_t *rb_aoff, int nrb,
                 const int *colidx, const TS *arena, const TC *z, const int *perm_t, int nr, int mu_base, int MU,
                 double *y, hipStream_t st);

} // namespace htg
