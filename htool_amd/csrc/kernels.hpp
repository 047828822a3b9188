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
// matrix in cluster order, written column-major with leading dimension R.
struct FillTile {
    uint64_t off;
    int row0, R, col0, ncols;
};

// One block handed to the batched ACA kernel.
struct AcaBlock {
    int off_t, M, off_s, N; // rows [off_t, off_t+M), cols [off_s, off_s+N) in cluster order
    int swap;               // sympartialACA with t.offset < s.offset: rows and columns exchange roles
    int cap;                // room for this many cross terms in the workspace
    uint64_t ws1, ws2;      // workspace offsets: side-1 vectors (n1*cap), side-2 vectors (n2*cap)
    double ctr[3];          // centre of the cluster whose points choose the first pivot
};

// Row-block piece of the product: a column-major R x ncols tile (R = rows of the
// row block) multiplied by a vector segment: x (dense leaf) or t (low-rank leaf).
struct Piece {
    uint64_t a_off;
    int ncols;
    int vec_off; // dense: first cluster column; low-rank: first entry of t
    int kind;    // 0 dense, 1 low-rank
    int pad;
};

// Column chunk of a low-rank leaf's V (k x n, row-major, leading dimension ld):
// computes k partial sums of V x over columns [x_off, x_off+ncols).
struct VTile {
    uint64_t v_off;
    int k, ld, ncols, x_off;
    int part_off; // partial sums of this tile: part[part_off + r*MU + s]
    int leaf;     // low-rank leaf id (completion counter index)
    int nparts;   // number of column chunks of this leaf
    int part_base;// partial sums of chunk 0 of this leaf
    int t_off;    // t[t_off + r*MU + s]
    int pad;
};

struct LrPack {
    uint64_t src_u, src_v; // workspace offsets of the U columns (len M) and V rows (len N)
    uint64_t dst_v;        // V arena offset (k x N row-major)
    int M, N, k, oi;       // oi: first cluster row of the leaf
    int piece_first;       // first U tile in the U-tile list
    int npieces;
};

struct UTileDst {
    uint64_t off; // destination in the tile arena
    int row0;     // first row of the tile relative to the leaf
    int R;
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

// Product phases. TS = storage type, TC = compute type inside a leaf tile.
template <typename TS, typename TC>
void launch_gather_x(const double *x, const int *perm_s, int nc, int mu_total, int mu_base, int MU, TC *xc,
                     hipStream_t st);
template <typename TS, typename TC>
void launch_lr_vx(const VTile *tiles, int ntiles, const TS *varena, const TC *xc, double *part, double *t,
                  unsigned *counters, int MU, hipStream_t st);
template <typename TS, typename TC>
void launch_row_apply(const int *rb_row0, const int *piece_off, const Piece *pieces, int nrb, const TS *arena,
                      const TC *xc, const double *t, const int *perm_t, int nr, int mu_base, int MU, double *y,
                      hipStream_t st);

} // namespace htg
