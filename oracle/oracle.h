/* CPU oracle: an FP64 restatement of the reference's H-matrix hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load this library, and only as the checker
 * (or as the timed CPU restatement in the baseline). The product library
 * (htool_amd/) never links, loads or calls it.
 *
 * Every function follows a reference file (paths relative to the reference
 * root); oracle.cpp cites file:line for each step.
 *
 * Pinning (what checks each function beyond retyping it):
 *   orc_kernel_block       closed forms; bitwise vs the reference driver's dense leaves
 *   orc_cluster_build      permutation bitwise vs the reference driver; invariants
 *                          (contiguous sons, every point once)
 *   orc_aca_size_bound     the reference driver's rank -1 / compressed outcome on blocks
 *                          either side of M N = 2^31 (46340^2, 46341^2)
 *   orc_aca                ranks/U/V vs the reference driver; cross interpolation
 *                          property; space-saving bands of the reference's lrmat tests;
 *                          Eckart-Young floor vs numpy SVD
 *   orc_hmat_build         leaf list identical to the reference driver; block cover
 *                          exactly once; admissibility recomputed per far leaf
 *   orc_hmat_mvprod /      vs the reference driver's y on the same tree (1e-12);
 *   orc_leaf_mvprod        vs numpy dense reconstruction; vs brute-force dense rows
 *   orc_dense_rows         closed form on tiny inputs (numpy)
 */
#ifndef HTOOL_AMD_ORACLE_H
#define HTOOL_AMD_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_hmat orc_hmat;

/* Arithmetic of partialACA's q (M + N) > M N exit (partialACA.hpp:74,
 * sympartialACA.hpp:98) for every later orc_aca / orc_hmat_build: 0 (default)
 * exact 64-bit; 1 the reference's int arithmetic, where M N > 2^31 - 1 wraps
 * modulo 2^32. orc_aca_size_bound returns the right side in use. */
void orc_set_ref_int_overflow(int on);
/* mintargetdepth / minsourcedepth of the block trees of later builds (default 0). */
void orc_set_min_depths(int mintargetdepth, int minsourcedepth);
long long orc_aca_size_bound(int M, int N);

/* One kernel entry family: 0 = 1/(4 pi r) (diag on i==j when same points),
 * 1 = 1/(delta+r), 2 = 1/(delta+4 pi r), 3 = 1/(4 pi sqrt(delta+r^2)), 6 = 0. */
void orc_kernel_block(int kernel, double delta, double diag, int same_points, const double *xt,
                      const double *xs, int M, int N, const int *rows, const int *cols, double *out);

/* PCA bisection cluster tree (geometric splitting, one MPI rank).
 * nodes_int: per node {offset, size, depth, rank, nb_sons, first_son} (6 ints),
 * nodes_dbl: per node {ctr0, ctr1, ctr2, rad}; capacity 2n+2 nodes. Returns node count. */
int orc_cluster_build(int n, const double *x, int minclustersize, int *perm, int *nodes_int, double *nodes_dbl);

/* Leaf compression of one block (partialACA when sym == 0, sympartialACA when sym == 1).
 * rows/cols are original indices. U is M x k column-major, V is k x N column-major
 * (V(r, j) at r + k*j), both with room for `cap` terms. Returns the rank (-1 = incompressible,
 * -2 = cap too small). */
int orc_aca(int sym, double eps, int M, int N, const int *rows, const int *cols, int kernel, double delta,
            double diag, int same_points, const double *xt, const double *xs, const double *ctr_t, int off_t,
            const double *ctr_s, int off_s, int reqrank, int cap, double *U, double *V);

orc_hmat *orc_hmat_build(int nt, const double *xt, int ns, const double *xs, int kernel, double delta, double diag,
                         int minclustersize, double eta, double eps, int sym_aca, int reqrank, double maxblocksize,
                         int uplo); /* 'L' / 'U': symmetric storage (one triangle, xs = NULL); else 'N' */
int orc_hmat_nleaves(const orc_hmat *h);
void orc_hmat_leaves(const orc_hmat *h, int *out5); /* off_i, m, off_j, n, rank(-1 dense); sorted by (off_i, off_j) */
long long orc_hmat_data_size(const orc_hmat *h);
void orc_hmat_data(const orc_hmat *h, double *out); /* dense m*n col-major | U m*k col-major then V k*n col-major */
void orc_hmat_perms(const orc_hmat *h, int *perm_t, int *perm_s);
int orc_hmat_false_positives(const orc_hmat *h);
void orc_hmat_mvprod(const orc_hmat *h, const double *x, double *y, int mu);
void orc_hmat_free(orc_hmat *h);
/* Complex kernel families (kernel 4: e^{ikr}/(4 pi r) with k passed as `delta`;
 * kernel 5: (1 + i)/(4 pi r)) build complex H-matrices: orc_hmat_data writes
 * (re, im) pairs (2 x orc_hmat_data_size doubles) and orc_hmat_mvprod takes and
 * returns complex vectors as (re, im) pairs. */
int orc_hmat_is_complex(const orc_hmat *h);

/* y = H x over an explicit leaf list with the data layout of orc_hmat_data.
 * x: nc x mu column-major, y: nr x mu column-major, both in original ordering. */
void orc_leaf_mvprod(int nr, int nc, const int *perm_t, const int *perm_s, int nleaves, const int *leaves5,
                     const double *data, const double *x, double *y, int mu);

/* The same over complex leaf data and vectors, (re, im) pairs. */
void orc_leaf_mvprod_c(int nr, int nc, const int *perm_t, const int *perm_s, int nleaves, const int *leaves5,
                       const double *data, const double *x, double *y, int mu);
/* Symmetric storage ('L' / 'U', square, one permutation): the kept leaves, the
 * transposes of the off-diagonal ones (no conjugation) and symv on the kept
 * triangle of the diagonal leaves; cplx: complex data and vectors. */
void orc_leaf_mvprod_sym(int n, const int *perm, int nleaves, const int *leaves5, const double *data,
                         const double *x, double *y, int mu, int uplo, int cplx);
/* Hermitian storage ('H', complex): the kept leaves, the conjugate transposes
 * of the off-diagonal ones, hemv on the kept triangle of the diagonal leaves
 * (the real part of their diagonal entries), hmatrix.hpp:864-904. */
void orc_leaf_mvprod_herm(int n, const int *perm, int nleaves, const int *leaves5, const double *data,
                          const double *x, double *y, int mu, int uplo);

/* Brute-force dense rows, complex kernels, (re, im) pairs. */
void orc_dense_rows_c(int kernel, double wavenumber, double diag, int same_points, const double *xt, const double *xs,
                      int ns, int nrows, const int *rows, const double *x, int mu, double *out);

/* Brute-force dense rows: out[r + nrows*c] = sum_j A(rows[r], j) x[j + ns*c]. */
void orc_dense_rows(int kernel, double delta, double diag, int same_points, const double *xt, const double *xs,
                    int ns, int nrows, const int *rows, const double *x, int mu, double *out);

/* Rank predictor network of the library (8 -> 64 -> 64 -> 1 perceptron, GELU
 * x Phi(x), no counterpart in the reference), restated in FP64. params: W1
 * [64][8], b1 [64], W2 [64][64], b2 [64], w3 [64], b3 (4801 values); inputs
 * (feat - mean) / stdv. orc_mlp_loss_grad: loss (1/n) sum 0.5 (out - target)^2
 * and its gradient (4801 values); orc_mlp_predict: the outputs. */
void orc_mlp_loss_grad(const double *params, const double *mean, const double *stdv, const float *feat,
                       const float *target, int n, double *loss, double *grad);
void orc_mlp_predict(const double *params, const double *mean, const double *stdv, const float *feat, int n,
                     double *out);

#ifdef __cplusplus
}
#endif
#endif
