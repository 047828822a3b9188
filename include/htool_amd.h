/* htool_amd: MI355X-native H-matrix product and partial-pivot ACA compression.
 *
 * C ABI of libhtool_amd.so. Plain pointers and sizes only; no exception ever
 * crosses it. Every entry point returns an htg_status (HTG_OK == 0) and leaves
 * a message for htg_last_error() on failure.
 *
 * Layout conventions follow the reference (Htool 0.8.1); for the complex kernel
 * families every vector and leaf entry is a complex<double> (two doubles, re
 * then im) and the sizes below count complex entries:
 *  - points: interleaved x[3*i + d], like HMatrix::build(A, xt, xs)
 *    (include/htool/types/hmatrix.hpp:285);
 *  - products: x is nc x mu and y is nr x mu, column-major, original (not
 *    cluster) ordering, like HMatrix::mvprod_global_to_global
 *    (include/htool/types/hmatrix.hpp:1073-1182);
 *  - leaves: (off_i, m, off_j, n, rank) in cluster ordering, rank -1 for a dense
 *    leaf, like HMatrix::save_plot (include/htool/types/hmatrix.hpp:1695-1712);
 *    leaf data: dense m x n column-major (SubMatrix, include/htool/types/matrix.hpp:572),
 *    low-rank U (m x k column-major) then V (k x n column-major, V(r,j) at r + k*j)
 *    (LowRankMatrix, include/htool/lrmat/lrmat.hpp:47-53).
 */
#ifndef HTOOL_AMD_H
#define HTOOL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum htg_status {
    HTG_OK               = 0,
    HTG_ERR_INVALID      = 1, /* bad argument (the reference throws std::invalid_argument) */
    HTG_ERR_HIP          = 2, /* HIP runtime error, or no usable GPU */
    HTG_ERR_OOM          = 3, /* device or host allocation failed */
    HTG_ERR_UNSUPPORTED  = 4, /* kernel family / mode not implemented */
    HTG_ERR_RANK_CAP     = 5  /* a block needed more ACA terms than the library allows */
} htg_status;

/* Kernel families evaluated on the device; they stand in for a user
 * VirtualGenerator::copy_submatrix (include/htool/types/virtual_generator.hpp:22). */
typedef enum htg_kernel {
    HTG_KERNEL_LAPLACE        = 0, /* 1/(4 pi r); `diag` on the diagonal when the point sets coincide */
    HTG_KERNEL_INV_DELTA_R    = 1, /* 1/(delta + r) */
    HTG_KERNEL_INV_DELTA_4PIR = 2, /* 1/(delta + 4 pi r): GeneratorTestDoubleSymmetric,
                                      include/htool/testing/generator_test.hpp:135-145 */
    HTG_KERNEL_SMOOTH         = 3, /* 1/(4 pi sqrt(delta + r^2)): examples/smallest_example.cpp:16-27 */
    /* Complex families: the H-matrix, x and y are complex<double> (interleaved re, im). */
    HTG_KERNEL_HELMHOLTZ      = 4, /* e^{i k r}/(4 pi r), k = `wavenumber`; `diag` on the diagonal when the
                                      point sets coincide */
    HTG_KERNEL_COMPLEX_TEST   = 5, /* (1 + i)/(4 pi r): GeneratorTestComplex,
                                      include/htool/testing/generator_test.hpp:110-120 */
    HTG_KERNEL_ZERO           = 6, /* real, every entry 0: GeneratorTestZero,
                                      tests/functional_tests/types/virtual_hmatrix/hmatrix/test_hmat_zero.hpp:18-65 */
    HTG_KERNEL_HERMITIAN_TEST = 7  /* complex, (1 + sign(x_i - x_j) i)/(1e-5 + 4 pi r) (x_i the first coordinate):
                                      GeneratorTestComplexHermitian, include/htool/testing/generator_test.hpp:165-180 */
} htg_kernel;

typedef enum htg_compressor {
    HTG_PARTIAL_ACA     = 0, /* include/htool/lrmat/partialACA.hpp */
    HTG_SYM_PARTIAL_ACA = 1  /* include/htool/lrmat/sympartialACA.hpp (the reference default) */
} htg_compressor;

/* Storage/arithmetic of the leaf products. Leaf entries are in two precision
 * classes: class 0 (dense leaves and the leading low-rank terms) and class 1
 * (the low-rank terms assigned to 16-bit storage at build time, see
 * htg_params.lowprec_tol / lowprec_all; empty otherwise). Arithmetic inside a
 * tile is FP32 except in FP64 mode; the sum over tiles and the output are FP64.
 * Multi-RHS products (32 real right-hand sides per pass; complex: 64 per pass
 * in FP32 arithmetic, 32 in FP64) run a 16-bit class on the 16-bit matrix
 * cores, with the FP32 vector entries split into 16-bit terms (FP16: two,
 * under a power-of-two scale per work item or pass; BF16: three) and FP32
 * accumulation. */
typedef enum htg_storage {
    HTG_STORAGE_FP64      = 0, /* class 0 and 1 in FP64 */
    HTG_STORAGE_FP32      = 1, /* class 0 and 1 in FP32 */
    HTG_STORAGE_FP32_FP16 = 2, /* class 0 FP32, class 1 FP16 (per-term scaled) */
    HTG_STORAGE_FP32_BF16 = 3  /* class 0 FP32, class 1 BF16 (per-term scaled) */
} htg_storage;

typedef struct htg_generator {
    int kernel;      /* htg_kernel */
    int same_points; /* 1: source points are the target points (square matrix) */
    double delta;
    double diag;
    double wavenumber; /* HTG_KERNEL_HELMHOLTZ */
} htg_generator;

/* A user generator evaluated on the host, for htg_hmatrix_build_ex: the C form
 * of VirtualGenerator<T>::copy_submatrix(M, N, rows, cols, ptr)
 * (include/htool/types/virtual_generator.hpp:22). It writes ptr[i + M*j] =
 * A(rows[i], cols[j]) (original indices, or cluster positions with
 * htg_params.no_permutation); complex entries are (re, im) pairs. The library
 * calls it from the building thread only, one call at a time: for every dense
 * leaf (m x n) and, per ACA cross term, one row (1 x n) or column (m x 1) of
 * each block being compressed. */
typedef void (*htg_copy_submatrix_fn)(int M, int N, const int *rows, const int *cols, double *ptr, void *ctx);
typedef struct htg_host_generator {
    htg_copy_submatrix_fn copy_submatrix;
    void *ctx;      /* passed back to every call */
    int is_complex; /* 1: complex<double> entries */
} htg_host_generator;

/* A cluster tree given by the caller (the layout htg_cluster_build writes:
 * Cluster::save_cluster's tree, cluster.hpp:433-484, or read_cluster's,
 * 486-583): n points, perm[cluster position] = original index, nnodes nodes
 * with nodes_int 6 per node (offset, size, depth, rank, nb_sons, first_son or
 * -1; node 0 the root, the sons of a node consecutive and covering its range)
 * and nodes_dbl 4 per node (centre x, y, z, radius). The tree of a single
 * process: the rank-0 clusters cover every point. */
typedef struct htg_cluster_tree {
    int n;
    const int *perm;
    int nnodes;
    const int *nodes_int;
    const double *nodes_dbl;
} htg_cluster_tree;

/* Parameters of HMatrix<T>(t, s, epsilon, eta, 'N', 'N', reqrank) plus the
 * setters the hot path reads (include/htool/types/hmatrix.hpp:115, 213-220)
 * and Cluster::set_minclustersize (include/htool/clustering/cluster.hpp:334). */
typedef struct htg_predictor htg_predictor;

typedef struct htg_params {
    double epsilon;     /* ACA tolerance (default 1e-6) */
    double eta;         /* admissibility parameter (default 10) */
    double maxblocksize;/* largest coalesced dense block, entries (default 1e6) */
    int minclustersize; /* default 10 */
    int compressor;     /* htg_compressor */
    int reqrank;        /* -1: epsilon-driven ranks; >0: fixed rank */
    int lowprec_all;    /* 1: every low-rank term in class 1 (16-bit storage) */
    double recompress_tol; /* > 0: FP64 QR+SVD recompression of each low-rank leaf, dropping
                              singular values whose Frobenius tail is <= tol relative (no
                              counterpart in the reference); 0: keep the ACA factors */
    double lowprec_tol;    /* > 0 (with recompression): trailing terms go to class 1 while
                              2 * 2^-11 * sum(sigma_j) <= lowprec_tol * ||sigma||_2 */
    const htg_predictor *predictor; /* NULL, or a trained rank predictor: presizes each block's
                                       ACA workspace and orders the blocks by expected work
                                       (no counterpart in the reference); the ranks do not change */
    double predictor_margin;        /* room above the prediction: cap = ceil((1 + margin) t) + 2 */
    int aca_wide_min;               /* blocks with M + N >= this run the multi-workgroup ACA (same
                                       factors and ranks); 0: 16384 (real entries) or 4096
                                       (complex entries), < 0: never */
    int symmetry;                   /* 'N' (or 0): general; 'S': symmetric storage, one triangle of
                                       blocks kept (HMatrix(..., 'S', UPLO), hmatrix.hpp:115-122); 'H':
                                       Hermitian storage, the same triangle, the product applying the
                                       conjugate transpose of each kept off-diagonal block
                                       (hmatrix.hpp:864-904); needs HTG_KERNEL_HERMITIAN_TEST or a real
                                       family (HTG_ERR_UNSUPPORTED otherwise) */
    int uplo;                       /* 'L' or 'U' with 'S' / 'H' (the kept triangle); 'N' (or 0) otherwise */
    int delay_dense;                /* 1: dense leaves left zero at build, filled later by
                                       htg_hmatrix_set_dense_blocks (set_delay_dense_computation) */
    int mintargetdepth, minsourcedepth; /* HMatrix::set_mintargetdepth / set_minsourcedepth
                                       (hmatrix.hpp:215-216; blocks.hpp:51): a block is a task or a merged
                                       dense block only when its target / source cluster is at least this
                                       deep (default 0) */
    int no_permutation;             /* 1: set_use_permutation(false) (hmatrix.hpp:218, 306-311): the
                                       generator's indices, the points and the product vectors are all in
                                       the cluster ordering (use with the caller's cluster trees, see
                                       htg_hmatrix_build_ex); 0 (default): original ordering */
    int ref_int_overflow;           /* 0: the ACA's "next rank not advantageous" exit q (M + N) > M N
                                       (partialACA.hpp:74, sympartialACA.hpp:98) in exact 64-bit
                                       arithmetic; 1: in 32-bit int as the reference evaluates it,
                                       where M N > 2^31 - 1 wraps and the block returns rank -1 (and
                                       is split) at once. htg_info.n_aca_int_wrap counts the blocks
                                       on which the two differ */
} htg_params;

typedef struct htg_info {
    int nr, nc;
    int n_dense, n_lowrank, n_false_positive;
    int rank_min, rank_max;
    double rank_mean;
    double compression_ratio; /* nr*nc / stored entries, include/htool/types/hmatrix.hpp:1604-1632 */
    int64_t dense_entries;    /* sum of m*n over dense leaves */
    int64_t lowrank_entries;  /* sum of k*(m+n) over low-rank leaves */
    int64_t lowprec_entries;  /* the part of lowrank_entries stored in class 1 */
    int64_t lowrank_u_entries;/* sum of k*m (the U part of lowrank_entries) */
    int64_t lowprec_u_entries;/* the U part of lowprec_entries */
    int n_row_blocks, n_columns, n_vtiles;
    double build_seconds;     /* block tree + assembly */
    int n_aca;                /* ACA runs of admissible blocks (false positives included) */
    int n_aca_retry;          /* runs that outgrew their first workspace and ran again */
    int64_t aca_ws_entries;   /* sum over first runs of (m + n) x workspace terms */
    int64_t aca_need_entries; /* sum over runs of (m + n) x terms actually needed */
    double aca_seconds;       /* batched ACA (kernels and host scheduling) */
    int is_complex;           /* entries, x and y are complex<double> */
    int n_aca_wide;           /* ACA runs on the multi-workgroup path (htg_params.aca_wide_min) */
    int symmetry, uplo;       /* as built ('N', or 'S' with the kept triangle 'L' / 'U') */
    int64_t mirrored_entries; /* symmetric storage: entries of the transposed copies of the kept
                                 off-diagonal leaves that the product also streams (the counts
                                 above are of the kept leaves, as the reference's) */
    int64_t mirrored_lowprec_entries; /* the part of mirrored_entries in class 1 */
    int64_t nb_mat_vec_prod;  /* products made through this handle, any entry point (the reference's
                                 infos["nb_mat_vec_prod"], hmatrix.hpp:1180) */
    double total_time_mat_vec_prod; /* wall seconds of the host-pointer products among them
                                 (infos["total_time_mat_vec_prod"], hmatrix.hpp:1181); device-pointer
                                 products are enqueued and return at once, adding no time */
    int n_aca_int_wrap;       /* ACA runs of blocks with M N > 2^31 - 1 on which the reference's 32-bit
                                 q (M + N) > M N test and the 64-bit one differ: with ref_int_overflow 0
                                 these blocks are compressed (the reference splits them), with 1 they
                                 are split as in the reference */
} htg_info;

/* Work decomposition of the product (no counterpart in the reference): rows
 * per row block of the near-field/U phase (16..128), target entries per V
 * chunk and per phase-1 wavefront, most chunks per wavefront, and most stacked
 * V rows per chunk (the rank bucket of a source-cluster panel; a panel no
 * wider than vchunk_el takes fewer rows per chunk instead of a column cut).
 * A build starts from {128, 8192, 2048, 4, 16}, with vrows 32 when it puts
 * terms in the 16-bit class. */
typedef struct htg_layout {
    int rb_max;
    int vchunk_el;
    int vgroup_el;
    int vgroup_tiles;
    int vrows;
} htg_layout;

typedef struct htg_hmatrix htg_hmatrix;
typedef struct htg_graph htg_graph;

const char *htg_last_error(void);
const char *htg_status_string(int status);

/* Replaces HMatrix<T>(t, s, eps, eta) + set_compression + build(A, xt, xs)
 * (include/htool/types/hmatrix.hpp:115-122, 220, 285-341) for the built-in
 * kernel families: cluster and block trees on the host, kernel entries and ACA
 * on the current HIP device. xs == NULL (or gen->same_points) uses the target points.
 * The ACA workspace stays allocated per device between builds (it only grows);
 * builds on one device from several host threads run one at a time. */
int htg_hmatrix_build(const htg_params *params, const htg_generator *gen, int nt, const double *xt, int ns,
                      const double *xs, htg_hmatrix **out);

/* HMatrix(t, s, ...) + build(A, xt, xs) in full: the generator is a built-in
 * device family (gen) or a user generator on the host (host_gen; exactly one of
 * the two, the other NULL), and the cluster trees are the caller's (tree_t,
 * tree_s; NULL: built from the points as htg_hmatrix_build does; tree_s is
 * ignored when the point sets coincide). With a host generator every ACA runs
 * the multi-workgroup kernels fed with the generator's rows and columns, so
 * ranks, pivots and factors are those of a device family that computes the
 * same entries, and its dense leaves are filled through it after the ACA (not
 * with delay_dense, which leaves them to htg_hmatrix_set_dense_blocks). Points
 * are needed for the first ACA pivots (partialACA.hpp:42-50). */
int htg_hmatrix_build_ex(const htg_params *params, const htg_generator *gen, const htg_host_generator *host_gen,
                         int nt, const double *xt, int ns, const double *xs, const htg_cluster_tree *tree_t,
                         const htg_cluster_tree *tree_s, htg_hmatrix **out);

/* VirtualGenerator::copy_submatrix(M, N, rows, cols, ptr) of a built-in family
 * (virtual_generator.hpp:22), its entries computed on the device:
 * ptr[i + M*j] = A(rows[i], cols[j]) (original indices into xt / xs; xs NULL
 * or gen->same_points: one point set, where the diagonal rule applies);
 * complex families write (re, im) pairs. For API compatibility (a generator
 * handed to code that calls copy_submatrix); builds evaluate the family inside
 * their own kernels. */
int htg_generator_block(const htg_generator *gen, int nt, const double *xt, int ns, const double *xs, int M, int N,
                        const int *rows, const int *cols, double *out);

/* Builds the product structures from leaves produced elsewhere (dense
 * SubMatrix blocks and LowRankMatrix U, V factors; the layout of the header
 * comment), e.g. the leaves of a reference HMatrix: replaces the leaf list that
 * HMatrix::ComputeBlocks leaves in MyComputedBlocks (include/htool/types/hmatrix.hpp:422-501). */
int htg_hmatrix_from_leaves(int nr, int nc, const int *perm_t, const int *perm_s, int nleaves, const int *leaves5,
                            const double *data, htg_hmatrix **out);

/* Many independent square H-matrices assembled as one: instance i has n[i]
 * points pts[i] (interleaved xyz) and generator gens[i]; the result is the
 * block-diagonal H-matrix diag(H_0, ..., H_{count-1}) of size sum(n), whose
 * product runs every instance's leaves in the same launches. x and y stack the
 * instances' vectors in instance order (original ordering within each). Each
 * instance follows HMatrix::build (include/htool/types/hmatrix.hpp:285-341). */
int htg_hmatrix_build_batch(const htg_params *params, int count, const htg_generator *gens, const int *n,
                            const double *const *pts, htg_hmatrix **out);

int htg_hmatrix_destroy(htg_hmatrix *h);

/* The H-matrix of the transposed matrix (no conjugation), sharing the
 * assembly: its products are HMatrix::mvprod_transp_global_to_global
 * (include/htool/types/virtual_hmatrix.hpp:58; hmatrix.hpp:1276-1455). The
 * leaves of A^T are the leaves of A with rows and columns exchanged (dense
 * blocks transposed, U and V exchanged). Matrices from htg_hmatrix_build /
 * _build_batch only; the result starts in FP64 storage and is destroyed with
 * htg_hmatrix_destroy, independently of h. */
int htg_hmatrix_transpose(const htg_hmatrix *h, htg_hmatrix **out);

/* Product work decomposition: read, or set and rebuild the product layout
 * (matrices from htg_hmatrix_build / _build_batch only). */
int htg_hmatrix_get_layout(const htg_hmatrix *h, htg_layout *layout);
int htg_hmatrix_set_layout(htg_hmatrix *h, const htg_layout *layout);

/* Autotuner: times `reps` products with `mu` right-hand sides in the current
 * storage for the current layout and 24 candidates (rows per row block x
 * (V chunk size, wavefront share) x stacked V rows), times the three fastest
 * again with 2 * reps products, keeps the fastest of that second timing and
 * returns it with its second time (ms per product) and the number of layouts
 * tried. */
int htg_hmatrix_autotune(htg_hmatrix *h, int mu, int reps, htg_layout *best, double *best_ms, int *ncandidates);

/* Captures one product of device buffers d_x -> d_y (mu right-hand sides, the
 * current storage) into a HIP graph on `stream` (a hipStream_t, not NULL);
 * htg_graph_launch replays it on a stream. The graph reads the H-matrix's
 * device data: capture again after htg_hmatrix_set_storage (to another mode),
 * _set_layout or _autotune (both rebuild the layout). htg_graph_launch refuses
 * (HTG_ERR_INVALID) a graph captured before one of these calls, or whose
 * H-matrix was destroyed. htg_hmatrix_set_dense_blocks writes in place: a graph
 * stays valid and sees the new blocks. */
int htg_graph_capture(htg_hmatrix *h, const double *d_x, double *d_y, int mu, void *stream, htg_graph **out);
int htg_graph_launch(htg_graph *g, void *stream);
int htg_graph_destroy(htg_graph *g);

/* Adds (or keeps) the storage variant used by the next products. */
int htg_hmatrix_set_storage(htg_hmatrix *h, int storage);

int htg_hmatrix_get_info(const htg_hmatrix *h, htg_info *info);

/* Leaf list sorted by (off_i, off_j) (MyComputedBlocks order, hmatrix.hpp:494-500),
 * 5 ints per leaf; leaf data in FP64 in the layout of the header comment. */
int htg_hmatrix_nleaves(const htg_hmatrix *h, int *nleaves, int64_t *data_size);
int htg_hmatrix_get_leaves(const htg_hmatrix *h, int *leaves5, double *data /* may be NULL */);
int htg_hmatrix_get_perms(const htg_hmatrix *h, int *perm_t, int *perm_s);

/* HMatrix::build_dense_blocks(VirtualDenseBlocksGenerator) (hmatrix.hpp:404-418,
 * virtual_dense_blocks_generator.hpp:11-17): the caller supplies every kept
 * dense leaf, in htg_hmatrix_get_leaves order, each m x n column-major with
 * entry (i, j) = A(perm_t[off_i + i], perm_s[off_j + j]) (the dense part of the
 * get_leaves data layout; complex: re, im pairs). They replace the dense
 * leaves (zero after a build with delay_dense), are kept through layout
 * changes and reach every storage mode. */
int htg_hmatrix_set_dense_blocks(htg_hmatrix *h, const double *data);

/* The cluster tree the builds use for a point set (host only, no device
 * needed): Cluster<PCAGeometricClustering>::build, cluster.hpp:57-92, with
 * the tree that Cluster::save_cluster writes (cluster.hpp:433-484). Call once
 * with nodes_int == NULL for *nnodes, then with arrays of *nnodes entries:
 * nodes_int 6 per node (offset, size, depth, rank, nb_sons, first_son index or
 * -1), nodes_dbl 4 per node (ctr x, y, z, radius); node 0 is the root, the
 * sons of a node are consecutive. perm (n entries, may be NULL): cluster
 * position -> original point index. */
int htg_cluster_build(int n, const double *x, int minclustersize, int *perm, int *nnodes, int *nodes_int,
                      double *nodes_dbl);

/* y = H x, replaces HMatrix::mvprod_global_to_global(in, out, mu)
 * (include/htool/types/hmatrix.hpp:1073-1182). Host pointers: copies over PCIe. */
int htg_hmatrix_mvprod(htg_hmatrix *h, const double *x, double *y, int mu);

/* Same product on device pointers, enqueued on `stream` (a hipStream_t, NULL
 * for the default stream); returns without synchronising. */
int htg_hmatrix_mvprod_device(htg_hmatrix *h, const double *d_x, double *d_y, int mu, void *stream);

/* y = H x in the cluster ordering with right-hand sides interleaved (entry
 * (j, s) at j * mu + s, as the reference's mymvprod_local_to_local takes and
 * returns them, hmatrix.hpp:979-996, layout 1117-1131; with one process local
 * = global), the product HPDDM's GMV calls (wrappers/wrapper_hpddm.hpp:41-60).
 * Device pointers, enqueued on `stream`; d_work: the caller's device workspace
 * of (nr + nc) * mu entries, or NULL for one allocated for the call (the
 * reference allocates `work` when given none; the call then synchronises the
 * stream before freeing it). */
int htg_hmatrix_mymvprod_local_to_local_device(htg_hmatrix *h, const double *d_in, double *d_out, int mu,
                                               double *d_work, void *stream);
/* The same on host pointers (copies over PCIe; synchronous). */
int htg_hmatrix_mymvprod_local_to_local(htg_hmatrix *h, const double *in, double *out, int mu);

/* HMatrix::mvprod_subrhs(in, out, mu, offset, size, margin)
 * (hmatrix.hpp:1457-1554): in holds the cluster-ordered source entries
 * [offset - margin, offset + size + margin) with interleaved right-hand sides;
 * every leaf whose columns meet [offset, offset + size] is applied, reading its
 * columns from in; out (nr * mu, cluster ordering, interleaved) is overwritten.
 * Host pointers. */
int htg_hmatrix_mvprod_subrhs(htg_hmatrix *h, const double *in, double *out, int mu, int offset, int size,
                              int margin);

/* The dense matrix of the leaves (nr x nc column-major): in the cluster
 * ordering (HMatrix::get_local_dense, hmatrix.hpp:1729-1758) or, with
 * original_order, in the original one (copy_local_dense_perm). Symmetric and
 * Hermitian storage: both triangles. Host memory of nr * nc entries. */
int htg_hmatrix_get_dense(const htg_hmatrix *h, double *out, int original_order);
/* Diagonal entries of a square H-matrix (get_local_diagonal / copy_local_diagonal):
 * cluster ordering, or original with original_order. */
int htg_hmatrix_get_diagonal(const htg_hmatrix *h, double *diag, int original_order);

/* Times each product kernel with HIP events over `reps` products of device
 * buffers on `stream`; ms[0..3] = average ms of (gather, V x, row apply, whole product). */
int htg_hmatrix_profile(htg_hmatrix *h, const double *d_x, double *d_y, int mu, int reps, void *stream, double *ms);

/* Frees the current device's ACA workspace and visited flags, which builds
 * keep allocated between builds (they only grow); the next build allocates
 * them again. Waits for a build running on the device to finish. No
 * counterpart in the reference (which allocates per block). */
int htg_release_workspace(void);

/* Device bandwidth probe (no counterpart in the reference): the best rate over
 * a few grid sizes of a streaming read (non-temporal 16-byte loads, four in flight per lane)
 * and of a streaming copy of `bytes` (>= 1 MiB), each averaged over `reps`
 * launches after a warm-up; GB/s = bytes read (copy: read + written) / time.
 * best_grid (may be NULL): workgroups of the fastest read. */
int htg_stream_probe(uint64_t bytes, int reps, double *read_gbps, double *copy_gbps, int *best_grid);
/* One copy of `bytes` (a multiple of 16) between device buffers with the
 * probe's copy kernel, synchronous: a transfer of known size for scaling HBM
 * traffic counters. */
int htg_stream_copy(const void *d_src, void *d_dst, uint64_t bytes);

/* Batched stand-alone leaf compression, replaces calling
 * VirtualLowRankGenerator::copy_low_rank_approximation once per block
 * (include/htool/lrmat/virtual_lrmat_generator.hpp:15, partialACA.hpp:39-171,
 * sympartialACA.hpp:38-209). Block b covers target points rows[row_off[b] ..
 * row_off[b]+M[b]) and source points cols[col_off[b] .. +N[b]) (original
 * indices into xt / xs); ctr_t/ctr_s (3 doubles per block) are the target and
 * source cluster centres and off_t/off_s their cluster offsets (used for the
 * first pivot and the sympartialACA orientation). Outputs: ranks[b] (-1 =
 * incompressible), U at u_off[b] (M x k column-major) and V at v_off[b]
 * (k x N column-major) into the caller's host arrays sized with `cap` terms per
 * block (u_off[b+1]-u_off[b] >= M[b]*cap). */
int htg_aca_batch(const htg_generator *gen, int compressor, double epsilon, int reqrank, int nt, const double *xt,
                  int ns, const double *xs, int nblocks, const int *M, const int *N, const int *rows,
                  const int *row_off, const int *cols, const int *col_off, const double *ctr_t, const double *ctr_s,
                  const int *off_t, const int *off_s, int cap, int *ranks, double *U, const int64_t *u_off, double *V,
                  const int64_t *v_off);

/* htg_aca_batch with no device allocation. No counterpart in the reference:
 * for callers that cannot call hipMalloc (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation). Same arguments, kernels and
 * results as htg_aca_batch (ranks and factors bit for bit); every device buffer
 * of the call (gathered points, block records, ACA workspace, visited flags,
 * ranks) lives in one statically allocated scratch array of 64 MiB per GPU,
 * held by the library until the process ends. A call that needs more (about
 * 8 * cap * sum(M + N) + 33 * sum(M + N) bytes) returns HTG_ERR_INVALID before
 * touching the device. Calls share the array: the library runs them one at a
 * time, each leaving its contents behind (the kernels clear what must start
 * at zero). Synchronous, on the default stream. */
int htg_aca_batch_noalloc(const htg_generator *gen, int compressor, double epsilon, int reqrank, int nt,
                          const double *xt, int ns, const double *xs, int nblocks, const int *M, const int *N,
                          const int *rows, const int *row_off, const int *cols, const int *col_off,
                          const double *ctr_t, const double *ctr_s, const int *off_t, const int *off_s, int cap,
                          int *ranks, double *U, const int64_t *u_off, double *V, const int64_t *v_off);

/* ACA records of the last build, one per ACA run of an admissible block (first
 * attempt): the 8 block features (ln m, ln n, ln d, ln diam_t, ln diam_s, eta,
 * ln epsilon, ln gap; d the centre distance, gap = max(d - rad_t - rad_s,
 * 1e-6 (rad_t + rad_s))), the cross terms the run needed (its rank, or
 * floor(mn/(m+n)) when the block was not compressible) and the workspace terms
 * of its first attempt. Call with feat == NULL to get the count in *n. */
int htg_hmatrix_aca_records(const htg_hmatrix *h, int *n, float *feat, int *terms, int *cap_first);

/* Learned ACA rank predictor (no counterpart in the reference): an
 * 8 -> 64 -> 64 -> 1 perceptron with GELU activations in FP32, mapping the
 * normalised features (f - mean) / std of an admissible block to ln(1 + t), t
 * the cross terms its ACA run needs. Parameters: HTG_MLP_NPARAMS floats, W1
 * [64][8], b1 [64], W2 [64][64], b2 [64], w3 [64], b3. Trained on the GPU with
 * the library's own forward and backward kernels and Adam. */
#define HTG_MLP_NFEAT 8
#define HTG_MLP_NPARAMS 4801
/* Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases from `seed`. */
int htg_predictor_create(unsigned long long seed, htg_predictor **out);
int htg_predictor_destroy(htg_predictor *p);
int htg_predictor_set_params(htg_predictor *p, const float *params, const float *mean, const float *stdv);
int htg_predictor_get_params(const htg_predictor *p, float *params, float *mean, float *stdv);
/* Loss (1/n) sum 0.5 (out_i - target_i)^2 and its gradient with respect to
 * the parameters at the current weights, over all n samples (host arrays). */
int htg_predictor_loss_grad(const htg_predictor *p, const float *feat, const float *target, int n, double *loss,
                            float *grad);
/* Adam on mini-batches of `batch` samples, shuffled from `seed`; fit_norm = 1
 * first sets mean and std from the data. epoch_loss[e] = mean loss of epoch e. */
int htg_predictor_train(htg_predictor *p, const float *feat, const float *target, int n, int epochs, int batch,
                        float lr, unsigned long long seed, int fit_norm, double *epoch_loss);
/* Network outputs (ln(1 + t) estimates) for n feature rows. */
int htg_predictor_predict(const htg_predictor *p, const float *feat, int n, float *out);

#ifdef __cplusplus
}
#endif
#endif
