// C++ host interface over the C ABI of libhtool_amd.so, mirroring the product
// surface of the reference's HMatrix<double> (include/htool/types/hmatrix.hpp,
// Htool 0.8.1): construction with (epsilon, eta, reqrank), the setters the hot
// path reads, build from points, mvprod_global_to_global, operator*, infos.
// Errors surface as exceptions with the reference's "[Htool error]" prefix.
// Header-only; link with -lhtool_amd.
#ifndef HTOOL_AMD_GPU_HMATRIX_HPP
#define HTOOL_AMD_GPU_HMATRIX_HPP

#include "../htool_amd.h"

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace htool_amd {

inline void check(int status) {
    if (status == HTG_OK)
        return;
    std::string msg = htg_last_error();
    if (msg.rfind("[Htool error]", 0) != 0)
        msg = "[Htool error] " + msg;
    if (status == HTG_ERR_INVALID)
        throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// A built-in kernel family evaluated on the GPU; stands in for a user
// VirtualGenerator<double> (include/htool/types/virtual_generator.hpp:11-29).
struct Generator {
    htg_kernel kernel = HTG_KERNEL_LAPLACE;
    double delta      = 0.0;
    double diag       = 0.0;
    double wavenumber = 0.0; // HTG_KERNEL_HELMHOLTZ
};

class GpuHMatrix {
  public:
    // HMatrix<T>(t, s, epsilon, eta, symmetry, UPLO, reqrank), hmatrix.hpp:115-122;
    // symmetry 'S' with UPLO 'L' / 'U' keeps one triangle of blocks (the
    // arguments are checked by the build, as the reference's constructor does).
    explicit GpuHMatrix(double epsilon = 1e-6, double eta = 10, char symmetry = 'N', char UPLO = 'N',
                        int reqrank = -1) {
        prm_.epsilon        = epsilon;
        prm_.eta            = eta;
        prm_.maxblocksize   = 1e6;
        prm_.minclustersize = 10;
        prm_.compressor     = HTG_SYM_PARTIAL_ACA; // the reference's default (hmatrix.hpp:297-299)
        prm_.reqrank        = reqrank;
        prm_.lowprec_all    = 0;
        prm_.recompress_tol = 0.0;
        prm_.lowprec_tol    = 0.0;
        prm_.predictor        = nullptr;
        prm_.predictor_margin = 0.25;
        prm_.aca_wide_min     = 0;
        prm_.symmetry         = symmetry;
        prm_.uplo             = UPLO;
        prm_.delay_dense      = 0;
        prm_.ref_int_overflow = 0;
    }
    GpuHMatrix(const GpuHMatrix &) = delete;
    GpuHMatrix &operator=(const GpuHMatrix &) = delete;
    ~GpuHMatrix() {
        htg_hmatrix_destroy(ht_);
        htg_hmatrix_destroy(h_);
    }

    // setters of hmatrix.hpp:213-220 and Cluster::set_minclustersize (cluster.hpp:334)
    void set_epsilon(double e) { prm_.epsilon = e; }
    void set_eta(double e) { prm_.eta = e; }
    void set_maxblocksize(unsigned int m) {
        if (m == 0)
            throw std::invalid_argument("[Htool error] MaxBlockSize parameter cannot be zero");
        prm_.maxblocksize = m;
    }
    void set_minclustersize(unsigned int m) {
        if (m == 0)
            throw std::invalid_argument("[Htool error] MinClusterSize parameter cannot be zero");
        prm_.minclustersize = (int)m;
    }
    void set_compression(htg_compressor c) { prm_.compressor = c; }
    void set_delay_dense_computation(bool d) { prm_.delay_dense = d ? 1 : 0; }
    // The ACA's q (M + N) > M N exit in 32-bit int as the reference computes it
    // (partialACA.hpp:74): blocks with M N > 2^31 - 1 are then split as there.
    void set_ref_int_overflow(bool on) { prm_.ref_int_overflow = on ? 1 : 0; }
    // build_dense_blocks (hmatrix.hpp:404-418): every kept dense leaf, in
    // get_leaves order, m x n column-major in cluster ordering
    // (the transpose, built from the blocks present when it was made, is dropped)
    void build_dense_blocks(const double *data) {
        check(htg_hmatrix_set_dense_blocks(handle(), data));
        htg_hmatrix_destroy(ht_);
        ht_ = nullptr;
    }
    // Build-time options of the mixed-precision product (no reference counterpart).
    void set_recompression(double tol) { prm_.recompress_tol = tol; }
    void set_lowprec(double tol, bool all_terms = false) {
        prm_.lowprec_tol = tol;
        prm_.lowprec_all = all_terms ? 1 : 0;
    }
    void set_storage(htg_storage s) {
        check(htg_hmatrix_set_storage(handle(), s));
        storage_ = s;
        if (ht_)
            check(htg_hmatrix_set_storage(ht_, s));
    }
    // Presizes ACA workspaces from a trained rank predictor (not owned; nullptr: none).
    void set_predictor(const htg_predictor *p, double margin = 0.25) {
        prm_.predictor        = p;
        prm_.predictor_margin = margin;
    }

    // HMatrix::build(A, xt, xs), hmatrix.hpp:285-341 (points interleaved xyz).
    void build(const Generator &g, int nt, const double *xt, int ns, const double *xs) {
        htg_generator gen{g.kernel, xs == nullptr ? 1 : 0, g.delta, g.diag, g.wavenumber};
        reset();
        check(htg_hmatrix_build(&prm_, &gen, nt, xt, ns, xs, &h_));
    }
    // Square build on one point set (same points as rows and columns).
    void build(const Generator &g, int n, const double *x) { build(g, n, x, n, nullptr); }

    // Leaves produced elsewhere (e.g. an existing HMatrix's SubMatrix / LowRankMatrix data).
    void from_leaves(int nr, int nc, const int *perm_t, const int *perm_s, int nleaves, const int *leaves5,
                     const double *data) {
        reset();
        check(htg_hmatrix_from_leaves(nr, nc, perm_t, perm_s, nleaves, leaves5, data, &h_));
    }

    // mvprod_global_to_global(in, out, mu), hmatrix.hpp:1073-1182: column-major
    // nc x mu in, nr x mu out, original ordering.
    void mvprod_global_to_global(const double *in, double *out, const int &mu = 1) const {
        check(htg_hmatrix_mvprod(handle(), in, out, mu));
    }
    // Device-resident variant (HIP device pointers, a hipStream_t or nullptr).
    void mvprod_device(const double *d_in, double *d_out, int mu, void *stream = nullptr) const {
        check(htg_hmatrix_mvprod_device(handle(), d_in, d_out, mu, stream));
    }
    // mvprod_transp_global_to_global(in, out, mu), virtual_hmatrix.hpp:58: the
    // transposed H-matrix is built on first use (same storage mode).
    void mvprod_transp_global_to_global(const double *in, double *out, const int &mu = 1) const {
        if (!ht_) {
            check(htg_hmatrix_transpose(handle(), &ht_));
            if (storage_ != HTG_STORAGE_FP64)
                check(htg_hmatrix_set_storage(ht_, storage_));
        }
        check(htg_hmatrix_mvprod(ht_, in, out, mu));
    }
    // operator*, hmatrix.hpp:1596-1602.
    std::vector<double> operator*(const std::vector<double> &x) const {
        std::vector<double> y(nb_rows());
        mvprod_global_to_global(x.data(), y.data(), 1);
        return y;
    }

    int nb_rows() const { return info().nr; }
    int nb_cols() const { return info().nc; }
    htg_info info() const {
        htg_info i{};
        check(htg_hmatrix_get_info(handle(), &i));
        return i;
    }
    // The keys of the reference's infos map this build fills (hmatrix.hpp:794-840).
    std::map<std::string, std::string> get_infos() const {
        const htg_info i = info();
        return {{"Number_of_lrmat", std::to_string(i.n_lowrank)},
                {"Number_of_dmat", std::to_string(i.n_dense)},
                {"Number_of_false_positive", std::to_string(i.n_false_positive)},
                {"Rank_min", std::to_string(i.rank_min)},
                {"Rank_max", std::to_string(i.rank_max)},
                {"Rank_mean", std::to_string(i.rank_mean)},
                {"Compression_ratio", std::to_string(i.compression_ratio)},
                {"Target_size", std::to_string(i.nr)},
                {"Source_size", std::to_string(i.nc)},
                {"Blocks_max", std::to_string(i.build_seconds)}};
    }
    std::vector<int> get_permt() const {
        std::vector<int> p(nb_rows());
        check(htg_hmatrix_get_perms(handle(), p.data(), nullptr));
        return p;
    }

  private:
    htg_hmatrix *handle() const {
        if (!h_)
            throw std::logic_error("[Htool error] HMatrix is not built");
        return h_;
    }
    void reset() {
        htg_hmatrix_destroy(ht_);
        htg_hmatrix_destroy(h_);
        h_       = nullptr;
        ht_      = nullptr;
        storage_ = HTG_STORAGE_FP64;
    }
    htg_params prm_{};
    htg_hmatrix *h_          = nullptr;
    mutable htg_hmatrix *ht_ = nullptr; // transpose, built on first use
    htg_storage storage_     = HTG_STORAGE_FP64;
};

} // namespace htool_amd
#endif
