// htool::VirtualHMatrix<T> implemented by libhtool_amd: the class a Htool user
// (or Htool itself) links to move an existing HMatrix<T> workload to the GPU
// without changing its callers (user code, HPDDM's GMV through
// mymvprod_local_to_local, tests written against VirtualHMatrix<T>).
//
// It needs the reference's headers (Htool 0.8.1, include/htool/...) and MPI on
// the include path; T is double or std::complex<double>. Interface followed:
// include/htool/types/virtual_hmatrix.hpp:11-78 (every pure virtual),
// include/htool/types/hmatrix.hpp:112-122 (constructors), 213-226 (setters).
//
// * build(A, xt, xs): A is any VirtualGenerator<T>; its copy_submatrix feeds
//   the GPU ACA row by row and column by column and fills the dense leaves
//   (htg_hmatrix_build_ex with a host generator). A htool_amd::KernelGenerator
//   (a built-in family) is evaluated on the device instead. The caller's
//   cluster trees t and s are the build's (their permutation, nodes, centres).
// * Compressors: partialACA<T> and sympartialACA<T> (the reference's default);
//   others (SVD, fullACA) are not on the GPU path and are refused.
// * One process: the communicator must have one rank (local = global).
// * Errors are std::invalid_argument / std::logic_error / std::runtime_error
//   with the reference's "[Htool error]" prefix.
#ifndef HTOOL_AMD_HTOOL_HMATRIX_HPP
#define HTOOL_AMD_HTOOL_HMATRIX_HPP

#include "../htool_amd.h"

#include <htool/clustering/virtual_cluster.hpp>
#include <htool/lrmat/partialACA.hpp>
#include <htool/lrmat/sympartialACA.hpp>
#include <htool/types/virtual_hmatrix.hpp>

#include <chrono>
#include <complex>
#include <iostream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace htool_amd {

inline void check_status(int status) {
    if (status == HTG_OK)
        return;
    std::string msg = htg_last_error();
    if (msg.rfind("[Htool error]", 0) != 0)
        msg = "[Htool error] " + msg;
    if (status == HTG_ERR_INVALID)
        throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

template <typename T>
struct is_complex_t : std::false_type {};
template <typename T>
struct is_complex_t<std::complex<T>> : std::true_type {};

// A built-in kernel family as a VirtualGenerator<T> (virtual_generator.hpp:11-29):
// copy_submatrix computes the entries on the device (htg_generator_block); a
// build given this generator evaluates the family inside its own kernels.
template <typename T>
class KernelGenerator : public htool::VirtualGenerator<T> {
  public:
    KernelGenerator(const htg_generator &g, int nr, const double *xt, int nc, const double *xs)
        : htool::VirtualGenerator<T>(nr, nc), gen_(g), xt_(xt), xs_(xs) {}
    void copy_submatrix(int M, int N, const int *const rows, const int *const cols, T *ptr) const override {
        check_status(htg_generator_block(&gen_, this->nr, xt_, this->nc, xs_, M, N, rows, cols,
                                         reinterpret_cast<double *>(ptr)));
    }
    const htg_generator &desc() const { return gen_; }

  private:
    htg_generator gen_;
    const double *xt_, *xs_;
};

template <typename T>
class HMatrixGpu final : public htool::VirtualHMatrix<T> {
    static_assert(std::is_same<T, double>::value || std::is_same<T, std::complex<double>>::value,
                  "HMatrixGpu<T>: T is double or std::complex<double>");

  public:
    // HMatrix(t, s, epsilon, eta, Symmetry, UPLO, reqrank, comm), hmatrix.hpp:115-122.
    HMatrixGpu(const std::shared_ptr<htool::VirtualCluster> &t, const std::shared_ptr<htool::VirtualCluster> &s,
               double epsilon = 1e-6, double eta = 10, char Symmetry = 'N', char UPLO = 'N', const int &reqrank = -1,
               MPI_Comm comm = MPI_COMM_WORLD)
        : t_(t), s_(s), comm_(comm) {
        if (!((Symmetry == 'N' || Symmetry == 'S' || Symmetry == 'H') && (UPLO == 'N' || UPLO == 'L' || UPLO == 'U') &&
              ((Symmetry == 'N' && UPLO == 'N') || (Symmetry != 'N' && UPLO != 'N'))))
            throw std::invalid_argument("[Htool error] Invalid arguments to create HMatrix");
        int size = 1;
        MPI_Comm_size(comm, &size);
        if (size != 1)
            throw std::invalid_argument("[Htool error] HMatrixGpu runs one process per GPU (communicator of one rank)");
        prm_                = htg_params{};
        prm_.epsilon        = epsilon;
        prm_.eta            = eta;
        prm_.maxblocksize   = 1e6;
        prm_.minclustersize = t->get_minclustersize();
        prm_.compressor     = HTG_SYM_PARTIAL_ACA; // the reference's default (hmatrix.hpp:297-299)
        prm_.reqrank        = reqrank;
        prm_.predictor_margin = 0.25;
        prm_.symmetry       = Symmetry;
        prm_.uplo           = UPLO;
    }
    HMatrixGpu(const HMatrixGpu &) = delete;
    HMatrixGpu &operator=(const HMatrixGpu &) = delete;
    ~HMatrixGpu() override { reset(); }

    // -- getters ---------------------------------------------------------------
    int nb_rows() const override { return info().nr; }
    int nb_cols() const override { return info().nc; }
    MPI_Comm get_comm() const override { return comm_; }
    int get_rankworld() const override { return 0; }
    int get_sizeworld() const override { return 1; }
    int get_local_size() const override { return info().nr; }
    int get_local_offset() const override { return 0; }
    char get_symmetry_type() const override { return (char)prm_.symmetry; }
    char get_storage_type() const override { return (char)prm_.uplo; }
    std::vector<T> get_local_diagonal(bool permutation = true) const override {
        std::vector<T> d(nb_rows());
        copy_local_diagonal(d.data(), permutation);
        return d;
    }
    void copy_local_diagonal(T *ptr, bool permutation = true) const override {
        if (t_ != s_)
            throw std::logic_error("[Htool error] Matrix is not square a priori, get_local_diagonal cannot be used");
        check_status(htg_hmatrix_get_diagonal(handle(), reinterpret_cast<double *>(ptr), permutation ? 1 : 0));
    }
    // one process: the local diagonal block is the whole matrix (hmatrix.hpp:1976-1981)
    htool::Matrix<T> get_local_diagonal_block(bool permutation = true) const override {
        htool::Matrix<T> m(nb_rows(), nb_cols());
        copy_local_diagonal_block(m.data(), permutation);
        return m;
    }
    void copy_local_diagonal_block(T *ptr, bool permutation = true) const override {
        if (t_ != s_)
            throw std::logic_error("[Htool error] Matrix is not square a priori, get_local_diagonal_block cannot be used");
        check_status(htg_hmatrix_get_dense(handle(), reinterpret_cast<double *>(ptr), permutation ? 1 : 0));
    }
    std::pair<int, int> get_max_size_blocks() const override {
        int nl = 0;
        int64_t ds = 0;
        check_status(htg_hmatrix_nleaves(handle(), &nl, &ds));
        std::vector<int> l5(5 * (size_t)nl);
        check_status(htg_hmatrix_get_leaves(handle(), l5.data(), nullptr));
        int mi = 0, mj = 0;
        for (int l = 0; l < nl; l++) {
            mi = std::max(mi, l5[5 * l + 1]);
            mj = std::max(mj, l5[5 * l + 3]);
        }
        return {mi, mj};
    }
    const htool::VirtualCluster *get_target_cluster() const override { return t_.get(); }
    const htool::VirtualCluster *get_source_cluster() const override { return s_.get(); }

    // -- parameters (hmatrix.hpp:207-220) ----------------------------------------
    double get_epsilon() const override { return prm_.epsilon; }
    double get_eta() const override { return prm_.eta; }
    int get_dimension() const override { return 1; }
    int get_minsourcedepth() const override { return prm_.minsourcedepth; }
    int get_mintargetdepth() const override { return prm_.mintargetdepth; }
    int get_maxblocksize() const override { return (int)prm_.maxblocksize; }
    void set_epsilon(double e) override { prm_.epsilon = e; }
    void set_eta(double e) override { prm_.eta = e; }
    void set_maxblocksize(unsigned int m) override {
        if (m == 0)
            throw std::invalid_argument("[Htool error] MaxBlockSize parameter cannot be zero");
        prm_.maxblocksize = m;
    }
    void set_minsourcedepth(unsigned int d) override { prm_.minsourcedepth = (int)d; }
    void set_mintargetdepth(unsigned int d) override { prm_.mintargetdepth = (int)d; }
    void set_use_permutation(bool choice) override { prm_.no_permutation = choice ? 0 : 1; }
    void set_compression(std::shared_ptr<htool::VirtualLowRankGenerator<T>> c) override {
        if (std::dynamic_pointer_cast<htool::sympartialACA<T>>(c))
            prm_.compressor = HTG_SYM_PARTIAL_ACA;
        else if (std::dynamic_pointer_cast<htool::partialACA<T>>(c))
            prm_.compressor = HTG_PARTIAL_ACA;
        else
            throw std::invalid_argument("[Htool error] HMatrixGpu compresses with partialACA or sympartialACA");
    }
    // set_delay_dense_computation (hmatrix.hpp:219); the blocks come with build_dense_blocks
    void set_delay_dense_computation(bool d) { prm_.delay_dense = d ? 1 : 0; }
    // Options of the GPU path (no counterpart in the reference).
    void set_storage(htg_storage s) {
        check_status(htg_hmatrix_set_storage(handle(), s));
        storage_ = s;
        if (ht_)
            check_status(htg_hmatrix_set_storage(ht_, s));
    }
    void set_recompression(double tol) { prm_.recompress_tol = tol; }
    void set_lowprec(double tol, bool all_terms = false) {
        prm_.lowprec_tol = tol;
        prm_.lowprec_all = all_terms ? 1 : 0;
    }
    htg_hmatrix *native() const { return handle(); }

    // -- build (hmatrix.hpp:285-402) ---------------------------------------------
    void build(htool::VirtualGenerator<T> &mat, const double *const xt, const double *const xs) override {
        build_impl(mat, xt, xs);
    }
    void build(htool::VirtualGenerator<T> &mat, const double *const xt) override { build_impl(mat, xt, nullptr); }

    // -- products ----------------------------------------------------------------
    void mvprod_global_to_global(const T *const in, T *const out, const int &mu = 1) const override {
        check_status(htg_hmatrix_mvprod(handle(), cd(in), d(out), mu));
    }
    // one process: the local vectors are the global ones (hmatrix.hpp:1184-1274)
    void mvprod_local_to_local(const T *const in, T *const out, const int &mu = 1, T * = nullptr) const override {
        mvprod_global_to_global(in, out, mu);
    }
    void mvprod_transp_global_to_global(const T *const in, T *const out, const int &mu = 1) const override {
        check_status(htg_hmatrix_mvprod(transposed(), cd(in), d(out), mu));
    }
    void mvprod_transp_local_to_local(const T *const in, T *const out, const int &mu = 1, T * = nullptr) const override {
        mvprod_transp_global_to_global(in, out, mu);
    }
    // cluster ordering, interleaved right-hand sides (hmatrix.hpp:979-996, 843-911)
    void mymvprod_local_to_local(const T *const in, T *const out, const int &mu = 1, T * = nullptr) const override {
        check_status(htg_hmatrix_mymvprod_local_to_local(handle(), cd(in), d(out), mu));
    }
    void mymvprod_global_to_local(const T *const in, T *const out, const int &mu = 1) const override {
        mymvprod_local_to_local(in, out, mu);
    }
    void mymvprod_transp_local_to_local(const T *const in, T *const out, const int &mu = 1,
                                        T * = nullptr) const override {
        check_status(htg_hmatrix_mymvprod_local_to_local(transposed(), cd(in), d(out), mu));
    }
    void mymvprod_transp_local_to_global(const T *const in, T *const out, const int &mu = 1) const override {
        mymvprod_transp_local_to_local(in, out, mu);
    }
    void mvprod_subrhs(const T *const in, T *const out, const int &mu, const int &offset, const int &size,
                       const int &margin) const override {
        check_status(htg_hmatrix_mvprod_subrhs(handle(), cd(in), d(out), mu, offset, size, margin));
    }
    // Device-resident products (HIP device pointers, a hipStream_t or nullptr).
    void mvprod_device(const T *d_in, T *d_out, int mu, void *stream = nullptr) const {
        check_status(htg_hmatrix_mvprod_device(handle(), cd(d_in), d(d_out), mu, stream));
    }
    void mymvprod_local_to_local_device(const T *d_in, T *d_out, int mu, T *d_work = nullptr,
                                        void *stream = nullptr) const {
        check_status(htg_hmatrix_mymvprod_local_to_local_device(handle(), cd(d_in), d(d_out), mu, d(d_work), stream));
    }

    // -- infos (hmatrix.hpp:223-226, 794-840, 1180-1181) -----------------------------
    const std::map<std::string, std::string> &get_infos() const override {
        const htg_info i = info();
        htg_info it{}; // products of the transpose count as this matrix's (mvprod_transp_*)
        if (ht_)
            check_status(htg_hmatrix_get_info(ht_, &it));
        infos_["Number_of_lrmat"]          = std::to_string(i.n_lowrank);
        infos_["Number_of_dmat"]           = std::to_string(i.n_dense);
        infos_["Number_of_false_positive"] = std::to_string(i.n_false_positive);
        infos_["Rank_min"]                 = std::to_string(i.rank_min);
        infos_["Rank_max"]                 = std::to_string(i.rank_max);
        infos_["Rank_mean"]                = std::to_string(i.rank_mean);
        infos_["Compression_ratio"]        = std::to_string(i.compression_ratio);
        infos_["Target_size"]              = std::to_string(i.nr);
        infos_["Source_size"]              = std::to_string(i.nc);
        infos_["Blocks_max"]               = std::to_string(i.build_seconds);
        infos_["nb_mat_vec_prod"]          = std::to_string(i.nb_mat_vec_prod + it.nb_mat_vec_prod);
        infos_["total_time_mat_vec_prod"]  = std::to_string(i.total_time_mat_vec_prod + it.total_time_mat_vec_prod);
        return infos_;
    }
    std::string get_infos(const std::string &key) const override {
        const auto &m = get_infos();
        const auto it = m.find(key);
        return it == m.end() ? std::string() : it->second;
    }
    void print_infos() const override {
        for (const auto &kv : get_infos())
            std::cout << kv.first << "\t" << kv.second << std::endl;
        std::cout << std::endl;
    }

    // -- dense conversions ---------------------------------------------------------
    htool::Matrix<T> get_local_dense_perm() const override {
        htool::Matrix<T> m(nb_rows(), nb_cols());
        copy_local_dense_perm(m.data());
        return m;
    }
    void copy_local_dense_perm(T *ptr) const override {
        check_status(htg_hmatrix_get_dense(handle(), reinterpret_cast<double *>(ptr), 1));
    }
    // get_local_dense (hmatrix.hpp:1728-1757): the cluster ordering
    htool::Matrix<T> get_local_dense() const {
        htool::Matrix<T> m(nb_rows(), nb_cols());
        check_status(htg_hmatrix_get_dense(handle(), reinterpret_cast<double *>(m.data()), 0));
        return m;
    }

  private:
    // std::complex<double> is laid out as double[2] ([complex.numbers]), the C ABI's (re, im) pairs
    static const double *cd(const T *p) { return reinterpret_cast<const double *>(p); }
    static double *d(T *p) { return reinterpret_cast<double *>(p); }

    htg_info info() const {
        htg_info i{};
        check_status(htg_hmatrix_get_info(handle(), &i));
        return i;
    }
    htg_hmatrix *handle() const {
        if (!h_)
            throw std::logic_error("[Htool error] HMatrix is not built");
        return h_;
    }
    htg_hmatrix *transposed() const {
        if (!ht_) {
            check_status(htg_hmatrix_transpose(handle(), &ht_));
            if (storage_ != HTG_STORAGE_FP64)
                check_status(htg_hmatrix_set_storage(ht_, storage_));
        }
        return ht_;
    }
    void reset() {
        htg_hmatrix_destroy(ht_);
        htg_hmatrix_destroy(h_);
        h_ = ht_ = nullptr;
        storage_ = HTG_STORAGE_FP64;
    }

    // A cluster tree of the reference as the C ABI's htg_cluster_tree: nodes in
    // breadth order, the sons of a node consecutive.
    struct Tree {
        std::vector<int> perm, ni;
        std::vector<double> nd;
        htg_cluster_tree c{};
    };
    static void flatten(const htool::VirtualCluster &root, Tree &out) {
        std::vector<const htool::VirtualCluster *> nodes{&root};
        for (size_t k = 0; k < nodes.size(); k++) {
            const htool::VirtualCluster &c = *nodes[k];
            const int first = c.get_nb_sons() > 0 ? (int)nodes.size() : -1;
            for (int p = 0; p < c.get_nb_sons(); p++)
                nodes.push_back(&c.get_son(p));
            const int v[6] = {c.get_offset(), c.get_size(), c.get_depth(), c.get_rank(), c.get_nb_sons(), first};
            out.ni.insert(out.ni.end(), v, v + 6);
            const std::vector<double> &ctr = c.get_ctr();
            for (int q = 0; q < 3; q++)
                out.nd.push_back(q < (int)ctr.size() ? ctr[q] : 0.0);
            out.nd.push_back(c.get_rad());
        }
        out.perm = root.get_perm();
        out.c    = htg_cluster_tree{(int)out.perm.size(), out.perm.data(), (int)nodes.size(), out.ni.data(),
                                 out.nd.data()};
    }

    static void trampoline(int M, int N, const int *rows, const int *cols, double *ptr, void *ctx) {
        auto *self = static_cast<Callback *>(ctx);
        try {
            self->gen->copy_submatrix(M, N, rows, cols, reinterpret_cast<T *>(ptr));
        } catch (...) {
            if (!self->error)
                self->error = std::current_exception();
        }
    }
    struct Callback {
        htool::VirtualGenerator<T> *gen;
        std::exception_ptr error;
    };

    void build_impl(htool::VirtualGenerator<T> &mat, const double *xt, const double *xs) {
        reset();
        Tree tt, ts;
        flatten(*t_->get_root(), tt);
        const bool same = xs == nullptr || t_ == s_;
        if (!same)
            flatten(*s_->get_root(), ts);
        const int nt = (int)tt.perm.size(), ns = same ? nt : (int)ts.perm.size();
        if (mat.nb_rows() != nt || mat.nb_cols() != ns)
            throw std::invalid_argument("[Htool error] Invalid size in arguments for building HMatrix");
        if (auto *kg = dynamic_cast<KernelGenerator<T> *>(&mat)) {
            htg_generator g = kg->desc();
            g.same_points   = same ? 1 : 0;
            check_status(htg_hmatrix_build_ex(&prm_, &g, nullptr, nt, xt, ns, same ? nullptr : xs, &tt.c,
                                              same ? nullptr : &ts.c, &h_));
            return;
        }
        Callback cb{&mat, nullptr};
        htg_host_generator hg{&HMatrixGpu::trampoline, &cb, is_complex_t<T>::value ? 1 : 0};
        const int st = htg_hmatrix_build_ex(&prm_, nullptr, &hg, nt, xt, ns, same ? nullptr : xs, &tt.c,
                                            same ? nullptr : &ts.c, &h_);
        if (cb.error) {
            reset();
            std::rethrow_exception(cb.error);
        }
        check_status(st);
    }

    std::shared_ptr<htool::VirtualCluster> t_, s_;
    MPI_Comm comm_;
    htg_params prm_{};
    htg_hmatrix *h_          = nullptr;
    mutable htg_hmatrix *ht_ = nullptr; // the transpose, made on first use
    htg_storage storage_     = HTG_STORAGE_FP64;
    mutable std::map<std::string, std::string> infos_;
};

} // namespace htool_amd
#endif
