// Drives htool_amd::HMatrixGpu<T> (include/htool_amd/htool_hmatrix.hpp) only
// through the reference's htool::VirtualHMatrix<T> interface, beside the
// reference's own HMatrix<T> on the same clusters and generator, and reports
// how far the two agree. Built against the reference's headers
// (tests/cpp/Makefile); the generators are the reference's own test
// generators (include/htool/testing/generator_test.hpp).
//
//   virtual_hmatrix_check case points_t.f64 points_s.f64 nt ns out_prefix
// case 0: GeneratorTestDouble, two point sets, partialACA, general storage
// case 1: GeneratorTestDoubleSymmetric, one point set, 'S' / 'L', sympartialACA
// case 2: GeneratorTestComplex, two point sets, partialACA (complex<double>)
// case 3: the same as 0 through a htool_amd::KernelGenerator (device family)
// Prints key=value lines; writes <out_prefix>_y.f64, the GPU product of the
// three columns x(j, c) = sin(1 + j + 0.37 c) (complex: + i cos(2 j - c)), for
// the oracle comparison.
// Exit codes: 0 ok, 3 library error (message on stderr).
#include <htool_amd/htool_hmatrix.hpp>

#include <htool/clustering/pca.hpp>
#include <htool/testing/generator_test.hpp>
#include <htool/types/hmatrix.hpp>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace htool;

static std::vector<double> read_file(const std::string &path, size_t n) {
    std::vector<double> v(n);
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(double), n, f) != n) {
        std::fprintf(stderr, "cannot read %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

template <typename T>
static double rel(const std::vector<T> &a, const std::vector<T> &b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); i++) {
        num += std::norm(a[i] - b[i]);
        den += std::norm(b[i]);
    }
    return std::sqrt(num / den);
}

template <typename T>
static T value(int j, int c) {
    return T(std::sin(1.0 + j + 0.37 * c));
}
template <>
std::complex<double> value<std::complex<double>>(int j, int c) {
    return {std::sin(1.0 + j + 0.37 * c), std::cos(2.0 * j - c)};
}

template <typename T>
static int run(int which, const std::vector<double> &p1, const std::vector<double> &p2, int nt, int ns,
               const std::string &out, VirtualGenerator<T> &A, VirtualGenerator<T> *device_gen) {
    const bool sym = which == 1;
    auto t = std::make_shared<Cluster<PCAGeometricClustering>>();
    t->set_minclustersize(10);
    t->build(nt, p1.data());
    std::shared_ptr<Cluster<PCAGeometricClustering>> s = t;
    if (!sym) {
        s = std::make_shared<Cluster<PCAGeometricClustering>>();
        s->set_minclustersize(10);
        s->build(ns, p2.data());
    }
    const double eps = 1e-6, eta = 10;
    std::shared_ptr<VirtualLowRankGenerator<T>> comp;
    if (sym)
        comp = std::make_shared<sympartialACA<T>>();
    else
        comp = std::make_shared<partialACA<T>>();
    // the reference
    HMatrix<T> HR(t, s, eps, eta, sym ? 'S' : 'N', sym ? 'L' : 'N');
    HR.set_compression(comp);
    if (sym)
        HR.build(A, p1.data());
    else
        HR.build(A, p1.data(), p2.data());
    // the GPU implementation, used only as a VirtualHMatrix<T>
    htool_amd::HMatrixGpu<T> G(t, s, eps, eta, sym ? 'S' : 'N', sym ? 'L' : 'N');
    VirtualHMatrix<T> &V = G;
    V.set_compression(comp);
    VirtualGenerator<T> &AG = device_gen ? *device_gen : A;
    if (sym)
        V.build(AG, p1.data());
    else
        V.build(AG, p1.data(), p2.data());

    const int mu = 3;
    std::vector<T> x(size_t(ns) * mu), yr(size_t(nt) * mu), yg(size_t(nt) * mu);
    for (int c = 0; c < mu; c++)
        for (int j = 0; j < ns; j++)
            x[j + size_t(ns) * c] = value<T>(j, c);
    HR.mvprod_global_to_global(x.data(), yr.data(), mu);
    V.mvprod_global_to_global(x.data(), yg.data(), mu);
    std::printf("global_to_global=%.3e\n", rel(yg, yr));
    std::vector<T> yl(yg.size());
    V.mvprod_local_to_local(x.data(), yl.data(), mu);
    std::printf("local_to_local=%.3e\n", rel(yl, yr));
    // cluster ordering, interleaved (hmatrix.hpp:979-996): the reference's GMV call
    std::vector<T> xc(x.size()), ycr(yr.size()), ycg(yr.size());
    for (int c = 0; c < mu; c++)
        for (int j = 0; j < ns; j++)
            xc[c + size_t(mu) * j] = x[s->get_perm(j) + size_t(ns) * c];
    HR.mymvprod_local_to_local(xc.data(), ycr.data(), mu);
    V.mymvprod_local_to_local(xc.data(), ycg.data(), mu);
    std::printf("mymvprod_local_to_local=%.3e\n", rel(ycg, ycr));
    if (!sym) {
        std::vector<T> xt_(size_t(nt) * mu), ztr(size_t(ns) * mu), ztg(size_t(ns) * mu);
        for (int c = 0; c < mu; c++)
            for (int i = 0; i < nt; i++)
                xt_[i + size_t(nt) * c] = value<T>(i + 7, c);
        HR.mvprod_transp_global_to_global(xt_.data(), ztr.data(), mu);
        V.mvprod_transp_global_to_global(xt_.data(), ztg.data(), mu);
        std::printf("transp_global_to_global=%.3e\n", rel(ztg, ztr));
    } else {
        // cluster ordering: the reference's trees built without a partition are not
        // "local" and refuse the permuted diagonal (hmatrix.hpp:1850-1852)
        std::vector<T> dr = HR.get_local_diagonal(false), dg = V.get_local_diagonal(false);
        std::printf("diagonal=%.3e\n", rel(dg, dr));
    }
    std::printf("nlrmat=%s %s\n", HR.get_infos("Number_of_lrmat").c_str(), V.get_infos("Number_of_lrmat").c_str());
    std::printf("ndmat=%s %s\n", HR.get_infos("Number_of_dmat").c_str(), V.get_infos("Number_of_dmat").c_str());
    std::printf("false_positive=%s %s\n", HR.get_infos("Number_of_false_positive").c_str(),
                V.get_infos("Number_of_false_positive").c_str());
    std::printf("nb_mat_vec_prod=%s\n", V.get_infos("nb_mat_vec_prod").c_str());
    const std::pair<int, int> mr = HR.get_max_size_blocks(), mg = V.get_max_size_blocks();
    std::printf("max_size_blocks=%d,%d %d,%d\n", mr.first, mr.second, mg.first, mg.second);
    FILE *f = std::fopen((out + "_y.f64").c_str(), "wb");
    if (!f)
        return 2;
    std::fwrite(yg.data(), sizeof(T), yg.size(), f);
    std::fclose(f);
    return 0;
}

int main(int argc, char **argv) {
    MPI_Init(&argc, &argv);
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s case points_t points_s nt ns out_prefix\n", argv[0]);
        return 2;
    }
    const int which = std::atoi(argv[1]), nt = std::atoi(argv[4]), ns = std::atoi(argv[5]);
    std::vector<double> p1 = read_file(argv[2], 3 * size_t(nt));
    std::vector<double> p2 = which == 1 ? p1 : read_file(argv[3], 3 * size_t(ns));
    const int nsrc = which == 1 ? nt : ns;
    int rc = 0;
    try {
        if (which == 0) {
            GeneratorTestDouble A(3, nt, nsrc, p1, p2);
            rc = run<double>(which, p1, p2, nt, nsrc, argv[6], A, nullptr);
        } else if (which == 1) {
            GeneratorTestDoubleSymmetric A(3, nt, p1);
            rc = run<double>(which, p1, p2, nt, nsrc, argv[6], A, nullptr);
        } else if (which == 2) {
            GeneratorTestComplex A(3, nt, nsrc, p1, p2);
            rc = run<std::complex<double>>(which, p1, p2, nt, nsrc, argv[6], A, nullptr);
        } else {
            GeneratorTestDouble A(3, nt, nsrc, p1, p2);
            htool_amd::KernelGenerator<double> K(htg_generator{HTG_KERNEL_LAPLACE, 0, 0.0, 0.0, 0.0}, nt, p1.data(),
                                                 nsrc, p2.data());
            // the device family and the reference's generator give the same entries
            std::vector<int> r{0, 1, 2}, c{3, 4};
            std::vector<double> a(6), k(6);
            A.copy_submatrix(3, 2, r.data(), c.data(), a.data());
            K.copy_submatrix(3, 2, r.data(), c.data(), k.data());
            std::printf("copy_submatrix=%.3e\n", rel(k, a));
            rc = run<double>(0, p1, p2, nt, nsrc, argv[6], A, &K);
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        rc = 3;
    }
    MPI_Finalize();
    return rc;
}
