// Driver that runs the reference Htool library (header-only, compiled from
// the reference's own headers by oracle/Makefile) on one input case and dumps
// what the tests and the CPU baseline need: permutations, the leaf list, the
// leaf data (dense blocks and U, V factors), y = H x and timings.
//
// It is test infrastructure: only tests/ and bench.py's CPU-baseline leg run it.
//
// Reference API used (file:line, paths relative to the reference root):
//   Cluster<PCAGeometricClustering>::build   include/htool/clustering/cluster.hpp:57-92
//   HMatrix<T>(t, s, eps, eta)               include/htool/types/hmatrix.hpp:115-122
//   HMatrix::set_compression                 include/htool/types/hmatrix.hpp:220
//   HMatrix::build(A, xt, xs)                include/htool/types/hmatrix.hpp:285-341
//   HMatrix::mvprod_global_to_global         include/htool/types/hmatrix.hpp:1073-1182
//   LowRankMatrix::build (stand-alone leaf)  include/htool/lrmat/lrmat.hpp:36-55
//
// Input file layout (little endian), written by tests/refcase.py:
//   int32  magic 0x48545246 ('HTRF'), version 1
//   int32  mode (0 = H-matrix build + product, 1 = one leaf compression of the whole t x s block)
//   int32  nt, ns, same_points, kernel, minclustersize, compressor (0 partialACA, 1 sympartialACA)
//   int32  mu, nreps, dump_data, reqrank, uplo (0: general; 'L' / 'U': HMatrix(..., 'S', uplo)
//          built with the single-point-set build(A, xt), hmatrix.hpp:343-400; 'H' << 8 added:
//          HMatrix(..., 'H', uplo))
//   (version 2 adds) int32 mintargetdepth, minsourcedepth (HMatrix::set_mintargetdepth /
//          set_minsourcedepth, hmatrix.hpp:215-216)
//   double eta, eps, delta, diag, maxblocksize
//   double xt[3 nt]; double xs[3 ns] (absent when same_points); double x[ns mu] (column-major)
// Complex kernel families (kernel 4: e^{ikr}/(4 pi r) with k = `delta`, `diag`
// on the diagonal; kernel 5: (1 + i)/(4 pi r); kernel 7: the reference's own
// GeneratorTestComplexHermitian) run HMatrix<std::complex<double>>: x, y and the
// leaf data are then complex<double> (re, im interleaved), mode 0 only. Kernel 6
// is the reference's GeneratorTestZero (every entry 0).
#include <htool/clustering/pca.hpp>
#include <htool/lrmat/partialACA.hpp>
#include <htool/lrmat/sympartialACA.hpp>
#include <htool/testing/generator_test.hpp>
#include <htool/types/hmatrix.hpp>

#include <chrono>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <omp.h>
#include <string>
#include <vector>

using namespace htool;

namespace {

// Kernel families shared with the HIP generator and the oracle. The arithmetic
// order is fixed (squared distance summed x, y, z from zero, then sqrt) so that
// all three produce identical doubles.
struct KernelGen final : public VirtualGenerator<double> {
    const double *xt, *xs;
    int kernel;
    bool same_points;
    double delta, diag;
    KernelGen(int nr, int nc, const double *xt0, const double *xs0, int kernel0, bool same0, double delta0, double diag0)
        : VirtualGenerator<double>(nr, nc), xt(xt0), xs(xs0), kernel(kernel0), same_points(same0), delta(delta0), diag(diag0) {}

    double coef(int i, int j) const {
        const double *a = xt + 3 * i, *b = xs + 3 * j;
        double r2 = 0;
        for (int d = 0; d < 3; d++) {
            double u = a[d] - b[d];
            r2       = r2 + u * u;
        }
        switch (kernel) {
        case 0:
            if (same_points && i == j)
                return diag;
            return 1. / (4 * M_PI * std::sqrt(r2));
        case 1:
            return 1. / (delta + std::sqrt(r2));
        case 2:
            return 1. / (delta + 4 * M_PI * std::sqrt(r2));
        case 6:
            return 0.0;
        default:
            return 1. / (4 * M_PI * std::sqrt(delta + r2));
        }
    }
    void copy_submatrix(int M, int N, const int *const rows, const int *const cols, double *ptr) const override {
        for (int i = 0; i < M; i++)
            for (int j = 0; j < N; j++)
                ptr[i + M * j] = coef(rows[i], cols[j]);
    }
};

// Complex families, with the distance summed as above. The Helmholtz kernel
// is the config's e^{ikr}/(4 pi r) (no generator of the reference computes it);
// kernel 5 is GeneratorTestComplex's (1 + i)/(4 pi r)
// (include/htool/testing/generator_test.hpp:110-120).
struct KernelGenC final : public VirtualGenerator<std::complex<double>> {
    const double *xt, *xs;
    int kernel;
    bool same_points;
    double wavenumber, diag;
    KernelGenC(int nr, int nc, const double *xt0, const double *xs0, int kernel0, bool same0, double k0, double diag0)
        : VirtualGenerator<std::complex<double>>(nr, nc), xt(xt0), xs(xs0), kernel(kernel0), same_points(same0),
          wavenumber(k0), diag(diag0) {}

    std::complex<double> coef(int i, int j) const {
        const double *a = xt + 3 * i, *b = xs + 3 * j;
        double r2 = 0;
        for (int d = 0; d < 3; d++) {
            double u = a[d] - b[d];
            r2       = r2 + u * u;
        }
        const double r = std::sqrt(r2), den = 4 * M_PI * r;
        if (kernel == 4) {
            if (same_points && i == j)
                return {diag, 0.0};
            return {std::cos(wavenumber * r) / den, std::sin(wavenumber * r) / den};
        }
        return {1. / den, 1. / den};
    }
    void copy_submatrix(int M, int N, const int *const rows, const int *const cols,
                        std::complex<double> *ptr) const override {
        for (int i = 0; i < M; i++)
            for (int j = 0; j < N; j++)
                ptr[i + M * j] = coef(rows[i], cols[j]);
    }
};

void push_entry(std::vector<double> &d, double v) { d.push_back(v); }
void push_entry(std::vector<double> &d, std::complex<double> v) {
    d.push_back(v.real());
    d.push_back(v.imag());
}

template <typename V>
void write_vec(const std::string &path, const std::vector<V> &v) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(3);
    }
    if (!v.empty())
        std::fwrite(v.data(), sizeof(V), v.size(), f);
    std::fclose(f);
}

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename T>
void run_hmatrix(std::shared_ptr<Cluster<PCAGeometricClustering>> t, std::shared_ptr<Cluster<PCAGeometricClustering>> s,
                 VirtualGenerator<T> &A, std::shared_ptr<VirtualLowRankGenerator<T>> comp, const double *xt,
                 const double *pxs, std::vector<T> &x, int nt, int mu, int nreps, int dump, int reqrank, double eps,
                 double eta, double maxblocksize, const std::string &out, std::FILE *info, char uplo, char symmetry,
                 int mintd, int minsd) {
    HMatrix<T> H(t, s, eps, eta, uplo == 'N' ? 'N' : symmetry, uplo, reqrank);
    H.set_mintargetdepth(mintd);
    H.set_minsourcedepth(minsd);
    H.set_compression(comp);
    H.set_maxblocksize(static_cast<unsigned int>(maxblocksize));
    double t0 = now();
    if (uplo == 'N')
        H.build(A, xt, pxs);
    else
        H.build(A, xt);
    double t_build = now() - t0;

    // Leaf list in save_plot order (near field, then far field).
    std::vector<int> leaves;
    std::vector<double> data;
    for (auto *d : H.get_MyNearFieldMats()) {
        leaves.insert(leaves.end(), {d->get_offset_i(), d->nb_rows(), d->get_offset_j(), d->nb_cols(), -1});
        if (dump)
            for (int j = 0; j < d->nb_cols(); j++)
                for (int i = 0; i < d->nb_rows(); i++)
                    push_entry(data, (*d)(i, j));
    }
    for (auto *l : H.get_MyFarFieldMats()) {
        int k = l->rank_of();
        leaves.insert(leaves.end(), {l->get_offset_i(), l->nb_rows(), l->get_offset_j(), l->nb_cols(), k});
        if (dump) {
            for (int r = 0; r < k; r++)
                for (int i = 0; i < l->nb_rows(); i++)
                    push_entry(data, l->get_U(i, r));
            for (int j = 0; j < l->nb_cols(); j++)
                for (int r = 0; r < k; r++)
                    push_entry(data, l->get_V(r, j));
        }
    }
    write_vec(out + "/leaves.i32", leaves);
    if (dump)
        write_vec(out + "/data.f64", data);

    // Products: three untimed warm-ups (one when no timing is asked for), then
    // nreps timed ones.
    std::vector<T> y(size_t(nt) * mu, T(0));
    for (int w = 0; w < (nreps > 0 ? 3 : 1); w++)
        H.mvprod_global_to_global(x.data(), y.data(), mu);
    double tm = 0;
    for (int r = 0; r < nreps; r++) {
        t0 = now();
        H.mvprod_global_to_global(x.data(), y.data(), mu);
        tm += now() - t0;
    }
    std::vector<double> yd;
    for (const T &v : y)
        push_entry(yd, v);
    write_vec(out + "/y.f64", yd);
    std::fprintf(info, "build_s=%.9g\nmatvec_s=%.9g\nreps=%d\nnlrmat=%d\nndmat=%d\nfalse_positive=%s\ncompression=%.9g\n",
                 t_build, nreps > 0 ? tm / nreps : 0.0, nreps, H.get_nlrmat(), H.get_ndmat(),
                 H.get_infos("Number_of_false_positive").c_str(), H.compression_ratio());
    std::fclose(info);
}

} // namespace

int main(int argc, char **argv) {
    MPI_Init(&argc, &argv);
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s case.bin outdir\n", argv[0]);
        return 2;
    }
    std::ifstream in(argv[1], std::ios::binary);
    std::string out = argv[2];
    int hdr[16] = {};
    double prm[5];
    in.read(reinterpret_cast<char *>(hdr), 14 * sizeof(int));
    if (in && hdr[1] == 2)
        in.read(reinterpret_cast<char *>(hdr + 14), 2 * sizeof(int));
    in.read(reinterpret_cast<char *>(prm), sizeof(prm));
    if (!in || hdr[0] != 0x48545246 || (hdr[1] != 1 && hdr[1] != 2)) {
        std::fprintf(stderr, "bad case file\n");
        return 2;
    }
    const int mode = hdr[2], nt = hdr[3], ns = hdr[4], same = hdr[5], kernel = hdr[6], mcs = hdr[7];
    const int compressor = hdr[8], mu = hdr[9], nreps = hdr[10], dump = hdr[11], reqrank = hdr[12];
    const int up0 = hdr[13] & 0xff;
    const char uplo = (up0 == 'L' || up0 == 'U') ? (char)up0 : 'N';
    const char symmetry = (hdr[13] >> 8) == 'H' ? 'H' : 'S';
    const bool cplx = kernel == 4 || kernel == 5 || kernel == 7;
    const int mintd = hdr[14], minsd = hdr[15];
    const double eta = prm[0], eps = prm[1], delta = prm[2], diag = prm[3], maxblocksize = prm[4];
    std::vector<double> xt(3 * size_t(nt)), xs, x(size_t(ns) * mu * (cplx ? 2 : 1));
    in.read(reinterpret_cast<char *>(xt.data()), xt.size() * sizeof(double));
    if (!same) {
        xs.resize(3 * size_t(ns));
        in.read(reinterpret_cast<char *>(xs.data()), xs.size() * sizeof(double));
    }
    in.read(reinterpret_cast<char *>(x.data()), x.size() * sizeof(double));
    if (!in) {
        std::fprintf(stderr, "truncated case file\n");
        return 2;
    }
    const double *pxs = same ? xt.data() : xs.data();
    KernelGen A(nt, ns, xt.data(), pxs, kernel, same != 0, delta, diag);

    std::shared_ptr<Cluster<PCAGeometricClustering>> t = std::make_shared<Cluster<PCAGeometricClustering>>();
    std::shared_ptr<Cluster<PCAGeometricClustering>> s;
    t->set_minclustersize(mcs);
    double t0 = now();
    t->build(nt, xt.data());
    if (same) {
        s = t;
    } else {
        s = std::make_shared<Cluster<PCAGeometricClustering>>();
        s->set_minclustersize(mcs);
        s->build(ns, pxs);
    }
    double t_cluster = now() - t0;
    write_vec(out + "/perm_t.i32", t->get_perm());
    write_vec(out + "/perm_s.i32", s->get_perm());

    std::shared_ptr<VirtualLowRankGenerator<double>> comp;
    if (compressor == 0)
        comp = std::make_shared<partialACA<double>>();
    else
        comp = std::make_shared<sympartialACA<double>>();

    std::FILE *info = std::fopen((out + "/info.txt").c_str(), "w");
    std::fprintf(info, "threads=%d\ncluster_s=%.9g\n", omp_get_max_threads(), t_cluster);

    if (mode == 1) {
        // One leaf compression of the whole t x s block (lrmat tests style).
        LowRankMatrix<double> L(1, t->get_perm(), s->get_perm(), reqrank, eps);
        L.build(A, *comp, *t, xt.data(), *s, pxs);
        int k = L.rank_of();
        std::vector<int> leaves = {0, nt, 0, ns, k};
        write_vec(out + "/leaves.i32", leaves);
        std::vector<double> data;
        for (int r = 0; r < k; r++)
            for (int i = 0; i < nt; i++)
                data.push_back(L.get_U(i, r));
        for (int j = 0; j < ns; j++)
            for (int r = 0; r < k; r++)
                data.push_back(L.get_V(r, j));
        write_vec(out + "/data.f64", data);
        std::fprintf(info, "rank=%d\n", k);
        std::fclose(info);
        MPI_Finalize();
        return 0;
    }

    if (cplx) {
        KernelGenC Ac(nt, ns, xt.data(), pxs, kernel, same != 0, delta, diag);
        // kernel 7 is the reference's own generator on the same points (3-d)
        std::vector<double> p1(xt), p2(same ? xt : xs);
        GeneratorTestComplexHermitian Ah(3, nt, ns, p1, p2);
        VirtualGenerator<std::complex<double>> &Ause = kernel == 7 ? static_cast<VirtualGenerator<std::complex<double>> &>(Ah) : Ac;
        std::vector<std::complex<double>> xc(size_t(ns) * mu);
        for (size_t e = 0; e < xc.size(); e++)
            xc[e] = {x[2 * e], x[2 * e + 1]};
        auto compc = compressor == 0
                         ? std::shared_ptr<VirtualLowRankGenerator<std::complex<double>>>(
                               std::make_shared<partialACA<std::complex<double>>>())
                         : std::make_shared<sympartialACA<std::complex<double>>>();
        run_hmatrix<std::complex<double>>(t, s, Ause, compc, xt.data(), pxs, xc, nt, mu, nreps, dump, reqrank, eps, eta,
                                          maxblocksize, out, info, uplo, symmetry, mintd, minsd);
    } else {
        run_hmatrix<double>(t, s, A, comp, xt.data(), pxs, x, nt, mu, nreps, dump, reqrank, eps, eta, maxblocksize,
                            out, info, uplo, symmetry, mintd, minsd);
    }
    MPI_Finalize();
    return 0;
}
