// Calls the C++ interface (include/htool_amd/gpu_hmatrix.hpp) the way a user
// of the reference's HMatrix would (examples/smallest_example.cpp): build from
// points, multiply, read infos. Reads points and x from binary files, writes y.
//   adapter_check points.f64 x.f64 y.f64 n kernel delta diag minclustersize compressor
// Exit codes: 0 ok, 3 library error (message on stderr).
#include <htool_amd/gpu_hmatrix.hpp>

#include <cstdio>
#include <cstdlib>
#include <exception>
#include <vector>

static std::vector<double> read_file(const char *path, size_t n) {
    std::vector<double> v(n);
    FILE *f = std::fopen(path, "rb");
    if (!f || std::fread(v.data(), sizeof(double), n, f) != n) {
        std::fprintf(stderr, "cannot read %s\n", path);
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

int main(int argc, char **argv) {
    if (argc < 10)
        return 2;
    const int n = std::atoi(argv[4]);
    std::vector<double> p = read_file(argv[1], 3 * (size_t)n);
    std::vector<double> x = read_file(argv[2], (size_t)n);
    try {
        htool_amd::GpuHMatrix H(1e-6, 10.0);
        H.set_minclustersize(std::atoi(argv[8]));
        H.set_compression(std::atoi(argv[9]) ? HTG_SYM_PARTIAL_ACA : HTG_PARTIAL_ACA);
        htool_amd::Generator g;
        g.kernel = static_cast<htg_kernel>(std::atoi(argv[5]));
        g.delta  = std::atof(argv[6]);
        g.diag   = std::atof(argv[7]);
        H.build(g, n, p.data());
        std::vector<double> y = H * x;
        std::printf("leaves dense=%s lowrank=%s\n", H.get_infos()["Number_of_dmat"].c_str(),
                    H.get_infos()["Number_of_lrmat"].c_str());
        FILE *f = std::fopen(argv[3], "wb");
        std::fwrite(y.data(), sizeof(double), y.size(), f);
        std::fclose(f);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 3;
    }
    return 0;
}
