// Bandwidth probe of the device (no counterpart in the reference): a
// streaming read (non-temporal 16-byte loads) and a streaming copy of a large
// buffer, four 16-byte words in flight per thread, timed with HIP events. It gives the
// achievable HBM rate the product kernels are compared with, beside the
// vendor's 8 TB/s, and a copy of known size for calibrating the FETCH_SIZE /
// WRITE_SIZE counters.
#include "common.hpp"

#include <algorithm>

namespace {

constexpr int PWG = 256;

typedef unsigned u4v __attribute__((ext_vector_type(4)));

// Reads n16 16-byte words (non-temporal loads: the fastest form measured on
// MI355X); the XOR of every word decides one store that a real run never
// makes (acc equal to the magic value), so the loads stay.
__global__ void __launch_bounds__(PWG) k_stream_read(const uint4 *__restrict__ src, size_t n16,
                                                     unsigned *__restrict__ sink, unsigned magic) {
    const size_t stride = (size_t)gridDim.x * PWG;
    size_t i            = (size_t)blockIdx.x * PWG + threadIdx.x;
    unsigned acc        = 0;
    const u4v *s        = reinterpret_cast<const u4v *>(src);
    for (; i + 3 * stride < n16; i += 4 * stride) {
        const u4v a = __builtin_nontemporal_load(s + i), b = __builtin_nontemporal_load(s + i + stride);
        const u4v c = __builtin_nontemporal_load(s + i + 2 * stride), d = __builtin_nontemporal_load(s + i + 3 * stride);
        acc ^= a.x ^ a.y ^ a.z ^ a.w ^ b.x ^ b.y ^ b.z ^ b.w ^ c.x ^ c.y ^ c.z ^ c.w ^ d.x ^ d.y ^ d.z ^ d.w;
    }
    for (; i < n16; i += stride) {
        const uint4 a = src[i];
        acc ^= a.x ^ a.y ^ a.z ^ a.w;
    }
    if (acc == magic)
        sink[blockIdx.x] = acc;
}

__global__ void __launch_bounds__(PWG) k_stream_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                     size_t n16) {
    const size_t stride = (size_t)gridDim.x * PWG;
    size_t i            = (size_t)blockIdx.x * PWG + threadIdx.x;
    for (; i + 3 * stride < n16; i += 4 * stride) {
        const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        dst[i]              = a;
        dst[i + stride]     = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n16; i += stride)
        dst[i] = src[i];
}

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    Events() {
        HTG_HIP(hipEventCreate(&a));
        HTG_HIP(hipEventCreate(&b));
    }
    ~Events() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

} // namespace

extern "C" int htg_stream_probe(uint64_t bytes, int reps, double *read_gbps, double *copy_gbps, int *best_grid) {
    return htg::guard([&] {
        if (bytes < (1u << 20) || reps <= 0 || !read_gbps || !copy_gbps)
            throw htg::Error(HTG_ERR_INVALID, "stream probe needs >= 1 MiB, reps > 0 and result pointers");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
            throw htg::Error(HTG_ERR_HIP, "no HIP device available");
        const size_t n16 = bytes / 16;
        htg::DevArray<uint4> src, dst;
        htg::DevArray<unsigned> sink;
        src.alloc(n16);
        dst.alloc(n16);
        src.zero();
        const int grids[] = {4096, 8192, 16384, 32768, 65536};
        sink.alloc(65536);
        Events ev;
        double best_r = 0, best_c = 0;
        int bg = 0;
        for (int g : grids) {
            hipLaunchKernelGGL(k_stream_read, dim3(g), dim3(PWG), 0, nullptr, src.p, n16, sink.p, 0x9e3779b9u);
            HTG_HIP(hipEventRecord(ev.a, nullptr));
            for (int r = 0; r < reps; r++)
                hipLaunchKernelGGL(k_stream_read, dim3(g), dim3(PWG), 0, nullptr, src.p, n16, sink.p, 0x9e3779b9u);
            HTG_HIP(hipEventRecord(ev.b, nullptr));
            HTG_HIP(hipEventSynchronize(ev.b));
            float ms = 0;
            HTG_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
            const double r = 16.0 * n16 * reps / (ms * 1e-3) / 1e9;
            if (r > best_r) {
                best_r = r;
                bg     = g;
            }
            hipLaunchKernelGGL(k_stream_copy, dim3(g), dim3(PWG), 0, nullptr, src.p, dst.p, n16);
            HTG_HIP(hipEventRecord(ev.a, nullptr));
            for (int r2 = 0; r2 < reps; r2++)
                hipLaunchKernelGGL(k_stream_copy, dim3(g), dim3(PWG), 0, nullptr, src.p, dst.p, n16);
            HTG_HIP(hipEventRecord(ev.b, nullptr));
            HTG_HIP(hipEventSynchronize(ev.b));
            HTG_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
            best_c = std::max(best_c, 32.0 * n16 * reps / (ms * 1e-3) / 1e9);
        }
        HTG_HIP(hipGetLastError());
        *read_gbps = best_r;
        *copy_gbps = best_c;
        if (best_grid)
            *best_grid = bg;
    });
}

// One copy of `bytes` (a multiple of 16) from d_src to d_dst on the default
// stream with the probe's copy kernel: the known-size transfer the FETCH_SIZE
// and WRITE_SIZE counters are scaled by.
extern "C" int htg_stream_copy(const void *d_src, void *d_dst, uint64_t bytes) {
    return htg::guard([&] {
        if (!d_src || !d_dst || bytes % 16)
            throw htg::Error(HTG_ERR_INVALID, "stream copy needs device pointers and a multiple of 16 bytes");
        hipLaunchKernelGGL(k_stream_copy, dim3(4096), dim3(PWG), 0, nullptr, static_cast<const uint4 *>(d_src),
                           static_cast<uint4 *>(d_dst), (size_t)(bytes / 16));
        HTG_HIP(hipGetLastError());
        HTG_HIP(hipDeviceSynchronize());
    });
}
