// HIP/CDNA4 kernels of the H-matrix hot path: the kernel-entry generator, the
// batched partial-pivot ACA, the packing of the compressed leaves into the
// product layout, and the three product phases (x gather, low-rank V x, row-block
// apply). One 256-thread workgroup (four 64-wide wavefronts) is the unit of work
// everywhere.
#include "kernels.hpp"

#include <cmath>

namespace htg {
namespace {

constexpr int WG    = 256;
constexpr int NWAVE = WG / 64;
constexpr int CAPMAX = 512; // largest number of ACA cross terms per block

// ---------------------------------------------------------------------------
// Kernel entries. The squared distance is summed x, y, z from zero with
// contraction off, then sqrt and an IEEE division, so entries equal the
// reference generators' doubles (include/htool/testing/generator_test.hpp:97,
// 139; examples/smallest_example.cpp:17).
__device__ __forceinline__ double kernel_entry(const KernelDesc &k, const DevPoints &T, int i, const DevPoints &S,
                                               int j) {
#pragma clang fp contract(off)
    double r2 = 0;
    double u  = T.x[i] - S.x[j];
    r2        = r2 + u * u;
    u         = T.y[i] - S.y[j];
    r2        = r2 + u * u;
    u         = T.z[i] - S.z[j];
    r2        = r2 + u * u;
    switch (k.kind) {
    case 0:
        if (k.same && T.perm[i] == S.perm[j])
            return k.diag;
        return 1. / (4 * M_PI * sqrt(r2));
    case 1:
        return 1. / (k.delta + sqrt(r2));
    case 2:
        return 1. / (k.delta + 4 * M_PI * sqrt(r2));
    default:
        return 1. / (4 * M_PI * sqrt(k.delta + r2));
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the workgroup, same value returned to every thread, fixed order.
__device__ double block_sum(double v, double *s_red) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        s_red[w] = v;
    __syncthreads();
    double total = 0;
    for (int q = 0; q < NWAVE; q++)
        total += s_red[q];
    __syncthreads();
    return total;
}

// Index of the largest |vec[b]| over unvisited b; ties go to the largest index,
// as the reference's `if (tmp < pivot) continue;` scan does
// (include/htool/lrmat/partialACA.hpp:89-100). -1 when every entry is visited.
__device__ int block_argmax_last(const double *vec, const unsigned char *visited, int n, double *s_v, int *s_i) {
    double best = -1.0;
    int bi      = -1;
    for (int b = threadIdx.x; b < n; b += WG) {
        if (visited[b])
            continue;
        const double a = fabs(vec[b]);
        if (a >= best) {
            best = a;
            bi   = b;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64);
        const int oi    = __shfl_xor(bi, o, 64);
        if (ob > best || (ob == best && oi > bi)) {
            best = ob;
            bi   = oi;
        }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_v[w] = best;
        s_i[w] = bi;
    }
    __syncthreads();
    double rb = s_v[0];
    int ri    = s_i[0];
    for (int q = 1; q < NWAVE; q++) {
        if (s_v[q] > rb || (s_v[q] == rb && s_i[q] > ri)) {
            rb = s_v[q];
            ri = s_i[q];
        }
    }
    __syncthreads();
    return ri;
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(WG) k_dense_fill(KernelDesc k, DevPoints T, DevPoints S, const FillTile *tiles,
                                                   double *arena) {
    const FillTile ft = tiles[blockIdx.x];
    const int total   = ft.R * ft.ncols;
    for (int e = threadIdx.x; e < total; e += WG) {
        const int i = e % ft.R, j = e / ft.R;
        arena[ft.off + e] = kernel_entry(k, T, ft.row0 + i, S, ft.col0 + j);
    }
}

// ---------------------------------------------------------------------------
// Batched partial-pivot ACA, one workgroup per block
// (include/htool/lrmat/partialACA.hpp:39-171; with `swap`, sympartialACA's
// exchanged roles, include/htool/lrmat/sympartialACA.hpp:46-61, 100-104, 126-130,
// 188-206). Side 1 is the side whose entries pick the next pivot row
// (length n1), side 2 the other one (length n2). Cross term q is kept as
// ws1[q*n1 + a] (scaled column) and ws2[q*n2 + b] (residual row). FP64
// throughout: pivots, residuals and the Frobenius estimate.
__global__ void __launch_bounds__(WG) k_aca(KernelDesc kd, DevPoints T, DevPoints S, const AcaBlock *blocks,
                                            double eps, int reqrank, double *ws, unsigned char *vis_ws,
                                            const uint64_t *vis_off, int *ranks) {
    __shared__ double s_coef[CAPMAX];
    __shared__ double s_prod[CAPMAX];
    __shared__ double s_v[NWAVE];
    __shared__ int s_i[NWAVE];
    __shared__ double s_red[NWAVE];

    const AcaBlock B = blocks[blockIdx.x];
    const int M = B.M, N = B.N;
    const bool swap = B.swap != 0;
    const int n1 = swap ? N : M, n2 = swap ? M : N;
    double *U1          = ws + B.ws1;
    double *V2          = ws + B.ws2;
    unsigned char *vis1 = vis_ws + vis_off[blockIdx.x];
    unsigned char *vis2 = vis1 + n1;
    const int tid       = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;

    for (int a = tid; a < n1 + n2; a += WG)
        vis1[a] = 0;

    // First pivot: the side-1 point nearest to the cluster centre, first on ties
    // (partialACA.hpp:42-50).
    {
        const DevPoints &X = swap ? S : T;
        const int off      = swap ? B.off_s : B.off_t;
        double best        = 1e300;
        int bi             = n1;
        for (int a = tid; a < n1; a += WG) {
#pragma clang fp contract(off)
            double s = 0, u;
            u        = X.x[off + a] - B.ctr[0];
            s        = s + u * u;
            u        = X.y[off + a] - B.ctr[1];
            s        = s + u * u;
            u        = X.z[off + a] - B.ctr[2];
            s        = s + u * u;
            const double d = sqrt(s);
            if (d < best) {
                best = d;
                bi   = a;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o, 64);
            const int oi    = __shfl_xor(bi, o, 64);
            if (ob < best || (ob == best && oi < bi)) {
                best = ob;
                bi   = oi;
            }
        }
        if (lane == 0) {
            s_v[w] = best;
            s_i[w] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            double rb = s_v[0];
            int ri    = s_i[0];
            for (int q = 1; q < NWAVE; q++)
                if (s_v[q] < rb || (s_v[q] == rb && s_i[q] < ri)) {
                    rb = s_v[q];
                    ri = s_i[q];
                }
            s_i[0] = ri;
        }
        __syncthreads();
    }
    int I1 = s_i[0] < n1 ? s_i[0] : 0;
    __syncthreads();

    auto entry = [&](int a, int b) -> double {
        return swap ? kernel_entry(kd, T, B.off_t + b, S, B.off_s + a) : kernel_entry(kd, T, B.off_t + a, S, B.off_s + b);
    };

    int I2 = 0, q = 0;
    double frob = 0, aux = 0;
    bool overflow = false;
    const long long MN = (long long)M * N;
    while (((reqrank > 0) && (q < min(reqrank, min(M, N)))) || ((reqrank < 0) && (q == 0 || sqrt(aux / frob) > eps))) {
        q += 1;
        if ((long long)q * (M + N) > MN) {
            q = -1;
            break;
        }
        if (q > B.cap || q > CAPMAX) {
            overflow = true;
            break;
        }
        const int p = q - 1; // number of previous cross terms
        double *u1  = V2 + (size_t)p * n2;
        double *u2  = U1 + (size_t)p * n1;

        // Residual row: generator row, minus the previous crosses in order (partialACA.hpp:83-87).
        for (int j = tid; j < p; j += WG)
            s_coef[j] = -U1[(size_t)j * n1 + I1];
        __syncthreads();
        for (int b = tid; b < n2; b += WG) {
            double val = entry(I1, b);
            for (int j = 0; j < p; j++)
                val = fma(s_coef[j], V2[(size_t)j * n2 + b], val);
            u1[b] = val;
        }
        __syncthreads();
        const int J = block_argmax_last(u1, vis2, n2, s_v, s_i);
        if (J >= 0)
            I2 = J;
        if (tid == 0)
            vis1[I1] = 1;
        const double piv   = u1[I2];
        const double gamma = 1. / piv;
        if (fabs(piv) > 1e-15) {
            // Residual column, scaled by 1/pivot (partialACA.hpp:108-113).
            for (int j = tid; j < p; j += WG)
                s_coef[j] = -V2[(size_t)j * n2 + I2];
            __syncthreads();
            for (int a = tid; a < n1; a += WG) {
                double val = entry(a, I2);
                for (int j = 0; j < p; j++)
                    val = fma(s_coef[j], U1[(size_t)j * n1 + a], val);
                u2[a] = val * gamma;
            }
            __syncthreads();
            const int I = block_argmax_last(u2, vis1, n1, s_v, s_i);
            if (I >= 0)
                I1 = I;
            if (tid == 0)
                vis2[I2] = 1;
            if (reqrank < 0) {
                // Frobenius estimate of the approximant (partialACA.hpp:127-141).
                double d22 = 0, d11 = 0;
                for (int a = tid; a < n1; a += WG)
                    d22 += u2[a] * u2[a];
                for (int b = tid; b < n2; b += WG)
                    d11 += u1[b] * u1[b];
                d22 = block_sum(d22, s_red);
                d11 = block_sum(d11, s_red);
                aux = fabs(d22) * fabs(d11);
                for (int j = w; j < p; j += NWAVE) {
                    double x1 = 0, x2 = 0;
                    for (int b = lane; b < n2; b += 64)
                        x1 += u1[b] * V2[(size_t)j * n2 + b];
                    for (int a = lane; a < n1; a += 64)
                        x2 += u2[a] * U1[(size_t)j * n1 + a];
                    x1 = wave_sum(x1);
                    x2 = wave_sum(x2);
                    if (lane == 0)
                        s_prod[j] = x1 * x2;
                }
                __syncthreads();
                double frob_aux = 0;
                for (int j = 0; j < p; j++)
                    frob_aux += s_prod[j];
                frob += aux + 2 * frob_aux;
            }
            __syncthreads();
        } else {
            q -= 1;
            if (q == 0)
                q = -1;
            break;
        }
    }
    if (tid == 0)
        ranks[blockIdx.x] = overflow ? -3 : q;
}

// ---------------------------------------------------------------------------
// Copies compressed leaves from the ACA workspace into the product layout:
// V as a k x N row-major block, U split into one R x k column-major tile per
// row block.
__global__ void __launch_bounds__(WG) k_lr_pack(const LrPack *packs, const UTileDst *utiles, const double *ws,
                                                double *arena, double *varena) {
    const LrPack p = packs[blockIdx.x];
    const size_t nv = (size_t)p.k * p.N;
    for (size_t e = threadIdx.x; e < nv; e += WG)
        varena[p.dst_v + e] = ws[p.src_v + e];
    for (int t = 0; t < p.npieces; t++) {
        const UTileDst d = utiles[p.piece_first + t];
        const int tot    = d.R * p.k;
        for (int e = threadIdx.x; e < tot; e += WG) {
            const int i = e % d.R, r = e / d.R;
            arena[d.off + e] = ws[p.src_u + (size_t)r * p.M + d.row0 + i];
        }
    }
}

__global__ void k_convert_f32(const double *src, float *dst, uint64_t n) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x)
        dst[e] = (float)src[e];
}

// ---------------------------------------------------------------------------
// Product phase 0: x (nc x mu column-major, original order) -> xc
// (cluster order, RHS-interleaved, compute type), for RHS [mu_base, mu_base+MU).
template <typename TC, int MU>
__global__ void k_gather_x(const double *x, const int *perm_s, int nc, int mu_base, TC *xc) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nc * MU)
        return;
    const int j = e / MU, s = e % MU;
    xc[e] = (TC)x[perm_s[j] + (size_t)nc * (mu_base + s)];
}

// Product phase 1: partial sums of V x per column chunk of each low-rank leaf;
// the last chunk to finish adds the partials in chunk order into t.
template <typename TS, typename TC, int MU>
__global__ void __launch_bounds__(WG) k_lr_vx(const VTile *tiles, const TS *varena, const TC *xc, double *part,
                                              double *t, unsigned *counters) {
    __shared__ int s_last;
    const VTile d = tiles[blockIdx.x];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // part_off, part_base and t_off count k-vectors; one k-vector holds MU RHS.
    double *out = d.nparts > 1 ? part + (size_t)d.part_off * MU : t + (size_t)d.t_off * MU;
    for (int r = w; r < d.k; r += NWAVE) {
        const TS *row = varena + d.v_off + (size_t)r * d.ld;
        TC acc[MU];
#pragma unroll
        for (int s = 0; s < MU; s++)
            acc[s] = 0;
#pragma unroll 4
        for (int j = lane; j < d.ncols; j += 64) {
            const TC a   = (TC)row[j];
            const TC *xv = xc + (size_t)(d.x_off + j) * MU;
#pragma unroll
            for (int s = 0; s < MU; s++)
                acc[s] += a * xv[s];
        }
#pragma unroll
        for (int s = 0; s < MU; s++) {
            const TC v = wave_sum(acc[s]);
            if (lane == 0)
                out[r * MU + s] = (double)v;
        }
    }
    if (d.nparts > 1) {
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned prev = atomicAdd(&counters[d.leaf], 1u);
            s_last              = (prev == (unsigned)(d.nparts - 1));
        }
        __syncthreads();
        if (s_last) {
            __threadfence();
            const int km       = d.k * MU;
            const double *pb   = part + (size_t)d.part_base * MU;
            double *tout       = t + (size_t)d.t_off * MU;
            for (int e = threadIdx.x; e < km; e += WG) {
                double s = 0;
                for (int c = 0; c < d.nparts; c++)
                    s += pb[(size_t)c * km + e];
                tout[e] = s;
            }
            if (threadIdx.x == 0)
                counters[d.leaf] = 0;
        }
    }
}

// Product phase 2: one workgroup per row block (<= 128 rows). Each piece is an
// R x ncols column-major tile times a vector segment; lanes run over rows,
// thread groups over columns. Per-tile sums are in the compute type, the sum
// over tiles in FP64; the result is scattered to original order.
template <typename TS, typename TC, int MU>
__global__ void __launch_bounds__(WG) k_row_apply(const int *rb_row0, const int *piece_off, const Piece *pieces,
                                                  const TS *arena, const TC *xc, const double *t,
                                                  const int *perm_t, int nr, int mu_base, double *y) {
    constexpr int SC = MU < 8 ? MU : 8; // RHS per LDS reduction pass
    __shared__ double s_red[WG * SC];
    const int rb = blockIdx.x;
    const int r0 = rb_row0[rb], R = rb_row0[rb + 1] - r0;
    const int pshift = R <= 32 ? 5 : (R <= 64 ? 6 : 7);
    const int P = 1 << pshift, G = WG >> pshift;
    const int i = threadIdx.x & (P - 1), g = threadIdx.x >> pshift;
    double acc[MU];
#pragma unroll
    for (int s = 0; s < MU; s++)
        acc[s] = 0;
    const int pend = piece_off[rb + 1];
    for (int pi = piece_off[rb]; pi < pend; pi++) {
        const Piece pc = pieces[pi];
        TC pacc[MU];
#pragma unroll
        for (int s = 0; s < MU; s++)
            pacc[s] = 0;
        if (i < R) {
            const TS *A = arena + pc.a_off + i;
            if (pc.kind == 0) {
                const TC *xv = xc + (size_t)pc.vec_off * MU;
#pragma unroll 4
                for (int j = g; j < pc.ncols; j += G) {
                    const TC a = (TC)A[(size_t)R * j];
#pragma unroll
                    for (int s = 0; s < MU; s++)
                        pacc[s] += a * xv[j * MU + s];
                }
            } else {
                const double *tv = t + (size_t)pc.vec_off * MU;
#pragma unroll 4
                for (int j = g; j < pc.ncols; j += G) {
                    const TC a = (TC)A[(size_t)R * j];
#pragma unroll
                    for (int s = 0; s < MU; s++)
                        pacc[s] += a * (TC)tv[j * MU + s];
                }
            }
        }
#pragma unroll
        for (int s = 0; s < MU; s++)
            acc[s] += (double)pacc[s];
    }
#pragma unroll
    for (int s0 = 0; s0 < MU; s0 += SC) {
#pragma unroll
        for (int s = 0; s < SC; s++)
            s_red[s * WG + threadIdx.x] = acc[s0 + s];
        __syncthreads();
        if (g == 0 && i < R) {
            const size_t row = (size_t)perm_t[r0 + i];
#pragma unroll
            for (int s = 0; s < SC; s++) {
                double v = 0;
                for (int gg = 0; gg < G; gg++)
                    v += s_red[s * WG + gg * P + i];
                y[row + (size_t)nr * (mu_base + s0 + s)] = v;
            }
        }
        __syncthreads();
    }
}

} // namespace

// ---------------------------------------------------------------------------
void launch_dense_fill(const KernelDesc &k, DevPoints T, DevPoints S, const FillTile *tiles, int ntiles,
                       double *arena, hipStream_t st) {
    if (ntiles > 0)
        hipLaunchKernelGGL(k_dense_fill, dim3(ntiles), dim3(WG), 0, st, k, T, S, tiles, arena);
}

void launch_aca(const KernelDesc &k, DevPoints T, DevPoints S, const AcaBlock *blocks, int nblocks, double eps,
                int reqrank, double *ws, unsigned char *vis_ws, const uint64_t *vis_off, int *ranks,
                hipStream_t st) {
    if (nblocks > 0)
        hipLaunchKernelGGL(k_aca, dim3(nblocks), dim3(WG), 0, st, k, T, S, blocks, eps, reqrank, ws, vis_ws, vis_off,
                           ranks);
}

void launch_lr_pack(const LrPack *packs, int npacks, const UTileDst *utiles, const double *ws, double *arena,
                    double *varena, hipStream_t st) {
    if (npacks > 0)
        hipLaunchKernelGGL(k_lr_pack, dim3(npacks), dim3(WG), 0, st, packs, utiles, ws, arena, varena);
}

void launch_convert_f32(const double *src, float *dst, uint64_t n, hipStream_t st) {
    if (n == 0)
        return;
    uint64_t blocks = (n + WG - 1) / WG;
    if (blocks > 65536)
        blocks = 65536;
    hipLaunchKernelGGL(k_convert_f32, dim3((unsigned)blocks), dim3(WG), 0, st, src, dst, n);
}

#define HTG_MU_SWITCH(MUV, CALL) \
    switch (MUV) {               \
    case 1: CALL(1); break;      \
    case 2: CALL(2); break;      \
    case 4: CALL(4); break;      \
    case 8: CALL(8); break;      \
    case 16: CALL(16); break;    \
    case 32: CALL(32); break;    \
    default: CALL(64); break;    \
    }

template <typename TS, typename TC>
void launch_gather_x(const double *x, const int *perm_s, int nc, int /*mu_total*/, int mu_base, int MU, TC *xc,
                     hipStream_t st) {
    const int n = nc * MU;
    const dim3 grid((n + WG - 1) / WG);
#define HTG_GATHER(M) hipLaunchKernelGGL((k_gather_x<TC, M>), grid, dim3(WG), 0, st, x, perm_s, nc, mu_base, xc)
    HTG_MU_SWITCH(MU, HTG_GATHER)
#undef HTG_GATHER
}

template <typename TS, typename TC>
void launch_lr_vx(const VTile *tiles, int ntiles, const TS *varena, const TC *xc, double *part, double *t,
                  unsigned *counters, int MU, hipStream_t st) {
    if (ntiles <= 0)
        return;
#define HTG_VX(M) \
    hipLaunchKernelGGL((k_lr_vx<TS, TC, M>), dim3(ntiles), dim3(WG), 0, st, tiles, varena, xc, part, t, counters)
    HTG_MU_SWITCH(MU, HTG_VX)
#undef HTG_VX
}

template <typename TS, typename TC>
void launch_row_apply(const int *rb_row0, const int *piece_off, const Piece *pieces, int nrb, const TS *arena,
                      const TC *xc, const double *t, const int *perm_t, int nr, int mu_base, int MU, double *y,
                      hipStream_t st) {
    if (nrb <= 0)
        return;
#define HTG_RA(M)                                                                                                  \
    hipLaunchKernelGGL((k_row_apply<TS, TC, M>), dim3(nrb), dim3(WG), 0, st, rb_row0, piece_off, pieces, arena, xc, \
                       t, perm_t, nr, mu_base, y)
    HTG_MU_SWITCH(MU, HTG_RA)
#undef HTG_RA
}

template void launch_gather_x<double, double>(const double *, const int *, int, int, int, int, double *, hipStream_t);
template void launch_gather_x<float, float>(const double *, const int *, int, int, int, int, float *, hipStream_t);
template void launch_lr_vx<double, double>(const VTile *, int, const double *, const double *, double *, double *,
                                           unsigned *, int, hipStream_t);
template void launch_lr_vx<float, float>(const VTile *, int, const float *, const float *, double *, double *,
                                         unsigned *, int, hipStream_t);
template void launch_row_apply<double, double>(const int *, const int *, const Piece *, int, const double *,
                                               const double *, const double *, const int *, int, int, int,
                                               double *, hipStream_t);
template void launch_row_apply<float, float>(const int *, const int *, const Piece *, int, const float *,
                                             const float *, const double *, const int *, int, int, int, double *,
                                             hipStream_t);

} // namespace htg
