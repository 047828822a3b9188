// HIP/CDNA4 kernels of the H-matrix hot path: the kernel-entry generator, the
// batched partial-pivot ACA, the packing of the compressed leaves into the
// product layout, and the three product phases (x gather, low-rank V x, row-block
// apply). One 256-thread workgroup (four 64-wide wavefronts) is the unit of work
// everywhere.
#include "kernels.hpp"

#include <cmath>
#include <type_traits>

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
    case HTG_KERNEL_ZERO:
        return 0.0;
    default:
        return 1. / (4 * M_PI * sqrt(k.delta + r2));
    }
}

// Complex kernel families (entries are std::complex<double>: re then im).
// Helmholtz e^{ikr}/(4 pi r) with `diag` on the diagonal of a square matrix;
// GeneratorTestComplex (1 + i)/(4 pi r) (include/htool/testing/generator_test.hpp:110-120).
// The distance is summed as in kernel_entry.
__device__ __forceinline__ cdbl kernel_entry_c(const KernelDesc &k, const DevPoints &T, int i, const DevPoints &S,
                                               int j) {
#pragma clang fp contract(off)
    double r2 = 0;
    double u  = T.x[i] - S.x[j];
    r2        = r2 + u * u;
    u         = T.y[i] - S.y[j];
    r2        = r2 + u * u;
    u         = T.z[i] - S.z[j];
    r2        = r2 + u * u;
    const double r = sqrt(r2);
    if (k.kind == HTG_KERNEL_HELMHOLTZ) {
        if (k.same && T.perm[i] == S.perm[j])
            return cdbl{k.diag, 0.0};
        double sn, cs;
        sincos(k.wavenumber * r, &sn, &cs);
        const double d = 4 * M_PI * r;
        return cdbl{cs / d, sn / d};
    }
    if (k.kind == HTG_KERNEL_HERMITIAN_TEST) {
        // (1 + sign(x_i - x_j) i) / (1e-5 + 4 pi r): the complex division by a
        // real denominator divides both parts
        const double dx = T.x[i] - S.x[j];
        const double sg = dx > 0 ? 1.0 : (dx < 0 ? -1.0 : 0.0);
        const double d  = 1e-5 + 4 * M_PI * r;
        return cdbl{1. / d, sg / d};
    }
    const double d = 4 * M_PI * r;
    return cdbl{1. / d, 1. / d};
}

// Scalar algebra of the ACA, real and complex. Complex products and sums are
// written out (no contraction) except the residual update, an FMA pair per
// component; the oracle restates the same operations.
__device__ __forceinline__ double axpy_s(double a, double x, double y) { return fma(a, x, y); }
__device__ __forceinline__ cdbl axpy_s(cdbl a, cdbl x, cdbl y) {
    return cdbl{fma(a.re, x.re, fma(-a.im, x.im, y.re)), fma(a.re, x.im, fma(a.im, x.re, y.im))};
}
__device__ __forceinline__ double mag(double x) { return fabs(x); }
__device__ __forceinline__ double mag(cdbl x) { return hypot(x.re, x.im); }
__device__ __forceinline__ double neg(double x) { return -x; }
__device__ __forceinline__ cdbl neg(cdbl x) { return cdbl{-x.re, -x.im}; }
__device__ __forceinline__ double recip(double x) { return 1. / x; }
// 1 / (a + ib) by Smith's method (Smith, CACM 5(8), 1962).
__device__ __forceinline__ cdbl recip(cdbl x) {
#pragma clang fp contract(off)
    if (fabs(x.re) >= fabs(x.im)) {
        const double r = x.im / x.re, d = x.re + x.im * r;
        return cdbl{1. / d, -r / d};
    }
    const double r = x.re / x.im, d = x.re * r + x.im;
    return cdbl{r / d, -1. / d};
}
__device__ __forceinline__ double mul(double a, double b) { return a * b; }
__device__ __forceinline__ cdbl mul(cdbl a, cdbl b) {
#pragma clang fp contract(off)
    return cdbl{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// acc + conj(a) b (conj is the identity for reals), as Blas<T>::dot
// (include/htool/wrappers/wrapper_blas.hpp:111-117) accumulates.
__device__ __forceinline__ double cdot_acc(double acc, double a, double b) { return acc + a * b; }
__device__ __forceinline__ cdbl cdot_acc(cdbl acc, cdbl a, cdbl b) {
#pragma clang fp contract(off)
    return cdbl{acc.re + (a.re * b.re + a.im * b.im), acc.im + (a.re * b.im - a.im * b.re)};
}
__device__ __forceinline__ double add(double a, double b) { return a + b; }
__device__ __forceinline__ cdbl add(cdbl a, cdbl b) { return cdbl{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ double real_part(double a) { return a; }
__device__ __forceinline__ double real_part(cdbl a) { return a.re; }
__device__ __forceinline__ double shfl_xor_s(double v, int o) { return __shfl_xor(v, o, 64); }
__device__ __forceinline__ cdbl shfl_xor_s(cdbl v, int o) {
    return cdbl{__shfl_xor(v.re, o, 64), __shfl_xor(v.im, o, 64)};
}
template <typename T>
__device__ __forceinline__ T zero_s() {
    return T{};
}

// A double from another lane through DPP (no LDS round trip): CTRL is the DPP
// control (quad_perm, row_mirror, row_half_mirror).
template <int CTRL>
__device__ __forceinline__ double dpp_double(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (unsigned long long)(unsigned)lo | (unsigned long long)(unsigned)hi << 32);
}

// The partner's value for a butterfly stage of offset o (o a power of two,
// wave-uniform) over groups of lanes: o = 1, 2: lane ^ o (DPP quad_perm); o =
// 4, 8: the mirrored lane of the 8- or 16-lane group (DPP row_half_mirror,
// row_mirror), which has the opposite bit o and the same higher bits, so a
// butterfly over offsets from high to low still sums every lane of the group
// exactly once (and leaves each lane's row selection unchanged); larger o:
// ds_bpermute. Cuts the reduction's LDS round trips in the single-vector
// kernels.
__device__ __forceinline__ double shfl_pair(double v, int o) {
    switch (o) {
    case 1:
        return dpp_double<0xB1>(v); // quad_perm [1, 0, 3, 2]
    case 2:
        return dpp_double<0x4E>(v); // quad_perm [2, 3, 0, 1]
    case 4:
        return dpp_double<0x141>(v); // row_half_mirror
    case 8:
        return dpp_double<0x140>(v); // row_mirror
    default:
        return __shfl_xor(v, o, 64);
    }
}

// Largest of a float over the wave.
__device__ __forceinline__ float wave_max(float m) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        m = fmaxf(m, __shfl_xor(m, o, 64));
    return m;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum_s(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = add(v, shfl_xor_s(v, o));
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
template <typename T>
__device__ int block_argmax_last(const T *vec, const unsigned char *visited, int n, double *s_v, int *s_i) {
    double best = -1.0;
    int bi      = -1;
    for (int b = threadIdx.x; b < n; b += WG) {
        if (visited[b])
            continue;
        const double a = mag(vec[b]);
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
        const int i = e / ft.ncols, j = e % ft.ncols;
        arena[ft.off + (size_t)i * ft.ld + j] =
            ft.tr ? kernel_entry(k, T, ft.col0 + j, S, ft.row0 + i) : kernel_entry(k, T, ft.row0 + i, S, ft.col0 + j);
    }
}

__global__ void __launch_bounds__(WG) k_dense_fill_c(KernelDesc k, DevPoints T, DevPoints S, const FillTile *tiles,
                                                     cdbl *arena) {
    const FillTile ft = tiles[blockIdx.x];
    const int total   = ft.R * ft.ncols;
    for (int e = threadIdx.x; e < total; e += WG) {
        const int i = e / ft.ncols, j = e % ft.ncols;
        arena[ft.off + (size_t)i * ft.ld + j] = ft.tr ? kernel_entry_c(k, T, ft.col0 + j, S, ft.row0 + i)
                                                      : kernel_entry_c(k, T, ft.row0 + i, S, ft.col0 + j);
    }
}

template <typename T>
__device__ __forceinline__ T entry_s(const KernelDesc &k, const DevPoints &T_, int i, const DevPoints &S, int j);
template <>
__device__ __forceinline__ double entry_s<double>(const KernelDesc &k, const DevPoints &T_, int i, const DevPoints &S,
                                                  int j) {
    return kernel_entry(k, T_, i, S, j);
}
template <>
__device__ __forceinline__ cdbl entry_s<cdbl>(const KernelDesc &k, const DevPoints &T_, int i, const DevPoints &S,
                                              int j) {
    return kernel_entry_c(k, T_, i, S, j);
}

// ---------------------------------------------------------------------------
// Batched partial-pivot ACA, one workgroup per block
// (include/htool/lrmat/partialACA.hpp:39-171; with `swap`, sympartialACA's
// exchanged roles, include/htool/lrmat/sympartialACA.hpp:46-61, 100-104, 126-130,
// 188-206). Side 1 is the side whose entries pick the next pivot row
// (length n1), side 2 the other one (length n2). Cross term q is kept as
// ws1[q*n1 + a] (scaled column) and ws2[q*n2 + b] (residual row). FP64
// throughout: pivots, residuals and the Frobenius estimate.
// Scalar type E: double, or cdbl for the complex kernel families (magnitudes
// |.| for the pivots, conjugated dot products in the Frobenius estimate,
// frob += aux + 2 Re(frob_aux), partialACA.hpp:127-141).
template <typename E>
__global__ void __launch_bounds__(WG) k_aca(KernelDesc kd, DevPoints T, DevPoints S, const AcaBlock *blocks,
                                            double eps, int reqrank, E *ws, unsigned char *vis_ws,
                                            const uint64_t *vis_off, int *ranks) {
    __shared__ E s_coef[CAPMAX];
    __shared__ E s_prod[CAPMAX];
    __shared__ double s_v[NWAVE];
    __shared__ int s_i[NWAVE];
    __shared__ double s_red[NWAVE];

    const AcaBlock B = blocks[blockIdx.x];
    const int M = B.M, N = B.N;
    const bool swap = B.swap != 0;
    const int n1 = swap ? N : M, n2 = swap ? M : N;
    E *U1               = ws + B.ws1;
    E *V2               = ws + B.ws2;
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

    auto entry = [&](int a, int b) -> E {
        return swap ? entry_s<E>(kd, T, B.off_t + b, S, B.off_s + a) : entry_s<E>(kd, T, B.off_t + a, S, B.off_s + b);
    };

    int I2 = 0, q = 0;
    double frob = 0, aux = 0;
    bool overflow = false;
    const long long MN = B.mn;
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
        E *u1       = V2 + (size_t)p * n2;
        E *u2       = U1 + (size_t)p * n1;

        // Residual row: generator row, minus the previous crosses in order (partialACA.hpp:83-87).
        for (int j = tid; j < p; j += WG)
            s_coef[j] = neg(U1[(size_t)j * n1 + I1]);
        __syncthreads();
        for (int b = tid; b < n2; b += WG) {
            E val = entry(I1, b);
            for (int j = 0; j < p; j++)
                val = axpy_s(s_coef[j], V2[(size_t)j * n2 + b], val);
            u1[b] = val;
        }
        __syncthreads();
        const int J = block_argmax_last(u1, vis2, n2, s_v, s_i);
        if (J >= 0)
            I2 = J;
        if (tid == 0)
            vis1[I1] = 1;
        const E piv   = u1[I2];
        const E gamma = recip(piv);
        if (mag(piv) > 1e-15) {
            // Residual column, scaled by 1/pivot (partialACA.hpp:108-113).
            for (int j = tid; j < p; j += WG)
                s_coef[j] = neg(V2[(size_t)j * n2 + I2]);
            __syncthreads();
            for (int a = tid; a < n1; a += WG) {
                E val = entry(a, I2);
                for (int j = 0; j < p; j++)
                    val = axpy_s(s_coef[j], U1[(size_t)j * n1 + a], val);
                u2[a] = mul(val, gamma);
            }
            __syncthreads();
            const int I = block_argmax_last(u2, vis1, n1, s_v, s_i);
            if (I >= 0)
                I1 = I;
            if (tid == 0)
                vis2[I2] = 1;
            if (reqrank < 0) {
                // Frobenius estimate of the approximant (partialACA.hpp:127-141);
                // conj(u) u is real, so only real parts enter aux.
                double d22 = 0, d11 = 0;
                for (int a = tid; a < n1; a += WG)
                    d22 = real_part(cdot_acc(E{d22}, u2[a], u2[a]));
                for (int b = tid; b < n2; b += WG)
                    d11 = real_part(cdot_acc(E{d11}, u1[b], u1[b]));
                d22 = block_sum(d22, s_red);
                d11 = block_sum(d11, s_red);
                aux = fabs(d22) * fabs(d11);
                for (int j = w; j < p; j += NWAVE) {
                    E x1 = zero_s<E>(), x2 = zero_s<E>();
                    for (int b = lane; b < n2; b += 64)
                        x1 = cdot_acc(x1, V2[(size_t)j * n2 + b], u1[b]);
                    for (int a = lane; a < n1; a += 64)
                        x2 = cdot_acc(x2, U1[(size_t)j * n1 + a], u2[a]);
                    x1 = wave_sum_s(x1);
                    x2 = wave_sum_s(x2);
                    if (lane == 0)
                        s_prod[j] = mul(x1, x2);
                }
                __syncthreads();
                E frob_aux = zero_s<E>();
                for (int j = 0; j < p; j++)
                    frob_aux = add(frob_aux, s_prod[j]);
                frob += aux + 2 * real_part(frob_aux);
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
// Multi-workgroup partial ACA for very large blocks: the steps of k_aca
// (partialACA.hpp:39-171, sympartialACA's exchanged roles with `swap`), with
// each residual row and column cut into chunks of ACAW_CHUNK entries, one
// workgroup per chunk. The host drives the cross terms; per term:
//   k_acaw_row  residual row of pivot I1 (side 2), per chunk its argmax over
//               unvisited entries, |row|^2 and the products with earlier rows;
//   k_acaw_piv  pivot column I2 (ties to the last index), reciprocal pivot or
//               the zero-pivot exit;
//   k_acaw_col  scaled residual column (side 1), the same per-chunk partials;
//   k_acaw_fin  next pivot row, Frobenius estimate (partials summed in chunk
//               order), stopping test for the next term.
// Residual entries are the same sums in the same order as in k_aca, so the
// factors and ranks are those of the one-workgroup kernel; only the Frobenius
// estimate's sums are taken in another order.

// Loop entry of k_aca for the next cross term: the `while` test, then the
// q (M + N) > MN exit and the workspace cap.
__device__ void acaw_enter(AcaWide &st, const AcaBlock &B, int reqrank, double eps) {
    const int p   = st.q;
    const bool go = (reqrank > 0 && p < min(reqrank, min(B.M, B.N))) ||
                    (reqrank < 0 && (p == 0 || sqrt(st.aux / st.frob) > eps));
    st.active = 0;
    if (!go) {
        st.rank = p;
    } else if ((long long)(p + 1) * (B.M + B.N) > B.mn) {
        st.rank = -1;
    } else if (p + 1 > B.cap || p + 1 > CAPMAX) {
        st.rank = -3;
    } else {
        st.active = 1;
    }
}

__global__ void __launch_bounds__(WG) k_acaw_init(const AcaBlock *blocks, AcaWide *states, int reqrank, double eps,
                                                  DevPoints T, DevPoints S, unsigned char *vis_ws,
                                                  const uint64_t *vis_off) {
    __shared__ double s_v[NWAVE];
    __shared__ int s_i[NWAVE];
    const AcaBlock B = blocks[blockIdx.x];
    const bool swap  = B.swap != 0;
    const int n1 = swap ? B.N : B.M, n2 = swap ? B.M : B.N;
    unsigned char *vis1 = vis_ws + vis_off[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int a = tid; a < n1 + n2; a += WG)
        vis1[a] = 0;
    // first pivot: the side-1 point nearest to the cluster centre, first on ties (partialACA.hpp:42-50)
    const DevPoints &X = swap ? S : T;
    const int off      = swap ? B.off_s : B.off_t;
    double best        = 1e300;
    int bi             = n1;
    for (int a = tid; a < n1; a += WG) {
#pragma clang fp contract(off)
        double d2 = 0, u;
        u         = X.x[off + a] - B.ctr[0];
        d2        = d2 + u * u;
        u         = X.y[off + a] - B.ctr[1];
        d2        = d2 + u * u;
        u         = X.z[off + a] - B.ctr[2];
        d2        = d2 + u * u;
        const double d = sqrt(d2);
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
        AcaWide st = states[blockIdx.x];
        st.I1      = ri < n1 ? ri : 0;
        st.I2      = 0;
        st.q       = 0;
        st.frob    = 0;
        st.aux     = 0;
        acaw_enter(st, B, reqrank, eps);
        states[blockIdx.x] = st;
    }
}

// Residual row (side 2, kRow) or scaled residual column (side 1) of the current
// cross term over one chunk, with the chunk's partial results at `slot`. The
// generator entries come from the device kernel family, or, with gv, from
// values the host generator wrote for this step: entry e of the block's
// vector at gv[gv_off[block] + e].
template <typename E, bool kRow>
__global__ void __launch_bounds__(WG) k_acaw_vec(KernelDesc kd, DevPoints T, DevPoints S, const AcaBlock *blocks,
                                                 const AcaWide *states, const AcawTask *tasks, int slot0, E *ws,
                                                 const unsigned char *vis_ws, const uint64_t *vis_off, double *pv,
                                                 int *pi, double *dd, E *xs, const E *gv, const uint64_t *gv_off) {
    __shared__ E s_coef[CAPMAX];
    __shared__ E s_u[ACAW_CHUNK];
    __shared__ double s_v[NWAVE];
    __shared__ int s_i[NWAVE];
    __shared__ double s_red[NWAVE];
    const AcawTask tk = tasks[blockIdx.x];
    const AcaWide st  = states[tk.block];
    if (!st.active)
        return;
    const AcaBlock B = blocks[tk.block];
    const bool swap  = B.swap != 0;
    const int n1 = swap ? B.N : B.M, n2 = swap ? B.M : B.N;
    E *U1 = ws + B.ws1, *V2 = ws + B.ws2;
    const unsigned char *vis = vis_ws + vis_off[tk.block] + (kRow ? n1 : 0);
    const int p   = st.q;
    const int len = kRow ? n2 : n1;
    const int e0 = tk.chunk * ACAW_CHUNK, e1 = min(len, e0 + ACAW_CHUNK);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // row: coefficients -u_j(I1) of the previous columns; column: -v_j(I2)
    for (int j = tid; j < p; j += WG)
        s_coef[j] = kRow ? neg(U1[(size_t)j * n1 + st.I1]) : neg(V2[(size_t)j * n2 + st.I2]);
    __syncthreads();
    const E *prev = kRow ? V2 : U1; // earlier vectors of this side
    E *out        = (kRow ? V2 : U1) + (size_t)p * len;
    E g;                            // reciprocal pivot (column only)
    if constexpr (std::is_same<E, double>::value)
        g = st.gre;
    else
        g = E{st.gre, st.gim};
    double best = -1.0, dsq = 0;
    int bi = -1;
    const E *hv = gv ? gv + gv_off[tk.block] : nullptr;
    for (int e = e0 + tid; e < e1; e += WG) {
        E val = hv ? hv[e]
                   : kRow ? (swap ? entry_s<E>(kd, T, B.off_t + e, S, B.off_s + st.I1)
                                  : entry_s<E>(kd, T, B.off_t + st.I1, S, B.off_s + e))
                          : (swap ? entry_s<E>(kd, T, B.off_t + st.I2, S, B.off_s + e)
                                  : entry_s<E>(kd, T, B.off_t + e, S, B.off_s + st.I2));
        for (int j = 0; j < p; j++)
            val = axpy_s(s_coef[j], prev[(size_t)j * len + e], val);
        if (!kRow)
            val = mul(val, g);
        out[e]      = val;
        s_u[e - e0] = val;
        if (!vis[e]) {
            const double a = mag(val);
            if (a >= best) {
                best = a;
                bi   = e;
            }
        }
        dsq = real_part(cdot_acc(E{dsq}, val, val));
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
    if (lane == 0) {
        s_v[w] = best;
        s_i[w] = bi;
    }
    dsq = block_sum(dsq, s_red); // synchronises: s_v, s_i and s_u complete
    const int slot = slot0 + blockIdx.x;
    if (tid == 0) {
        double rb = s_v[0];
        int ri    = s_i[0];
        for (int q = 1; q < NWAVE; q++)
            if (s_v[q] > rb || (s_v[q] == rb && s_i[q] > ri)) {
                rb = s_v[q];
                ri = s_i[q];
            }
        pv[slot] = rb;
        pi[slot] = ri;
        dd[slot] = dsq;
    }
    // products with the earlier vectors of this side, for the Frobenius estimate
    for (int j = w; j < p; j += NWAVE) {
        E x = zero_s<E>();
        for (int e = e0 + lane; e < e1; e += 64)
            x = cdot_acc(x, prev[(size_t)j * len + e], s_u[e - e0]);
        x = wave_sum_s(x);
        if (lane == 0)
            xs[(size_t)slot * CAPMAX + j] = x;
    }
}

// Pivot column from the row chunks' partials, one thread per block.
template <typename E>
__global__ void __launch_bounds__(WG) k_acaw_piv(const AcaBlock *blocks, AcaWide *states, int nblocks, const E *ws,
                                                 unsigned char *vis_ws, const uint64_t *vis_off, const double *pv,
                                                 const int *pi) {
    const int bk = blockIdx.x * WG + threadIdx.x;
    if (bk >= nblocks)
        return;
    AcaWide st = states[bk];
    if (!st.active)
        return;
    const AcaBlock B = blocks[bk];
    const int n2     = B.swap ? B.M : B.N;
    double best      = -1.0;
    int bi           = -1;
    for (int c = 0; c < st.nch2; c++) {
        const double v = pv[st.task2 + c];
        const int i    = pi[st.task2 + c];
        if (i >= 0 && (v > best || (v == best && i > bi))) {
            best = v;
            bi   = i;
        }
    }
    if (bi >= 0)
        st.I2 = bi;
    vis_ws[vis_off[bk] + st.I1] = 1;
    const E piv = ws[B.ws2 + (size_t)st.q * n2 + st.I2];
    if (mag(piv) > 1e-15) {
        const E g = recip(piv);
        if constexpr (std::is_same<E, double>::value) {
            st.gre = g;
            st.gim = 0;
        } else {
            st.gre = g.re;
            st.gim = g.im;
        }
    } else { // k_aca: q -= 1, and q == 0 means not compressible
        st.active = 0;
        st.rank   = st.q == 0 ? -1 : st.q;
    }
    states[bk] = st;
}

// Next pivot row, Frobenius estimate and stopping test, one workgroup per block.
template <typename E>
__global__ void __launch_bounds__(WG) k_acaw_fin(const AcaBlock *blocks, AcaWide *states, int reqrank, double eps,
                                                 unsigned char *vis_ws, const uint64_t *vis_off, const double *pv,
                                                 const int *pi, const double *dd, const E *xs) {
    __shared__ E s_prod[CAPMAX];
    __shared__ int s_act;
    const int bk = blockIdx.x;
    AcaWide st   = states[bk];
    if (threadIdx.x == 0)
        s_act = st.active;
    __syncthreads();
    if (!s_act)
        return;
    const AcaBlock B = blocks[bk];
    const bool swap  = B.swap != 0;
    const int n1     = swap ? B.N : B.M;
    const int p      = st.q;
    if (reqrank < 0) {
        for (int j = threadIdx.x; j < p; j += WG) {
            E x1 = zero_s<E>(), x2 = zero_s<E>();
            for (int c = 0; c < st.nch2; c++)
                x1 = add(x1, xs[(size_t)(st.task2 + c) * CAPMAX + j]);
            for (int c = 0; c < st.nch1; c++)
                x2 = add(x2, xs[(size_t)(st.task1 + c) * CAPMAX + j]);
            s_prod[j] = mul(x1, x2);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double best = -1.0;
        int bi      = -1;
        for (int c = 0; c < st.nch1; c++) {
            const double v = pv[st.task1 + c];
            const int i    = pi[st.task1 + c];
            if (i >= 0 && (v > best || (v == best && i > bi))) {
                best = v;
                bi   = i;
            }
        }
        if (bi >= 0)
            st.I1 = bi;
        vis_ws[vis_off[bk] + n1 + st.I2] = 1;
        if (reqrank < 0) {
            double d22 = 0, d11 = 0;
            for (int c = 0; c < st.nch1; c++)
                d22 += dd[st.task1 + c];
            for (int c = 0; c < st.nch2; c++)
                d11 += dd[st.task2 + c];
            st.aux     = fabs(d22) * fabs(d11);
            E frob_aux = zero_s<E>();
            for (int j = 0; j < p; j++)
                frob_aux = add(frob_aux, s_prod[j]);
            st.frob += st.aux + 2 * real_part(frob_aux);
        }
        st.q = p + 1;
        acaw_enter(st, B, reqrank, eps);
        states[bk] = st;
    }
}

// ---------------------------------------------------------------------------
// Copies compressed leaves (from the ACA workspace or the compressed store)
// into another layout: V row by row, U into one tile per destination.
__device__ __forceinline__ double conj_if(double v, int) { return v; }
__device__ __forceinline__ cdbl conj_if(cdbl v, int c) { return c ? cdbl{v.re, -v.im} : v; }

template <typename E>
__global__ void __launch_bounds__(WG) k_lr_pack(const LrPack *packs, const UTileDst *utiles, const E *ws, E *arena,
                                                E *varena) {
    const LrPack p  = packs[blockIdx.x];
    const size_t nv = (size_t)p.k * p.N;
    for (size_t e = threadIdx.x; e < nv; e += WG) {
        const size_t r = e / p.N, j = e % p.N;
        varena[p.dst_v + r * p.dst_ldv + j] = conj_if(ws[p.src_v + r * p.src_ldv + j], p.conj);
    }
    for (int t = 0; t < p.npieces; t++) {
        const UTileDst d = utiles[p.piece_first + t];
        const int tot    = d.R * p.k;
        for (int e = threadIdx.x; e < tot; e += WG) {
            const int i = e / p.k, r = e % p.k;
            arena[d.off + (size_t)i * d.si + (size_t)r * d.sr] =
                conj_if(ws[p.src_u + (size_t)r * p.M + d.row0 + i], p.conj);
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 recompression of one low-rank leaf per workgroup (no counterpart in the
// reference; it follows the QR + SVD recompression of low-rank blocks,
// e.g. Bebendorf, "Hierarchical Matrices", 2008, sec. 1.1.5). With U = Qu Ru
// and V^T = Qv Rv (Ru, Rv from Cholesky factors of the Gram matrices, columns
// scaled to unit norm first), C = Ru Rv^T = P S Q^T by one-sided Jacobi, the
// leaf becomes U' S_r V' with U' = U Ru^-1 P_r and V' = Q_r^T Rv^-T V: unit
// columns / rows and the kept singular values S_r, r the smallest rank whose
// dropped singular values have a Frobenius norm <= tol * ||S||.
namespace rc {
constexpr int K = RC_KMAX;
constexpr int CH = 32; // rows (or columns) staged per pass

// Gram of the k vectors of a leaf factor. col_major: vector a is X[a*ld + i]
// (entries i < len); otherwise entries are strided rows of a row-major k x len block.
__device__ void gram(const double *X, size_t ld, int len, int k, double *G, double *sT) {
    double acc[(K * K + WG - 1) / WG];
    const int nent = k * k;
#pragma unroll
    for (int q = 0; q < (K * K + WG - 1) / WG; q++)
        acc[q] = 0;
    for (int i0 = 0; i0 < len; i0 += CH) {
        const int nb = min(CH, len - i0);
        for (int e = threadIdx.x; e < CH * k; e += WG) {
            const int ii = e % CH, a = e / CH;
            sT[ii * K + a] = ii < nb ? X[(size_t)a * ld + i0 + ii] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < (K * K + WG - 1) / WG; q++) {
            const int e = threadIdx.x + q * WG;
            if (e < nent) {
                const int a = e / k, b = e % k;
                double s = 0;
                for (int ii = 0; ii < CH; ii++)
                    s += sT[ii * K + a] * sT[ii * K + b];
                acc[q] += s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < (K * K + WG - 1) / WG; q++) {
        const int e = threadIdx.x + q * WG;
        if (e < nent)
            G[(e / k) * K + e % k] = acc[q];
    }
    __syncthreads();
}

// In place: G (k x k, row stride K) -> R upper triangular with R^T R = D^-1 G D^-1,
// then R <- R D (D = sqrt(diag G)). Returns false when not positive definite.
__device__ bool cholesky_scaled(double *G, int k, double *sd, int *s_ok) {
    if (threadIdx.x < k)
        sd[threadIdx.x] = sqrt(G[threadIdx.x * K + threadIdx.x]);
    if (threadIdx.x == 0)
        *s_ok = 1;
    __syncthreads();
    for (int e = threadIdx.x; e < k * k; e += WG) {
        const int a = e / k, b = e % k;
        G[a * K + b] = G[a * K + b] / (sd[a] * sd[b]);
    }
    __syncthreads();
    for (int j = 0; j < k; j++) {
        if (threadIdx.x == 0) {
            double d = G[j * K + j];
            for (int i = 0; i < j; i++)
                d -= G[i * K + j] * G[i * K + j];
            if (!(d > 1e-13))
                *s_ok = 0;
            G[j * K + j] = sqrt(fmax(d, 1e-300));
        }
        __syncthreads();
        const int c = j + 1 + threadIdx.x;
        if (c < k) {
            double v = G[j * K + c];
            for (int i = 0; i < j; i++)
                v -= G[i * K + j] * G[i * K + c];
            G[j * K + c] = v / G[j * K + j];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < k * k; e += WG) {
        const int a = e / k, b = e % k;
        G[a * K + b] = b >= a ? G[a * K + b] * sd[b] : 0.0;
    }
    const bool ok = *s_ok != 0;
    __syncthreads(); // every thread has read s_ok before the next call resets it
    return ok;
}

// Solves R X = B (R upper triangular k x k, B k x r, both row stride K) in place in B.
__device__ void upper_solve(const double *R, double *B, int k, int r) {
    const int c = threadIdx.x;
    if (c < r) {
        for (int i = k - 1; i >= 0; i--) {
            double v = B[i * K + c];
            for (int l = i + 1; l < k; l++)
                v -= R[i * K + l] * B[l * K + c];
            B[i * K + c] = v / R[i * K + i];
        }
    }
    __syncthreads();
}
} // namespace rc

__global__ void __launch_bounds__(WG) k_recompress(const RcLeaf *leaves, double *store, double tol, int *rank_out,
                                                   int *ok_out, double *sigma_out) {
    using namespace rc;
    __shared__ double sRu[K * K], sRv[K * K], sW[K * K], sQ[K * K], sT[CH * K];
    __shared__ double s_sig[K], s_d[K];
    __shared__ int s_idx[K], s_ok, s_rot, s_r;
    const RcLeaf L = leaves[blockIdx.x];
    const int k    = L.k;
    double *U      = store + L.u_off;
    double *V      = store + L.v_off;
    if (k < 2 || k > K) {
        if (threadIdx.x == 0) {
            rank_out[blockIdx.x] = k;
            ok_out[blockIdx.x]   = 0;
        }
        return;
    }
    // Gram matrices: U^T U (vector a = column a of U) and V V^T (vector a = row a of V).
    gram(U, (size_t)L.M, L.M, k, sRu, sT);
    gram(V, (size_t)L.N, L.N, k, sRv, sT);
    const bool oku = cholesky_scaled(sRu, k, s_d, &s_ok);
    const bool okv = cholesky_scaled(sRv, k, s_d, &s_ok);
    if (!oku || !okv) {
        if (threadIdx.x == 0) {
            rank_out[blockIdx.x] = k;
            ok_out[blockIdx.x]   = 0;
        }
        return;
    }
    // C = Ru Rv^T into sW (row stride K); sQ = identity.
    for (int e = threadIdx.x; e < k * k; e += WG) {
        const int a = e / k, b = e % k;
        double v = 0;
        for (int i = max(a, b); i < k; i++)
            v += sRu[a * K + i] * sRv[b * K + i];
        sW[a * K + b] = v;
        sQ[a * K + b] = a == b ? 1.0 : 0.0;
    }
    __syncthreads();
    // One-sided Jacobi on the columns of C (round-robin pairs, kk even).
    const int kk = k + (k & 1);
    if (k & 1) {
        for (int a = threadIdx.x; a < k; a += WG) {
            sW[a * K + k] = 0.0;
            sQ[a * K + k] = 0.0;
        }
        if (threadIdx.x == 0)
            sQ[k * K + k] = 1.0;
        for (int b = threadIdx.x; b < k; b += WG)
            sQ[k * K + b] = 0.0;
        __syncthreads();
    }
    const int rows = k; // rows of W; Q has kk rows when padded
    for (int sweep = 0; sweep < 40; sweep++) {
        if (threadIdx.x == 0)
            s_rot = 0;
        __syncthreads();
        for (int round = 0; round < kk - 1; round++) {
            const int t = threadIdx.x;
            if (t < kk / 2) {
                // round-robin: position 0 fixed, the others rotate
                auto at = [&](int pos) { return pos == 0 ? 0 : 1 + (pos - 1 + round) % (kk - 1); };
                int p = at(t), q = at(kk - 1 - t);
                if (p > q) {
                    const int tmp = p;
                    p = q, q = tmp;
                }
                double al = 0, be = 0, ga = 0;
                for (int a = 0; a < rows; a++) {
                    const double wp = sW[a * K + p], wq = sW[a * K + q];
                    al += wp * wp;
                    be += wq * wq;
                    ga += wp * wq;
                }
                if (fabs(ga) > 1e-15 * sqrt(al * be) && ga != 0.0) {
                    const double zeta = (be - al) / (2 * ga);
                    const double tt   = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
                    const double c = 1 / sqrt(1 + tt * tt), sn = c * tt;
                    for (int a = 0; a < rows; a++) {
                        const double wp = sW[a * K + p], wq = sW[a * K + q];
                        sW[a * K + p] = c * wp - sn * wq;
                        sW[a * K + q] = sn * wp + c * wq;
                    }
                    for (int a = 0; a < kk; a++) {
                        const double vp = sQ[a * K + p], vq = sQ[a * K + q];
                        sQ[a * K + p] = c * vp - sn * vq;
                        sQ[a * K + q] = sn * vp + c * vq;
                    }
                    s_rot = 1;
                }
            }
            __syncthreads();
        }
        if (!s_rot)
            break;
        __syncthreads();
    }
    // Singular values (column norms of W), sorted descending; rank r by the tail test.
    if (threadIdx.x < kk) {
        double v = 0;
        for (int a = 0; a < rows; a++)
            v += sW[a * K + threadIdx.x] * sW[a * K + threadIdx.x];
        s_d[threadIdx.x] = sqrt(v);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 0; j < kk; j++)
            s_idx[j] = j;
        for (int j = 1; j < kk; j++) { // insertion sort, descending
            const int id = s_idx[j];
            int l        = j - 1;
            while (l >= 0 && s_d[s_idx[l]] < s_d[id]) {
                s_idx[l + 1] = s_idx[l];
                l--;
            }
            s_idx[l + 1] = id;
        }
        double tot = 0;
        for (int j = 0; j < k; j++)
            tot += s_d[s_idx[j]] * s_d[s_idx[j]];
        int r      = k;
        double tail = 0;
        for (int j = k - 1; j >= 1; j--) {
            tail += s_d[s_idx[j]] * s_d[s_idx[j]];
            if (tail <= tol * tol * tot)
                r = j;
            else
                break;
        }
        for (int j = 0; j < k; j++)
            s_sig[j] = s_d[s_idx[j]];
        s_r = r;
    }
    __syncthreads();
    const int r = s_r;
    // A = Ru^-1 P_r (P_r columns = W columns / sigma) into sT-free scratch: reuse sW's
    // right part is unsafe, so build P_r in sQ's transpose order first.
    // Bt = Rv^-1 Q_r.
    __shared__ double sA[K * K], sB[K * K];
    for (int e = threadIdx.x; e < k * r; e += WG) {
        const int a = e / r, c = e % r;
        const int id = s_idx[c];
        sA[a * K + c] = sW[a * K + id] / s_sig[c];
        sB[a * K + c] = sQ[a * K + id];
    }
    __syncthreads();
    upper_solve(sRu, sA, k, r);
    upper_solve(sRv, sB, k, r);
    // U <- U A (rows staged through sT), V <- B^T V (columns staged through sT).
    for (int i0 = 0; i0 < L.M; i0 += CH) {
        const int nb = min(CH, L.M - i0);
        for (int e = threadIdx.x; e < CH * k; e += WG) {
            const int ii = e % CH, a = e / CH;
            sT[ii * K + a] = ii < nb ? U[(size_t)a * L.M + i0 + ii] : 0.0;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < CH * r; e += WG) {
            const int ii = e % CH, c = e / CH;
            if (ii < nb) {
                double v = 0;
                for (int a = 0; a < k; a++)
                    v += sT[ii * K + a] * sA[a * K + c];
                U[(size_t)c * L.M + i0 + ii] = v;
            }
        }
        __syncthreads();
    }
    for (int j0 = 0; j0 < L.N; j0 += CH) {
        const int nb = min(CH, L.N - j0);
        for (int e = threadIdx.x; e < CH * k; e += WG) {
            const int jj = e % CH, a = e / CH;
            sT[jj * K + a] = jj < nb ? V[(size_t)a * L.N + j0 + jj] : 0.0;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < CH * r; e += WG) {
            const int jj = e % CH, c = e / CH;
            if (jj < nb) {
                double v = 0;
                for (int a = 0; a < k; a++)
                    v += sB[a * K + c] * sT[jj * K + a];
                V[(size_t)c * L.N + j0 + jj] = v;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        rank_out[blockIdx.x] = r;
        ok_out[blockIdx.x]   = 1;
    }
    for (int j = threadIdx.x; j < r; j += WG)
        sigma_out[(size_t)blockIdx.x * K + j] = s_sig[j];
}


// ---------------------------------------------------------------------------
// Complex FP64 recompression of one low-rank leaf per workgroup: the real
// kernel's steps with Hermitian Grams and unitary rotations. A = U V with
// U^H U = Ru^H Ru and V V^H = Rv^H Rv (Cholesky of the scaled Grams), so
// A = (U Ru^-1)(Ru Rv^H)(Rv^-H V) with orthonormal outer factors; C = Ru Rv^H =
// P S Q^H by one-sided Jacobi (each pair of columns first turned so that their
// inner product is real, then the real rotation), U' = U Ru^-1 P_r and
// V' = (Rv^-1 Q_r)^H V: unit columns / rows, the kept singular values as
// scales, r the smallest rank whose dropped tail is <= tol * ||S||.
namespace rcc {
constexpr int K  = RC_KMAX_C;
constexpr int CH = 32;

__device__ __forceinline__ cdbl cj(cdbl a) { return cdbl{a.re, -a.im}; }
__device__ __forceinline__ cdbl cmul(cdbl a, cdbl b) {
    return cdbl{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cdbl csub(cdbl a, cdbl b) { return cdbl{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cdbl cadd(cdbl a, cdbl b) { return cdbl{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cdbl cdiv_real(cdbl a, double d) { return cdbl{a.re / d, a.im / d}; }
// a / b for b != 0 (Smith's method, as the ACA's reciprocal)
__device__ __forceinline__ cdbl cdiv(cdbl a, cdbl b) {
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return cdbl{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im, d = b.re * r + b.im;
    return cdbl{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

// The Hermitian Gram of k vectors of length len, vector a at X[a * ld + i]:
// G[a][b] = sum_i conj(x_a[i]) x_b[i] for U's columns (U^H U), or
// sum_i x_a[i] conj(x_b[i]) for V's rows (V V^H, vrows).
__device__ void gram(const cdbl *X, size_t ld, int len, int k, bool vrows, cdbl *G, cdbl *sT) {
    constexpr int NQ = (K * K + WG - 1) / WG;
    cdbl acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++)
        acc[q] = cdbl{0, 0};
    const int nent = k * k;
    for (int i0 = 0; i0 < len; i0 += CH) {
        const int nb = min(CH, len - i0);
        for (int e = threadIdx.x; e < CH * k; e += WG) {
            const int ii = e % CH, a = e / CH;
            sT[ii * K + a] = ii < nb ? X[(size_t)a * ld + i0 + ii] : cdbl{0, 0};
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int e = threadIdx.x + q * WG;
            if (e < nent) {
                const int a = e / k, b = e % k;
                cdbl s{0, 0};
                for (int ii = 0; ii < CH; ii++)
                    s = cadd(s, vrows ? cmul(sT[ii * K + a], cj(sT[ii * K + b]))
                                      : cmul(cj(sT[ii * K + a]), sT[ii * K + b]));
                acc[q] = cadd(acc[q], s);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int e = threadIdx.x + q * WG;
        if (e < nent)
            G[(e / k) * K + e % k] = acc[q];
    }
    __syncthreads();
}

// In place: Hermitian positive definite G -> upper triangular R with
// R^H R = D^-1 G D^-1, then R <- R D (D = sqrt(diag G), real). false when not
// positive definite.
__device__ bool cholesky_scaled(cdbl *G, int k, double *sd, int *s_ok) {
    if (threadIdx.x < k)
        sd[threadIdx.x] = sqrt(G[threadIdx.x * K + threadIdx.x].re);
    if (threadIdx.x == 0)
        *s_ok = 1;
    __syncthreads();
    for (int e = threadIdx.x; e < k * k; e += WG) {
        const int a = e / k, b = e % k;
        G[a * K + b] = cdiv_real(G[a * K + b], sd[a] * sd[b]);
    }
    __syncthreads();
    for (int j = 0; j < k; j++) {
        if (threadIdx.x == 0) {
            double d = G[j * K + j].re;
            for (int i = 0; i < j; i++)
                d -= G[i * K + j].re * G[i * K + j].re + G[i * K + j].im * G[i * K + j].im;
            if (!(d > 1e-13))
                *s_ok = 0;
            G[j * K + j] = cdbl{sqrt(fmax(d, 1e-300)), 0.0};
        }
        __syncthreads();
        const int c = j + 1 + threadIdx.x;
        if (c < k) {
            cdbl v = G[j * K + c];
            for (int i = 0; i < j; i++)
                v = csub(v, cmul(cj(G[i * K + j]), G[i * K + c]));
            G[j * K + c] = cdiv_real(v, G[j * K + j].re);
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < k * k; e += WG) {
        const int a = e / k, b = e % k;
        G[a * K + b] = b >= a ? cdbl{G[a * K + b].re * sd[b], G[a * K + b].im * sd[b]} : cdbl{0, 0};
    }
    const bool ok = *s_ok != 0;
    __syncthreads();
    return ok;
}

// R X = B in place in B (R upper triangular k x k, B k x r).
__device__ void upper_solve(const cdbl *R, cdbl *B, int k, int r) {
    const int c = threadIdx.x;
    if (c < r) {
        for (int i = k - 1; i >= 0; i--) {
            cdbl v = B[i * K + c];
            for (int l = i + 1; l < k; l++)
                v = csub(v, cmul(R[i * K + l], B[l * K + c]));
            B[i * K + c] = cdiv(v, R[i * K + i]);
        }
    }
    __syncthreads();
}
} // namespace rcc

__global__ void __launch_bounds__(WG) k_recompress_c(const RcLeaf *leaves, cdbl *store, double tol, int *rank_out,
                                                     int *ok_out, double *sigma_out) {
    using namespace rcc;
    __shared__ cdbl sRu[K * K], sRv[K * K], sW[K * K], sQ[K * K], sA[K * K], sT[CH * K];
    __shared__ double s_sig[K], s_d[K];
    __shared__ int s_idx[K], s_ok, s_rot, s_r;
    const RcLeaf L = leaves[blockIdx.x];
    const int k    = L.k;
    cdbl *U        = store + L.u_off;
    cdbl *V        = store + L.v_off;
    if (k < 2 || k > K) {
        if (threadIdx.x == 0) {
            rank_out[blockIdx.x] = k;
            ok_out[blockIdx.x]   = 0;
        }
        return;
    }
    gram(U, (size_t)L.M, L.M, k, false, sRu, sT); // U^H U
    gram(V, (size_t)L.N, L.N, k, true, sRv, sT);  // V V^H
    const bool oku = cholesky_scaled(sRu, k, s_d, &s_ok);
    const bool okv = cholesky_scaled(sRv, k, s_d, &s_ok);
    if (!oku || !okv) {
        if (threadIdx.x == 0) {
            rank_out[blockIdx.x] = k;
            ok_out[blockIdx.x]   = 0;
        }
        return;
    }
    // C = Ru Rv^H into sW; sQ = identity.
    for (int e = threadIdx.x; e < k * k; e += WG) {
        const int a = e / k, b = e % k;
        cdbl v{0, 0};
        for (int i = max(a, b); i < k; i++)
            v = cadd(v, cmul(sRu[a * K + i], cj(sRv[b * K + i])));
        sW[a * K + b] = v;
        sQ[a * K + b] = cdbl{a == b ? 1.0 : 0.0, 0.0};
    }
    __syncthreads();
    const int kk = k + (k & 1);
    if (k & 1) {
        for (int a = threadIdx.x; a < k; a += WG) {
            sW[a * K + k] = cdbl{0, 0};
            sQ[a * K + k] = cdbl{0, 0};
        }
        if (threadIdx.x == 0)
            sQ[k * K + k] = cdbl{1.0, 0.0};
        for (int b = threadIdx.x; b < k; b += WG)
            sQ[k * K + b] = cdbl{0, 0};
        __syncthreads();
    }
    const int rows = k;
    for (int sweep = 0; sweep < 40; sweep++) {
        if (threadIdx.x == 0)
            s_rot = 0;
        __syncthreads();
        for (int round = 0; round < kk - 1; round++) {
            const int t = threadIdx.x;
            if (t < kk / 2) {
                auto at = [&](int pos) { return pos == 0 ? 0 : 1 + (pos - 1 + round) % (kk - 1); };
                int p = at(t), q = at(kk - 1 - t);
                if (p > q) {
                    const int tmp = p;
                    p = q, q = tmp;
                }
                double al = 0, be = 0;
                cdbl ga{0, 0}; // w_p^H w_q
                for (int a = 0; a < rows; a++) {
                    const cdbl wp = sW[a * K + p], wq = sW[a * K + q];
                    al += wp.re * wp.re + wp.im * wp.im;
                    be += wq.re * wq.re + wq.im * wq.im;
                    ga = cadd(ga, cmul(cj(wp), wq));
                }
                const double g = sqrt(ga.re * ga.re + ga.im * ga.im);
                if (g > 1e-15 * sqrt(al * be) && g != 0.0) {
                    // column q times e^{-i arg ga}: the pair's inner product becomes g (real)
                    const cdbl ph{ga.re / g, -ga.im / g};
                    const double zeta = (be - al) / (2 * g);
                    const double tt   = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
                    const double c = 1 / sqrt(1 + tt * tt), sn = c * tt;
                    for (int a = 0; a < rows; a++) {
                        const cdbl wp = sW[a * K + p], wq = cmul(sW[a * K + q], ph);
                        sW[a * K + p] = cdbl{c * wp.re - sn * wq.re, c * wp.im - sn * wq.im};
                        sW[a * K + q] = cdbl{sn * wp.re + c * wq.re, sn * wp.im + c * wq.im};
                    }
                    for (int a = 0; a < kk; a++) {
                        const cdbl vp = sQ[a * K + p], vq = cmul(sQ[a * K + q], ph);
                        sQ[a * K + p] = cdbl{c * vp.re - sn * vq.re, c * vp.im - sn * vq.im};
                        sQ[a * K + q] = cdbl{sn * vp.re + c * vq.re, sn * vp.im + c * vq.im};
                    }
                    s_rot = 1;
                }
            }
            __syncthreads();
        }
        if (!s_rot)
            break;
        __syncthreads();
    }
    if (threadIdx.x < kk) {
        double v = 0;
        for (int a = 0; a < rows; a++)
            v += sW[a * K + threadIdx.x].re * sW[a * K + threadIdx.x].re +
                 sW[a * K + threadIdx.x].im * sW[a * K + threadIdx.x].im;
        s_d[threadIdx.x] = sqrt(v);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 0; j < kk; j++)
            s_idx[j] = j;
        for (int j = 1; j < kk; j++) {
            const int id = s_idx[j];
            int l        = j - 1;
            while (l >= 0 && s_d[s_idx[l]] < s_d[id]) {
                s_idx[l + 1] = s_idx[l];
                l--;
            }
            s_idx[l + 1] = id;
        }
        double tot = 0;
        for (int j = 0; j < k; j++)
            tot += s_d[s_idx[j]] * s_d[s_idx[j]];
        int r       = k;
        double tail = 0;
        for (int j = k - 1; j >= 1; j--) {
            tail += s_d[s_idx[j]] * s_d[s_idx[j]];
            if (tail <= tol * tol * tot)
                r = j;
            else
                break;
        }
        for (int j = 0; j < k; j++)
            s_sig[j] = s_d[s_idx[j]];
        s_r = r;
    }
    __syncthreads();
    const int r = s_r;
    // A = Ru^-1 P_r (P_r: W's columns over sigma, sorted); B = Rv^-1 Q_r written
    // over W once A has been read out of it.
    for (int e = threadIdx.x; e < k * r; e += WG) {
        const int a = e / r, c = e % r;
        sA[a * K + c] = cdiv_real(sW[a * K + s_idx[c]], s_sig[c]);
    }
    __syncthreads();
    cdbl *sB = sW;
    for (int e = threadIdx.x; e < k * r; e += WG) {
        const int a = e / r, c = e % r;
        sB[a * K + c] = sQ[a * K + s_idx[c]];
    }
    __syncthreads();
    upper_solve(sRu, sA, k, r);
    upper_solve(sRv, sB, k, r);
    // U <- U A, V <- B^H V.
    for (int i0 = 0; i0 < L.M; i0 += CH) {
        const int nb = min(CH, L.M - i0);
        for (int e = threadIdx.x; e < CH * k; e += WG) {
            const int ii = e % CH, a = e / CH;
            sT[ii * K + a] = ii < nb ? U[(size_t)a * L.M + i0 + ii] : cdbl{0, 0};
        }
        __syncthreads();
        for (int e = threadIdx.x; e < CH * r; e += WG) {
            const int ii = e % CH, c = e / CH;
            if (ii < nb) {
                cdbl v{0, 0};
                for (int a = 0; a < k; a++)
                    v = cadd(v, cmul(sT[ii * K + a], sA[a * K + c]));
                U[(size_t)c * L.M + i0 + ii] = v;
            }
        }
        __syncthreads();
    }
    for (int j0 = 0; j0 < L.N; j0 += CH) {
        const int nb = min(CH, L.N - j0);
        for (int e = threadIdx.x; e < CH * k; e += WG) {
            const int jj = e % CH, a = e / CH;
            sT[jj * K + a] = jj < nb ? V[(size_t)a * L.N + j0 + jj] : cdbl{0, 0};
        }
        __syncthreads();
        for (int e = threadIdx.x; e < CH * r; e += WG) {
            const int jj = e % CH, c = e / CH;
            if (jj < nb) {
                cdbl v{0, 0};
                for (int a = 0; a < k; a++)
                    v = cadd(v, cmul(cj(sB[a * K + c]), sT[jj * K + a]));
                V[(size_t)c * L.N + j0 + jj] = v;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        rank_out[blockIdx.x] = r;
        ok_out[blockIdx.x]   = 1;
    }
    for (int j = threadIdx.x; j < r; j += WG)
        sigma_out[(size_t)blockIdx.x * RC_KMAX + j] = s_sig[j];
}


// ---------------------------------------------------------------------------
// Product phase 0: x (nc x mu column-major, original order) -> x part of z
// (cluster order, compute type) for RHS [mu_base, mu_base+MU). Entries
// [nc, tpad) are written as zeros on every call: they meet the zero padding of
// V rows, and must hold finite values whatever an earlier product of another
// width or type left in the buffer.
// Layouts of z: RHS-interleaved, entry j of RHS s at j*MU + s; or, for the
// 32-wide matrix-core passes, blocked by 4 entries: zblk(j, s).
__device__ __forceinline__ size_t zblk(size_t j, int s) { return (j >> 2) * 128 + (size_t)s * 4 + (j & 3); }

template <typename TC, int MU, bool BLK>
__global__ void k_gather_x(const double *x, const int *perm_s, int nc, int tpad, int mu_base, TC *z) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= tpad * MU)
        return;
    int j, s;
    if constexpr (BLK) {
        j = (e >> 7) * 4 + (e & 3);
        s = (e >> 2) & 31;
    } else {
        j = e / MU;
        s = e % MU;
    }
    z[e] = j < nc ? (TC)x[perm_s[j] + (size_t)nc * (mu_base + s)] : (TC)0;
}

// The same for a 32-wide blocked pass, in two coalesced steps: (1) RHS s of
// original entry i goes to row iperm[i] of an RHS-interleaved staging copy
// (consecutive i read consecutive x; each thread writes 32 contiguous bytes);
// (2) the staging copy is reordered into the blocked layout, zeros from nc on.
__global__ void k_scatter_x32(const double *x, const int *iperm_s, int nc, int mu_base, float *stage) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc)
        return;
    const int s0 = 8 * blockIdx.y;
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; u++)
        v[u] = (float)x[i + (size_t)nc * (mu_base + s0 + u)];
    float4 *o = reinterpret_cast<float4 *>(stage + (size_t)iperm_s[i] * 32 + s0);
    o[0]      = make_float4(v[0], v[1], v[2], v[3]);
    o[1]      = make_float4(v[4], v[5], v[6], v[7]);
}

// FP64 32-wide passes keep z RHS-interleaved, the staging layout itself: one
// coalesced step, RHS s0..s0+7 of original entry i to z[iperm[i] * 32 + s0..],
// and zeros for the padding entries [nc, tpad).
__global__ void k_scatter_x64(const double *x, const int *iperm_s, int nc, int tpad, int mu_base, double *z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= tpad)
        return;
    const int s0 = 8 * blockIdx.y;
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++)
        v[u] = i < nc ? x[i + (size_t)nc * (mu_base + s0 + u)] : 0.0;
    double2 *o = reinterpret_cast<double2 *>(z + (size_t)(i < nc ? iperm_s[i] : i) * 32 + s0);
#pragma unroll
    for (int u = 0; u < 4; u++)
        o[u] = make_double2(v[2 * u], v[2 * u + 1]);
}

// Workgroup b (256 threads, 32 entries) also writes the largest |x| of its
// entries to xmax[b] when xmax is given (the FP16 scale of the 32-RHS V phase),
// and workgroup 0 clears tmax.
__global__ void __launch_bounds__(256) k_block_x32(const float *stage, int nc, int tpad, float *z, float *xmax,
                                                   unsigned *tmax) {
    __shared__ float s_m[4];
    const int e = blockIdx.x * blockDim.x + threadIdx.x; // (group, RHS)
    if (tmax && e < TMAX_SLOTS)
        tmax[e] = 0u; // the V phase that follows keeps its largest class-1 |t| here
    float m     = 0.0f;
    if (e < tpad * 8) {
        const int g = e >> 5, s = e & 31;
        float v[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int j = 4 * g + c;
            v[c]        = j < nc ? stage[(size_t)j * 32 + s] : 0.0f;
            m           = fmaxf(m, fabsf(v[c]));
        }
        reinterpret_cast<float4 *>(z)[e] = make_float4(v[0], v[1], v[2], v[3]);
    }
    if (xmax) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1)
            m = fmaxf(m, __shfl_xor(m, o, 64));
        if ((threadIdx.x & 63) == 0)
            s_m[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0)
            xmax[blockIdx.x] = fmaxf(fmaxf(s_m[0], s_m[1]), fmaxf(s_m[2], s_m[3]));
    }
}

// 16-byte vector loads of the stored entries.
template <typename TS>
struct Vec;
// `load` is an ordinary 16-byte load (vector entries, reused from the caches);
// `stream` a non-temporal one for matrix entries, which every product reads
// exactly once: measured on MI355X, a streaming read of 1-4 GiB with
// non-temporal 16-byte loads runs at 6.8-7.0 TB/s against 5.5-6.2 TB/s with
// ordinary ones, and the entries no longer displace the vector entries from L2.
typedef unsigned nt_u4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ nt_u4 nt_load16(const void *p) {
    return __builtin_nontemporal_load(reinterpret_cast<const nt_u4 *>(p));
}
template <>
struct Vec<float> {
    static constexpr int W = 4;
    __device__ static void load(const float *p, float (&a)[4]) {
        const float4 v = *reinterpret_cast<const float4 *>(p);
        a[0] = v.x, a[1] = v.y, a[2] = v.z, a[3] = v.w;
    }
    __device__ static void stream(const float *p, float (&a)[4]) {
        const nt_u4 v = nt_load16(p);
        a[0] = __uint_as_float(v.x), a[1] = __uint_as_float(v.y), a[2] = __uint_as_float(v.z),
        a[3] = __uint_as_float(v.w);
    }
};
template <>
struct Vec<double> {
    static constexpr int W = 2;
    __device__ static void load(const double *p, double (&a)[2]) {
        const double2 v = *reinterpret_cast<const double2 *>(p);
        a[0] = v.x, a[1] = v.y;
    }
    __device__ static void stream(const double *p, double (&a)[2]) {
        const nt_u4 v = nt_load16(p);
        a[0] = __hiloint2double((int)v.y, (int)v.x);
        a[1] = __hiloint2double((int)v.w, (int)v.z);
    }
};
template <>
struct Vec<__half> {
    static constexpr int W = 8;
    __device__ static void load(const __half *p, __half (&a)[8]) {
        const uint4 v = *reinterpret_cast<const uint4 *>(p);
        const __half *h = reinterpret_cast<const __half *>(&v);
#pragma unroll
        for (int u = 0; u < 8; u++)
            a[u] = h[u];
    }
    __device__ static void stream(const __half *p, __half (&a)[8]) {
        const nt_u4 v = nt_load16(p);
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
        const __half *h = reinterpret_cast<const __half *>(w);
#pragma unroll
        for (int u = 0; u < 8; u++)
            a[u] = h[u];
    }
};
template <>
struct Vec<__hip_bfloat16> {
    static constexpr int W = 8;
    __device__ static void load(const __hip_bfloat16 *p, __hip_bfloat16 (&a)[8]) {
        const uint4 v = *reinterpret_cast<const uint4 *>(p);
        const __hip_bfloat16 *h = reinterpret_cast<const __hip_bfloat16 *>(&v);
#pragma unroll
        for (int u = 0; u < 8; u++)
            a[u] = h[u];
    }
    __device__ static void stream(const __hip_bfloat16 *p, __hip_bfloat16 (&a)[8]) {
        const nt_u4 v = nt_load16(p);
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
        const __hip_bfloat16 *h = reinterpret_cast<const __hip_bfloat16 *>(w);
#pragma unroll
        for (int u = 0; u < 8; u++)
            a[u] = h[u];
    }
};

// Stored entries as raw 16-byte words, converted to FP32 only when used
// (16-bit formats stay packed two to a register until then).
template <typename TS>
struct Raw;
template <>
struct Raw<float> {
    static constexpr int W = 4;
    __device__ static void cvt(const nt_u4 &v, float (&f)[4]) {
        f[0] = __uint_as_float(v.x), f[1] = __uint_as_float(v.y), f[2] = __uint_as_float(v.z),
        f[3] = __uint_as_float(v.w);
    }
};
template <>
struct Raw<__half> {
    static constexpr int W = 8;
    __device__ static void cvt(const nt_u4 &v, float (&f)[8]) {
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            f[2 * k]     = __half2float(__ushort_as_half((unsigned short)(w[k] & 0xffffu)));
            f[2 * k + 1] = __half2float(__ushort_as_half((unsigned short)(w[k] >> 16)));
        }
    }
};
template <>
struct Raw<__hip_bfloat16> {
    static constexpr int W = 8;
    __device__ static void cvt(const nt_u4 &v, float (&f)[8]) {
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            f[2 * k]     = __uint_as_float(w[k] << 16);
            f[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
        }
    }
};

// W consecutive entries of the compute type (16-byte aligned), in 16-byte loads.
template <typename TC, int W>
__device__ __forceinline__ void load_n(const TC *p, TC (&a)[W]) {
    constexpr int VW = Vec<TC>::W;
    static_assert(W % VW == 0, "vector width");
#pragma unroll
    for (int h = 0; h < W / VW; h++) {
        TC t[VW];
        Vec<TC>::load(p + h * VW, t);
#pragma unroll
        for (int u = 0; u < VW; u++)
            a[h * VW + u] = t[u];
    }
}

// Stored entry -> compute type.
template <typename TC, typename TS>
__device__ __forceinline__ TC cvt(TS v) {
    return (TC)v;
}
template <>
__device__ __forceinline__ float cvt<float, __half>(__half v) {
    return __half2float(v);
}
template <>
__device__ __forceinline__ float cvt<float, __hip_bfloat16>(__hip_bfloat16 v) {
    return __bfloat162float(v);
}

// Lanes per row for a row of nvec 16-byte vectors: a power of two between 8
// (one 128-byte line per load) and 64, aiming at about four loads per lane.
__device__ __forceinline__ int lanes_per_row(int nvec) {
    int lpr = 8;
    while (lpr < 64 && lpr * 4 < nvec)
        lpr <<= 1;
    return lpr;
}


// Row-major block times an LDS-staged vector segment. Rows [0, nrows) with
// row stride ld, nvec 16-byte vectors of columns per row. Each row is taken by
// a group of lanes of one wave (lanes_per_row), so short rows still fill the
// wave; the group sums its products in TC, reduces in FP64 with shuffles, and
// its first lane adds the row total into s_acc[row*MU + s]. stage() fills s_z
// (every thread calls it once; it ends with a workgroup barrier): for one FP32
// right-hand side the first matrix vectors are requested before it, so their
// latency overlaps the staging's.
template <typename TS, typename TC, int MU, typename Stage>
__device__ __forceinline__ void rows_gemv(const TS *base, size_t ld, int nrows, int nvec, const TC *s_z,
                                          double *s_acc, Stage stage) {
    constexpr int W = Vec<TS>::W;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lpr  = lanes_per_row(nvec);
    const int rpw  = 64 / lpr;
    const int sl = lane & (lpr - 1), sub = lane / lpr;
    if constexpr (MU == 1 && std::is_same<TC, float>::value) {
        // Two rows per lane group, software pipelined: the loads of step
        // it + 1 (one 16-byte vector of both rows per lane) are issued before
        // the products and the row reduction of step it. Step it takes row
        // pair it / nseg and vectors [seg lpr, (seg + 1) lpr) of its rows.
        const int wu      = __builtin_amdgcn_readfirstlane(w);
        const int nseg    = (nvec + lpr - 1) / lpr;
        const int rstep   = NWAVE * 2 * rpw;
        const int r00     = wu * 2 * rpw + sub;
        const int npair   = r00 - sub < nrows ? (nrows - (r00 - sub) + rstep - 1) / rstep : 0;
        const int nit     = npair * nseg;
        const nt_u4 zero4 = {0u, 0u, 0u, 0u};
        auto issue = [&](int it, nt_u4 &ra, nt_u4 &rb) {
            const int pr = it / nseg, q = (it - pr * nseg) * lpr + sl;
            const int r0 = r00 + pr * rstep, r1 = r0 + rpw;
            ra = zero4;
            rb = zero4;
            if (q < nvec && r0 < nrows)
                ra = nt_load16(base + (size_t)r0 * ld + (size_t)q * W);
            if (q < nvec && r1 < nrows)
                rb = nt_load16(base + (size_t)r1 * ld + (size_t)q * W);
        };
        float acc0 = 0.f, acc1 = 0.f;
        auto consume = [&](int it, const nt_u4 &ra, const nt_u4 &rb) {
            const int pr = it / nseg, sg = it - pr * nseg, q = sg * lpr + sl;
            float xq[W], fa[W], fb[W];
#pragma unroll
            for (int u = 0; u < W; u++)
                xq[u] = s_z[(q < nvec ? q : 0) * W + u];
            Raw<TS>::cvt(ra, fa);
            Raw<TS>::cvt(rb, fb);
#pragma unroll
            for (int u = 0; u < W; u++) {
                acc0 += fa[u] * xq[u];
                acc1 += fb[u] * xq[u];
            }
            if (sg == nseg - 1) {
                const int r0 = r00 + pr * rstep, r1 = r0 + rpw;
                double v0 = (double)acc0, v1 = (double)acc1;
                for (int o = lpr >> 1; o > 0; o >>= 1) {
                    v0 += __shfl_xor(v0, o, 64);
                    v1 += __shfl_xor(v1, o, 64);
                }
                if (sl == 0 && r0 < nrows)
                    s_acc[r0] += v0;
                if (sl == 0 && r1 < nrows)
                    s_acc[r1] += v1;
                acc0 = 0.f, acc1 = 0.f;
            }
        };
        nt_u4 a0, b0, a1, b1;
        if (nit > 0)
            issue(0, a0, b0);
        if (nit > 1)
            issue(1, a1, b1);
        stage();
        for (int it = 0; it < nit; it += 2) {
            consume(it, a0, b0);
            if (it + 2 < nit)
                issue(it + 2, a0, b0);
            if (it + 1 < nit) {
                consume(it + 1, a1, b1);
                if (it + 3 < nit)
                    issue(it + 3, a1, b1);
            }
        }
        return;
    }
    stage();
    if constexpr (MU == 1) {
        // two rows per lane group: one staged vector read serves both
        for (int rbase = w * 2 * rpw; rbase < nrows; rbase += NWAVE * 2 * rpw) {
            const int r0 = rbase + sub, r1 = r0 + rpw;
            const bool ok0 = r0 < nrows, ok1 = r1 < nrows;
            const TS *row0 = base + (size_t)(ok0 ? r0 : 0) * ld;
            const TS *row1 = base + (size_t)(ok1 ? r1 : 0) * ld;
            TC acc0 = 0, acc1 = 0;
            if (ok0) {
#pragma unroll 2
                for (int q = sl; q < nvec; q += lpr) {
                    TS a[W], b[W];
                    Vec<TS>::stream(row0 + (size_t)q * W, a);
                    Vec<TS>::stream(row1 + (size_t)q * W, b); // row 0 again when r1 is past the end
#pragma unroll
                    for (int u = 0; u < W; u++) {
                        const TC zz = s_z[q * W + u];
                        acc0 += cvt<TC>(a[u]) * zz;
                        acc1 += cvt<TC>(b[u]) * zz;
                    }
                }
            }
            double v0 = (double)acc0, v1 = (double)acc1;
            for (int o = lpr >> 1; o > 0; o >>= 1) {
                v0 += __shfl_xor(v0, o, 64);
                v1 += __shfl_xor(v1, o, 64);
            }
            if (sl == 0 && ok0)
                s_acc[r0] += v0;
            if (sl == 0 && ok1)
                s_acc[r1] += v1;
        }
        return;
    }
    for (int rbase = w * rpw; rbase < nrows; rbase += NWAVE * rpw) {
        const int r = rbase + sub;
        TC part[MU];
#pragma unroll
        for (int s = 0; s < MU; s++)
            part[s] = 0;
        if (r < nrows) {
            const TS *row = base + (size_t)r * ld;
#pragma unroll 4
            for (int q = sl; q < nvec; q += lpr) {
                TS a[W];
                Vec<TS>::stream(row + (size_t)q * W, a);
#pragma unroll
                for (int u = 0; u < W; u++) {
                    const TC *zz = s_z + (q * W + u) * MU;
#pragma unroll
                    for (int s = 0; s < MU; s++)
                        part[s] += cvt<TC>(a[u]) * zz[s];
                }
            }
        }
        double v[MU];
#pragma unroll
        for (int s = 0; s < MU; s++)
            v[s] = (double)part[s];
        for (int o = lpr >> 1; o > 0; o >>= 1) {
#pragma unroll
            for (int s = 0; s < MU; s++)
                v[s] += __shfl_xor(v[s], o, 64);
        }
        if (sl == 0 && r < nrows) {
#pragma unroll
            for (int s = 0; s < MU; s++)
                s_acc[r * MU + s] += v[s];
        }
    }
}

constexpr int ZCAP   = 2048; // vector entries staged in LDS per pass (a multiple of 4 * MU)
constexpr int ROWMAX = 128;  // rows of a row block / V rows per pass

// Product phase 1: t = V x for every low-rank leaf. Each wavefront walks its
// own list of V chunks (no workgroup barrier anywhere). V rows are stored
// aligned to the global column index, so a lane reads 16 bytes of V and the
// matching 16 bytes of x from the L2-resident z with the same vector index.
// Rows are taken by lane groups (lanes_per_row), reduced in FP64 with
// shuffles; the group's first lane writes the row's t entry. A leaf cut into
// several chunks writes one partial sum per chunk; k_vx_sum adds them in
// chunk order.
template <typename TS, typename TC, int MU>
__device__ __forceinline__ void vx_store(const VTile &d, int r, const double (&v0)[MU], TC *z, double *part,
                                         const double *tscale) {
    double v[MU];
    const double sc = (tscale && d.nparts == 1) ? tscale[d.t_index + r] : 1.0;
#pragma unroll
    for (int s = 0; s < MU; s++)
        v[s] = v0[s] * sc;
    if (d.nparts > 1) {
        double *o = part + (size_t)(d.part_off + r) * MU;
#pragma unroll
        for (int s = 0; s < MU; s++)
            o[s] = v[s];
    } else {
        TC *o = z + (size_t)(d.t_index + r) * MU;
#pragma unroll
        for (int s = 0; s < MU; s++)
            o[s] = (TC)v[s];
    }
}

// One V chunk, one right-hand side. Rows go in blocks of RB (8 to 32): lane
// group g of LPR lanes (LPR = 64 / RPW) takes rows g, g + RPW, ... (NI = RB /
// RPW <= 8 rows, which bounds the registers) and lane l of the group the vectors l,
// l + LPR, ... of every row, so the lane loads its x entries once per block
// and reuses them for all NI rows. The NI partial sums per lane are reduced by
// recursive halving: at each step a lane keeps half of its rows and adds the
// partner lane's copy of that half (NI - 1 shuffles), then SH = LPR / NI lanes
// hold partial sums of the same row and finish with log2(SH) shuffles; lane l
// ends with row g + RPW (l / SH).
template <typename TS, typename TC, int LPR>
__device__ __forceinline__ void vx_chunk_rows(const VTile &d, const TS *varena, TC *z, double *part,
                                              const double *tscale, int lane) {
    constexpr int W   = Vec<TS>::W;
    constexpr int RPW = 64 / LPR;
    constexpr int NI  = LPR == 64 ? 8 : (RPW >= 16 ? 1 : 16 / RPW); // at most 8 rows in flight per lane
    constexpr int RB  = NI * RPW;
    constexpr int SH  = LPR / NI;
    const int sl = lane & (LPR - 1), sub = lane / LPR;
    const int nvec = d.ncols / W;
    const TC *xv   = z + d.x_off;
    const TS *vb   = varena + d.v_off;
#pragma unroll 1
    for (int rb0 = 0; rb0 < d.k; rb0 += RB) {
        const int rfin  = rb0 + sub + RPW * (sl / SH);
        const bool wr   = sl % SH == 0 && rfin < d.k;
        const double sc = (tscale && d.nparts == 1 && wr) ? tscale[d.t_index + rfin] : 1.0;
        TC acc[NI];
#pragma unroll
        for (int i = 0; i < NI; i++)
            acc[i] = 0;
#pragma unroll 1
        for (int q0 = 0; q0 < nvec; q0 += LPR) {
            const int q    = q0 + sl;
            const bool okq = q < nvec;
            TC xq[W];
            load_n<TC, W>(xv + (size_t)(okq ? q : 0) * W, xq);
#pragma unroll
            for (int u = 0; u < W; u++)
                xq[u] = okq ? xq[u] : (TC)0;
            TS a[NI][W];
#pragma unroll
            for (int i = 0; i < NI; i++) {
                const int r = rb0 + sub + RPW * i;
                Vec<TS>::stream(vb + (size_t)(r < d.k ? r : 0) * d.ld + (size_t)(okq ? q : 0) * W, a[i]);
            }
#pragma unroll
            for (int i = 0; i < NI; i++) {
                const bool ok = rb0 + sub + RPW * i < d.k;
#pragma unroll
                for (int u = 0; u < W; u++)
                    acc[i] += ok ? cvt<TC>(a[i][u]) * xq[u] : (TC)0;
            }
        }
#pragma unroll
        for (int h = NI / 2, o = LPR / 2; h >= 1; h >>= 1, o >>= 1) {
            const bool up = (sl & o) != 0;
#pragma unroll
            for (int j = 0; j < h; j++) {
                const TC send = up ? acc[j] : acc[j + h];
                const TC keep = up ? acc[j + h] : acc[j];
                acc[j]        = keep + __shfl_xor(send, o, 64);
            }
        }
#pragma unroll
        for (int o = SH / 2; o >= 1; o >>= 1)
            acc[0] += __shfl_xor(acc[0], o, 64);
        if (wr) {
            const double v = (double)acc[0] * sc;
            if (d.nparts > 1)
                part[(size_t)(d.part_off + rfin)] = v;
            else
                z[(size_t)(d.t_index + rfin)] = (TC)v;
        }
    }
}


// One V chunk, one right-hand side, FP32 arithmetic, rows of at most 256
// vectors. Every lane keeps 8 16-byte loads in flight per step, as R rows of Q
// vectors (Q = 1, 2 or 4 vectors of each row per lane, R = 8 / Q rows per lane
// group), and holds its Q x vectors in registers for the whole chunk. Lane
// group g (of lpr lanes, rpw groups per wave) takes rows g, g + rpw, ... of the
// step, so each load instruction reads rpw consecutive rows. The R row sums
// of a lane are reduced in FP64 by recursive halving (R - 1 shuffles, then
// log2(lpr / R) more); lane sl ends with row g + rpw (sl / (lpr / R)).
template <typename TS, int Q>
__device__ __forceinline__ void vx_chunk_q(const VTile &d, const TS *varena, float *z, double *part,
                                           const double *tscale, int nvec, int lpr, int lane) {
    constexpr int W = Raw<TS>::W;
    constexpr int R = 8 / Q;
    const int rpw = 64 / lpr, SH = lpr / R;
    const int sl = lane & (lpr - 1), sub = lane / lpr;
    const float *xv = z + d.x_off;
    const TS *vb    = varena + d.v_off;
    float xq[Q][W];
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const int q = sl + i * lpr;
        if (q < nvec) {
            load_n<float, W>(xv + (size_t)q * W, xq[i]);
        } else {
#pragma unroll
            for (int u = 0; u < W; u++)
                xq[i][u] = 0.f;
        }
    }
    const nt_u4 zero4 = {0u, 0u, 0u, 0u};
    for (int rbase = 0; rbase < d.k; rbase += R * rpw) {
        const int rfin   = rbase + sub + rpw * (sl / SH);
        const bool wr    = sl % SH == 0 && rfin < d.k;
        const double sc  = (tscale && d.nparts == 1 && wr) ? tscale[d.t_index + rfin] : 1.0; // in flight early
        nt_u4 ra[R][Q];
#pragma unroll
        for (int j = 0; j < R; j++) {
            const int r   = rbase + sub + rpw * j;
            const TS *row = vb + (size_t)(r < d.k ? r : 0) * d.ld;
#pragma unroll
            for (int i = 0; i < Q; i++) {
                const int q = sl + i * lpr;
                ra[j][i]    = zero4;
                if (q < nvec && r < d.k)
                    ra[j][i] = nt_load16(row + (size_t)q * W);
            }
        }
        double acc[R];
#pragma unroll
        for (int j = 0; j < R; j++) {
            float a = 0.f;
#pragma unroll
            for (int i = 0; i < Q; i++) {
                float f[W];
                Raw<TS>::cvt(ra[j][i], f);
#pragma unroll
                for (int u = 0; u < W; u++)
                    a += f[u] * xq[i][u];
            }
            acc[j] = (double)a;
        }
#pragma unroll
        for (int h = R / 2, o = lpr / 2; h >= 1; h >>= 1, o >>= 1) {
            const bool up = (sl & o) != 0;
#pragma unroll
            for (int j = 0; j < h; j++) {
                const double send = up ? acc[j] : acc[j + h];
                const double keep = up ? acc[j + h] : acc[j];
                acc[j]            = keep + shfl_pair(send, o);
            }
        }
        for (int o = SH / 2; o >= 1; o >>= 1)
            acc[0] += shfl_pair(acc[0], o);
        if (wr) {
            if (d.nparts > 1)
                part[(size_t)(d.part_off + rfin)] = acc[0];
            else
                z[(size_t)(d.t_index + rfin)] = (float)(acc[0] * sc);
        }
    }
}

// One V chunk: rows taken by lane groups (lanes_per_row), FP64 shuffle reduction.
template <typename TS, typename TC, int MU>
__device__ __forceinline__ void vx_chunk(const VTile &d, const TS *varena, TC *z, double *part, const double *tscale,
                                         int lane) {
    constexpr int W = Vec<TS>::W;
    const int nvec  = d.ncols / W;
    const int lpr   = lanes_per_row(nvec);
    const int rpw   = 64 / lpr;
    const int sl = lane & (lpr - 1), sub = lane / lpr;
    const TC *xv = z + (size_t)d.x_off * MU;
    if constexpr (MU == 1 && std::is_same<TS, double>::value) {
        // FP64 rows: x held in registers across up to 8 rows (faster here;
        // for 32- and 16-bit rows the two-row form below is faster)
        if (nvec > 32)
            vx_chunk_rows<TS, TC, 64>(d, varena, z, part, tscale, lane);
        else if (nvec > 16)
            vx_chunk_rows<TS, TC, 32>(d, varena, z, part, tscale, lane);
        else if (nvec > 8)
            vx_chunk_rows<TS, TC, 16>(d, varena, z, part, tscale, lane);
        else if (nvec > 4)
            vx_chunk_rows<TS, TC, 8>(d, varena, z, part, tscale, lane);
        else if (nvec > 2)
            vx_chunk_rows<TS, TC, 4>(d, varena, z, part, tscale, lane);
        else
            vx_chunk_rows<TS, TC, 2>(d, varena, z, part, tscale, lane);
    } else if constexpr (MU == 1 && std::is_same<TC, float>::value) {
        // rows of at most 16 vectors: eight lanes a row, two vectors each;
        // up to 256 vectors: four vectors a lane (x held in registers)
        if (nvec <= 16) {
            vx_chunk_q<TS, 2>(d, varena, z, part, tscale, nvec, 8, lane);
            return;
        }
        if (nvec <= 4 * lpr) {
            vx_chunk_q<TS, 4>(d, varena, z, part, tscale, nvec, lpr, lane);
            return;
        }
        // longer rows: two rows per lane group, x reloaded per row pair
        for (int rbase = 0; rbase < d.k; rbase += 2 * rpw) {
            const int r0 = rbase + sub, r1 = r0 + rpw;
            const bool ok0 = r0 < d.k, ok1 = r1 < d.k;
            const TS *row0 = varena + d.v_off + (size_t)(ok0 ? r0 : 0) * d.ld;
            const TS *row1 = varena + d.v_off + (size_t)(ok1 ? r1 : 0) * d.ld;
            TC acc0 = 0, acc1 = 0;
            if (ok0) {
#pragma unroll 2
                for (int q = sl; q < nvec; q += lpr) {
                    TS a[W], b[W];
                    TC xq[W];
                    load_n<TC, W>(xv + (size_t)q * W, xq);
                    Vec<TS>::stream(row0 + (size_t)q * W, a);
                    Vec<TS>::stream(row1 + (size_t)q * W, b); // row 0 again when r1 is past the end
#pragma unroll
                    for (int u = 0; u < W; u++) {
                        acc0 += cvt<TC>(a[u]) * xq[u];
                        acc1 += cvt<TC>(b[u]) * xq[u];
                    }
                }
            }
            double v0[1] = {(double)acc0}, v1[1] = {(double)acc1};
            for (int o = lpr >> 1; o > 0; o >>= 1) {
                v0[0] += __shfl_xor(v0[0], o, 64);
                v1[0] += __shfl_xor(v1[0], o, 64);
            }
            if (sl == 0 && ok0)
                vx_store<TS, TC, 1>(d, r0, v0, z, part, tscale);
            if (sl == 0 && ok1)
                vx_store<TS, TC, 1>(d, r1, v1, z, part, tscale);
        }
    } else {
        for (int rbase = 0; rbase < d.k; rbase += rpw) {
            const int r = rbase + sub;
            TC acc[MU];
#pragma unroll
            for (int s = 0; s < MU; s++)
                acc[s] = 0;
            if (r < d.k) {
                const TS *row = varena + d.v_off + (size_t)r * d.ld;
#pragma unroll 2
                for (int q = sl; q < nvec; q += lpr) {
                    TS a[W];
                    Vec<TS>::stream(row + (size_t)q * W, a);
#pragma unroll
                    for (int u = 0; u < W; u++) {
                        const TC *xj = xv + (size_t)(q * W + u) * MU;
#pragma unroll
                        for (int s = 0; s < MU; s++)
                            acc[s] += cvt<TC>(a[u]) * xj[s];
                    }
                }
            }
            double v[MU];
#pragma unroll
            for (int s = 0; s < MU; s++)
                v[s] = (double)acc[s];
            for (int o = lpr >> 1; o > 0; o >>= 1) {
#pragma unroll
                for (int s = 0; s < MU; s++)
                    v[s] += __shfl_xor(v[s], o, 64);
            }
            if (sl == 0 && r < d.k)
                vx_store<TS, TC, MU>(d, r, v, z, part, tscale);
        }
    }
}

// Product phase 1: t = V x for every low-rank leaf. Each wavefront walks its
// own list of V chunks (no workgroup barrier anywhere). V rows are stored
// aligned to the global column index, so a lane reads 16 bytes of V and the
// matching x entries from the L2-resident z with the same vector index. A
// leaf cut into several chunks writes one partial sum per chunk; k_vx_sum adds
// them in chunk order.
template <typename TS0, typename TS1, typename TC, int MU>
__global__ void __launch_bounds__(WG) k_vx(const int *wv_off, int nwaves, const VTile *tiles, const TS0 *varena0,
                                           const TS1 *varena1, TC *z, double *part, const double *tscale) {
    // wave-uniform: the descriptors of the wave's chunks go to scalar registers
    const int wv = __builtin_amdgcn_readfirstlane(blockIdx.x * NWAVE + (threadIdx.x >> 6));
    if (wv >= nwaves)
        return;
    const int lane = threadIdx.x & 63;
    const int tend = wv_off[wv + 1];
    int ti         = wv_off[wv];
    VTile d        = tiles[ti];
    for (; ti < tend; ti++) {
        const VTile next = tiles[ti + 1 < tend ? ti + 1 : ti]; // descriptor of the next chunk, loaded early
        if (d.cls == 0)
            vx_chunk<TS0, TC, MU>(d, varena0, z, part, tscale, lane);
        else
            vx_chunk<TS1, TC, MU>(d, varena1, z, part, tscale, lane);
        d = next;
    }
}

// Adds the chunk partial sums of each multi-chunk row group into its t
// entries, in chunk order: one wavefront per group for one RHS (at most a few
// dozen sums), one workgroup per group otherwise.
template <int MU>
constexpr int vx_sum_groups_per_block() {
    return MU == 1 ? NWAVE : 1;
}

// In-order sum of n partial values (E doubles each, stride apart), with up to
// eight loads in flight: the sum's latency is not one load per chunk.
template <int E>
__device__ __forceinline__ void sum_parts(const double *p, size_t stride, int n, double (&v)[E]) {
    constexpr int U = 8;
    int c = 0;
    for (; c + U <= n; c += U) {
        double w[U][E];
#pragma unroll
        for (int u = 0; u < U; u++) {
            if constexpr (E == 1)
                w[u][0] = p[(size_t)(c + u) * stride];
            else
                Vec<double>::load(p + (size_t)(c + u) * stride, w[u]);
        }
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int i = 0; i < E; i++)
                v[i] += w[u][i];
    }
    for (; c < n; c++)
#pragma unroll
        for (int i = 0; i < E; i++)
            v[i] += p[(size_t)c * stride + i];
}

template <typename TC, int MU, bool BLK = false>
__global__ void __launch_bounds__(WG) k_vx_sum(const VTile *first, int nmulti,
                                               const double *part, TC *z, const double *tscale,
                                               unsigned *tmax = nullptr) {
    constexpr int GPB = vx_sum_groups_per_block<MU>();
    constexpr int TPG = WG / GPB; // threads per group
    const int gi      = blockIdx.x * GPB + (int)threadIdx.x / TPG;
    if (gi >= nmulti)
        return;
    const VTile d    = first[gi];
    const int km     = d.k * MU;
    const double *pb = part + (size_t)d.part_base * MU;
    // gridDim.y workgroups share one group's entries
    float m = 0.0f;
    for (int e = threadIdx.x % TPG + TPG * blockIdx.y; e < km; e += TPG * gridDim.y) {
        double v[1] = {0};
        sum_parts<1>(pb + e, km, d.nparts, v);
        const int r = e / MU, s = e % MU;
        const TC o  = (TC)(tscale ? v[0] * tscale[d.t_index + r] : v[0]);
        z[BLK ? zblk(d.t_index + r, s) : (size_t)(d.t_index + r) * MU + s] = o;
        m = fmaxf(m, fabsf((float)o));
    }
    if (tmax && d.cls == 1) { // one group per workgroup here (MU = 32)
        m = wave_max(m);
        if ((threadIdx.x & 63) == 0)
            atomicMax(tmax + ((blockIdx.x + blockIdx.y) & (TMAX_SLOTS - 1)), __float_as_uint(m));
    }
}

// Product phase 2: one workgroup per row block (<= 128 rows). In each
// precision class the row block's dense-leaf rows and U rows form one
// row-major R x C block (row stride C rounded up to 4, zero padded); the vector
// entry of each column (an x entry for a dense column, a t entry for a U
// column) is staged through LDS. Both classes add into the same FP64 row
// sums, which are scattered to original order.
template <typename TS, typename TC, int MU>
__device__ __forceinline__ void rows_class(const TS *arena, const RowClass &rc, int rb, int lo, int R, const TC *z,
                                           TC *s_z, double *s_acc) {
    constexpr int ZCOLS = ZCAP / MU;
    constexpr int W     = Vec<TS>::W;
    const int *colidx   = rc.colidx;
    const int c0 = rc.rb_col0[rb], C = rc.rb_col0[rb + 1] - c0;
    const int Cp = (C + rc.align - 1) & ~(rc.align - 1);    // columns read (vector width)
    const int Cs = (C + rc.salign - 1) & ~(rc.salign - 1);  // row stride (128-byte rows)
    const uint64_t *rb_aoff = rc.rb_aoff;
    const TS *A  = arena + rb_aoff[rb] + (size_t)lo * Cs;
    constexpr int ZPT = ZCAP / WG; // staged entries per thread
    for (int cc = 0; cc < Cp; cc += ZCOLS) {
        const int nz = min(ZCOLS, Cp - cc);
        auto stage = [&]() {
            if constexpr (MU == 1 && sizeof(TC) == 4) {
                // all column indices of the pass, then all vector entries, in
                // flight together: two memory latencies per pass (FP32 vectors;
                // for FP64 the extra registers cost more occupancy than this saves)
                int idx[ZPT];
#pragma unroll
                for (int i = 0; i < ZPT; i++) {
                    const int c = threadIdx.x + i * WG;
                    idx[i]      = c < nz && cc + c < C ? colidx[c0 + cc + c] : -1;
                }
                TC v[ZPT];
#pragma unroll
                for (int i = 0; i < ZPT; i++)
                    v[i] = idx[i] >= 0 ? z[idx[i]] : (TC)0;
#pragma unroll
                for (int i = 0; i < ZPT; i++) {
                    const int c = threadIdx.x + i * WG;
                    if (c < nz)
                        s_z[c] = v[i];
                }
            } else {
                for (int e = threadIdx.x; e < nz * MU; e += WG) {
                    const int c = cc + e / MU;
                    s_z[e]      = c < C ? z[(size_t)colidx[c0 + c] * MU + e % MU] : (TC)0;
                }
            }
            __syncthreads();
        };
        rows_gemv<TS, TC, MU>(A + cc, (size_t)Cs, R, nz / W, s_z, s_acc, stage);
        __syncthreads();
    }
}

template <typename TS0, typename TS1, typename TC, int MU>
__global__ void __launch_bounds__(WG) k_rows(const int *rb_order, const int *rb_row0, RowClass c0, const TS0 *arena0,
                                             RowClass c1, const TS1 *arena1, const TC *z, const int *perm_t, int nr,
                                             int mu_base, double *y) {
    __shared__ TC s_z[ZCAP];
    __shared__ double s_acc[ROWMAX * MU];
    // entry: row block (bits 0-23); a nonzero part count np (bits 28-31)
    // selects rows [R p / np, R (p + 1) / np) of it, p in bits 24-27
    const int ent = rb_order[blockIdx.x];
    if ((ent & 0xffffff) == 0xffffff) // padding entry of an XCD-ordered list
        return;
    const int rb = ent & 0xffffff, p = (ent >> 24) & 15, np = (int)((unsigned)ent >> 28);
    const int rf0 = rb_row0[rb], Rf = rb_row0[rb + 1] - rf0;
    const int lo = np ? Rf * p / np : 0, hi = np ? Rf * (p + 1) / np : Rf;
    const int r0 = rf0 + lo, R = hi - lo;
    for (int e = threadIdx.x; e < R * MU; e += WG)
        s_acc[e] = 0;
    rows_class<TS0, TC, MU>(arena0, c0, rb, lo, R, z, s_z, s_acc);
    if (arena1)
        rows_class<TS1, TC, MU>(arena1, c1, rb, lo, R, z, s_z, s_acc);
    __syncthreads();
    for (int e = threadIdx.x; e < R * MU; e += WG) {
        const int r = e / MU, s = e % MU;
        y[(size_t)perm_t[r0 + r] + (size_t)nr * (mu_base + s)] = s_acc[e];
    }
}

// ---------------------------------------------------------------------------
// 32-RHS products on the matrix cores (FP32 MFMA, wave64). The reduction index
// of an MFMA is a column of the stored block: lane l feeds one row and one
// column of the A side and the same column for one right-hand side on the B
// side. Each lane reads VW consecutive columns of its row with one 16-byte
// load; every load of a pass is issued before the pass's first MFMA, so a wave
// waits for memory once per pass. Out-of-range rows and columns read a valid
// address and are masked to zero (columns) or never written (rows). Products
// accumulate in FP32 within a pass, then in FP64.
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int MFMA_MU    = 32;  // right-hand sides of one matrix-core pass
constexpr int VPASS16 = 64; // V columns per pass, 16-row tiles (complex passes)
constexpr int VPASS32 = 32; // V columns per pass, 32-row tiles (complex passes)

// Output row of accumulator i of lane l in a 32x32 tile (v_mfma_f32_32x32x2f32).
__device__ __forceinline__ int mfma32_row(int i, int lane) { return (i & 3) + 4 * (lane >> 5) + 8 * (i >> 2); }

// 16-bit matrix-core products for 16-bit stored entries: v_mfma_f32_32x32x16
// and 16x16x32 (f16 or bf16), eight entries of one row per lane, the lane's
// 16-byte load as it is. The FP32 vector entries are split into 16-bit terms
// whose sum keeps their precision: FP16 entries times a power-of-two pass
// scale s = h1 + h2 (s brings the pass's largest magnitude into [2^14, 2^15),
// inside FP16's range), BF16 entries = b1 + b2 + b3 (BF16 has FP32's range, no
// scale). Each term is one MFMA into the same FP32 accumulators; a pass's sums
// are multiplied by 1 / s when added into the FP64 totals. Output layouts are
// those of the FP32 forms (32x32: mfma32_row; 16x16: row 4 (l / 16) + i). The
// eight k slots of a lane may be any eight columns as long as A and B agree.
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
template <typename TS>
struct Mma16 {
    static constexpr bool on = false, scaled = false;
};
template <>
struct Mma16<__half> {
    static constexpr bool on = true, scaled = true;
    static constexpr int NB = 2;
    using V                 = f16x8;
    using E                 = _Float16;
    __device__ static _Float16 el(__half x, bool in) { return in ? __builtin_bit_cast(_Float16, x) : (_Float16)0.0f; }
    __device__ static V a(const __half (&x)[8], bool in) {
        V v;
#pragma unroll
        for (int j = 0; j < 8; j++)
            v[j] = in ? __builtin_bit_cast(_Float16, x[j]) : (_Float16)0.0f;
        return v;
    }
    __device__ static void split1(float b, float sc, E (&o)[NB]) {
        const float t    = b * sc;
        const _Float16 h = (_Float16)t;
        o[0]             = h;
        o[1]             = (_Float16)(t - (float)h);
    }
    __device__ static void split(const float (&b)[8], float sc, V (&o)[NB]) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            E e[NB];
            split1(b[j], sc, e);
            o[0][j] = e[0];
            o[1][j] = e[1];
        }
    }
    __device__ static floatx16 mma32(V a, V b, floatx16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }
    __device__ static floatx4 mma16(V a, V b, floatx4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};
template <>
struct Mma16<__hip_bfloat16> {
    static constexpr bool on = true, scaled = false;
    static constexpr int NB = 3;
    using V                 = bf16x8;
    using E                 = __bf16;
    __device__ static __bf16 el(__hip_bfloat16 x, bool in) { return in ? __builtin_bit_cast(__bf16, x) : (__bf16)0.0f; }
    __device__ static V a(const __hip_bfloat16 (&x)[8], bool in) {
        V v;
#pragma unroll
        for (int j = 0; j < 8; j++)
            v[j] = in ? __builtin_bit_cast(__bf16, x[j]) : (__bf16)0.0f;
        return v;
    }
    __device__ static void split1(float t, float, E (&o)[NB]) {
        const __bf16 h = (__bf16)t;
        const float r  = t - (float)h;
        const __bf16 m = (__bf16)r;
        o[0]           = h;
        o[1]           = m;
        o[2]           = (__bf16)(r - (float)m);
    }
    __device__ static void split(const float (&b)[8], float, V (&o)[NB]) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            E e[NB];
            split1(b[j], 1.0f, e);
            o[0][j] = e[0];
            o[1][j] = e[1];
            o[2][j] = e[2];
        }
    }
    __device__ static floatx16 mma32(V a, V b, floatx16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
    __device__ static floatx4 mma16(V a, V b, floatx4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};

// Real and imaginary parts of two 16-byte loads of complex 16-bit entries
// (four (re, im) pairs each) as two 8-entry operands; entries of a load that
// lies past the pass's columns are zero when its flag is false.
template <typename TS>
__device__ __forceinline__ void deinterleave16(const TS (&p)[8], const TS (&q)[8], bool inp, bool inq,
                                               typename Mma16<TS>::V &re, typename Mma16<TS>::V &im) {
    using M = Mma16<TS>;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        re[j]     = M::el(p[2 * j], inp);
        im[j]     = M::el(p[2 * j + 1], inp);
        re[j + 4] = M::el(q[2 * j], inq);
        im[j + 4] = M::el(q[2 * j + 1], inq);
    }
}

// Power-of-two pass scale for a largest magnitude m (1 when m is zero or not finite: the products then carry the
// infinity or NaN as FP32 ones would).
__device__ __forceinline__ float pass_scale(float m) {
    return m > 0.0f && m <= 3.0e38f ? ldexpf(1.0f, 14 - ilogbf(m)) : 1.0f;
}

typedef double doublex4 __attribute__((ext_vector_type(4))); // v_mfma_f64_16x16x4f64 accumulators

// ---------------------------------------------------------------------------
// 32-RHS V phase by row ranges. A work item is at most RR_ROWS (128) stacked V
// rows of one panel in one precision class over a range of its columns; one
// workgroup each, wave w owning the 16-row tiles w and w + 4 of the range. The
// workgroup walks the range in chunks of RR_CHUNK columns (RR_CHUNK64 in FP64):
// each chunk's x entries (all 32 right-hand sides) are staged in LDS, double
// buffered, the next chunk read into registers while the current one is used.
// A panel stacks the V rows of every low-rank leaf of its source cluster (107
// on average at N = 1e5), so each x entry is fetched once per 128 rows rather
// than once per 16. Each wave streams its tiles' rows in units of four column
// steps on 16x16 matrix-core tiles, the next unit's entries loaded while the
// current one runs (register ping-pong; every load is issued unconditionally,
// so the wait before a unit covers that unit's loads only, and the one barrier
// per chunk waits on LDS alone). The row sums stay in registers over the range
// and are stored once at the end: the t entries when the range spans the
// panel's columns, else partial sums that k_vx_sum adds.
constexpr int RR_ROWS    = RR_ITEM_ROWS;    // rows per work item (one 16-row tile per wave)
static_assert((TMAX_SLOTS & (TMAX_SLOTS - 1)) == 0, "slots of the largest class-1 |t|: a power of two");
static_assert(TMAX_SLOTS % 64 == 0, "the row kernels read the slots 64 to a load");
constexpr int RR_CHUNK   = RR_ITEM_CHUNK32; // columns per staged chunk, FP32 arithmetic
constexpr int RR_CHUNK64 = RR_ITEM_CHUNK64; // columns per staged chunk, FP64
static_assert(RR_ROWS == NWAVE * 16, "one 16-row tile per wave");

// LDS bytes of one staged chunk in the form class storage TS multiplies: FP32
// (blocked by 4 columns) or the 16-bit terms of the split (blocked by 8).
template <typename TS>
constexpr int rr_stage_bytes() {
    if constexpr (Mma16<TS>::on)
        return Mma16<TS>::NB * RR_CHUNK * MFMA_MU * 2;
    else
        return RR_CHUNK * MFMA_MU * 4;
}

// One wave's walk over a work item: chunk 0 staged, then units u = 0 .. U - 1
// (UPC per chunk), unit u's matrix entries in one register buffer while unit
// u - 1 is consumed from the other (every load issued unconditionally, so the wait before a unit
// covers that unit's loads only). At the first unit of chunk j the next
// chunk's x is read into registers (xload), after its last unit it is stored
// to the other LDS buffer (xstore) and the workgroup meets at the chunk's one
// barrier (which waits on LDS, not on the loads in flight).
template <int UPC, typename Buf, typename Load, typename Run, typename XLoad, typename XStore>
__device__ __forceinline__ void rr_walk(int nchunk, Load load, Run run, XLoad xload, XStore xstore) {
    const int U = nchunk * UPC;
    Buf A, B;
    // the first unit's entries are requested before chunk 0 is staged: the two
    // latencies overlap
    load(0, A);
    xload(0);
    xstore(0);
    __syncthreads();
    auto half = [&](int u, const Buf &cur, Buf &nxt) {
        const int j = u / UPC;
        if (u % UPC == 0)
            xload(min(j + 1, nchunk - 1));
        load(min(u + 1, U - 1), nxt);
        run(u, cur);
        if (u % UPC == UPC - 1) {
            if (j + 1 < nchunk)
                xstore(j + 1);
            __syncthreads();
        }
    };
    for (int u = 0; u < U; u += 2) {
        half(u, A, B);
        if (u + 1 >= U)
            break;
        half(u + 1, B, A);
    }
}

// rr_walk for one unit per chunk with every load two units ahead: while unit
// u runs, the matrix entries of units u + 1 and u + 2 and the x of chunks
// u + 1 (in registers until its store to LDS after unit u) and u + 2 are in
// flight, so each load's latency is covered by two units of work. Three
// register buffers for the entries, two for x (xload / xstore take the x
// buffer); the loop is unrolled to six units so that every buffer is a fixed
// set of registers.
template <typename Buf, typename XS, typename Load, typename Run, typename XLoad, typename XStore>
__device__ __forceinline__ void rr_walk2(int nchunk, Load load, Run run, XLoad xload, XStore xstore) {
    const int U = nchunk;
    Buf A0, A1, A2;
    XS X0, X1;
    load(0, A0);
    xload(0, X0);
    load(min(1, U - 1), A1);
    xload(min(1, U - 1), X1);
    xstore(0, X0);
    __syncthreads();
    auto step = [&](int u, const Buf &cur, Buf &nxt2, XS &xl, const XS &xs) {
        xload(min(u + 2, U - 1), xl);
        load(min(u + 2, U - 1), nxt2);
        run(u, cur);
        if (u + 1 < U)
            xstore(u + 1, xs);
        __syncthreads();
    };
    for (int u = 0; u < U; u += 6) {
        step(u, A0, A2, X0, X1);
        if (u + 1 >= U)
            break;
        step(u + 1, A1, A0, X1, X0);
        if (u + 2 >= U)
            break;
        step(u + 2, A2, A1, X0, X1);
        if (u + 3 >= U)
            break;
        step(u + 3, A0, A2, X1, X0);
        if (u + 4 >= U)
            break;
        step(u + 4, A1, A0, X0, X1);
        if (u + 5 >= U)
            break;
        step(u + 5, A2, A1, X1, X0);
    }
}

// Rows [0, d.k) of a work item (<= 64: wave w owns the 16-row tile w), class
// storage TS, FP32 arithmetic. Steps are 16 columns for FP32 entries
// (v_mfma_f32_16x16x4f32: lane (r, g) = (l % 16, l / 16) loads row r's columns
// 4 g .. 4 g + 3 of the step and feeds them to four MFMAs, k = g in each, B =
// x of the same column from LDS) and 32 columns for 16-bit entries
// (v_mfma_f32_16x16x32: lane (r, g) loads columns 8 g .. 8 g + 7 as its A
// operand, B = the staged 16-bit terms, one MFMA per term); two MFMAs per A
// operand (right-hand sides 0-15 and 16-31). Every step of a unit runs: past
// the range's columns the staged x is zero and the clamped loads finite. FP16
// terms carry a power-of-two scale per chunk from xmax (the largest |x| of
// each 32-column block of z, written by the gather). Unit sums are FP32, then
// FP64.
template <typename TS>
__device__ __forceinline__ void rr_rows32(const VTile &d, const TS *varena, const float *z, float *zil,
                                          const float *xmax, unsigned *tmax, unsigned char *lds, double *part,
                                          const double *tscale) {
    constexpr bool H16 = Mma16<TS>::on;
    constexpr int VW   = Vec<TS>::W;           // entries per 16-byte load
    constexpr int SC   = H16 ? 32 : 16;        // columns per step
    constexpr int NS   = 4;                    // steps per unit
    constexpr int UPC  = RR_CHUNK / (SC * NS); // units per chunk (2 or 1)
    constexpr int BUF  = rr_stage_bytes<TS>();
    constexpr int XPT  = RR_CHUNK * MFMA_MU / 4 / WG; // staged float4 per thread (4)
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    struct Buf {
        u4 a[NS];
    };
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 15, g = lane >> 4;
    const int K      = d.k;
    const int nstep  = (d.ncols + SC - 1) / SC;
    const int nchunk = (d.ncols + RR_CHUNK - 1) / RR_CHUNK;
    const int ngrp   = (d.ncols + 3) >> 2;       // 4-column groups of the range
    const bool real  = 16 * w < K;               // the wave's tile holds rows
    const TS *rp = varena + d.v_off + (size_t)min(16 * w + r, K - 1) * d.ld + VW * g;
    auto load = [&](int u, Buf &b) {
#pragma unroll
        for (int st = 0; st < NS; st++)
            b.a[st] = nt_load16(rp + (size_t)min(u * NS + st, nstep - 1) * SC);
    };
    const float4 *xsrc = reinterpret_cast<const float4 *>(z + (size_t)(d.x_off >> 2) * 128);
    float4 xv[XPT];
    // FP16 terms: the chunk's power-of-two scale, from the xmax of the (at most
    // NXB) 32-column blocks it touches. Those are loaded with the chunk's x and
    // the scale kept in a register from the chunk's store to its last unit: a
    // load at the store or in a unit would wait on every load issued before it
    // (loads complete in order), the next unit's matrix entries included.
    constexpr bool SCL = H16 && Mma16<TS>::scaled;
    constexpr int NXB  = RR_CHUNK / 32 + 1;
    float xm[NXB];
    double inv_cur = 1.0; // 1 / the scale of the chunk in LDS being used
    auto xload = [&](int j) {
#pragma unroll
        for (int i = 0; i < XPT; i++) {
            const int f = (int)threadIdx.x + WG * i;
            xv[i]       = xsrc[(size_t)min(j * (RR_CHUNK / 4) + (f >> 5), ngrp - 1) * MFMA_MU + (f & 31)];
        }
        if constexpr (SCL) {
            const int c0 = d.x_off + j * RR_CHUNK, c1 = min(c0 + RR_CHUNK, d.x_off + d.ncols) - 1;
#pragma unroll
            for (int k = 0; k < NXB; k++)
                xm[k] = xmax[min((c0 >> 5) + k, c1 >> 5)];
        }
    };
    auto xstore = [&](int j) {
        unsigned char *buf = lds + (j & 1) * BUF;
        float sc = 1.0f;
        if constexpr (SCL) {
            float m = 0.0f;
#pragma unroll
            for (int k = 0; k < NXB; k++)
                m = fmaxf(m, xm[k]);
            sc      = pass_scale(m);
            inv_cur = 1.0 / (double)sc;
        }
#pragma unroll
        for (int i = 0; i < XPT; i++) {
            const int f = (int)threadIdx.x + WG * i, grp = f >> 5, s = f & 31;
            const int c = j * RR_CHUNK + grp * 4; // column of the range
            float4 v    = xv[i];
            v.x = c + 0 < d.ncols ? v.x : 0.0f;
            v.y = c + 1 < d.ncols ? v.y : 0.0f;
            v.z = c + 2 < d.ncols ? v.z : 0.0f;
            v.w = c + 3 < d.ncols ? v.w : 0.0f;
            if constexpr (!H16) {
                reinterpret_cast<float4 *>(buf)[f] = v;
            } else {
                using M  = Mma16<TS>;
                using EL = typename M::E;
                const float b[4] = {v.x, v.y, v.z, v.w};
                EL t[M::NB][4];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    EL o[M::NB];
                    M::split1(b[e], sc, o);
#pragma unroll
                    for (int u = 0; u < M::NB; u++)
                        t[u][e] = o[u];
                }
#pragma unroll
                for (int u = 0; u < M::NB; u++) {
                    // [term][grp / 2][s][8]: this group is entries (grp % 2) * 4 .. + 3
                    EL *p = reinterpret_cast<EL *>(buf) +
                            ((size_t)(u * (RR_CHUNK / 8) + (grp >> 1)) * MFMA_MU + s) * 8 + (grp & 1) * 4;
                    typedef EL el4 __attribute__((ext_vector_type(4)));
                    el4 wv;
                    wv[0] = t[u][0], wv[1] = t[u][1], wv[2] = t[u][2], wv[3] = t[u][3];
                    *reinterpret_cast<el4 *>(p) = wv;
                }
            }
        }
    };
    double accd[2][4] = {}; // [RHS half][row]
    auto run = [&](int u, const Buf &b) {
        if (!real)
            return;
        const int j = u / UPC, s = u % UPC;
        const unsigned char *buf = lds + (j & 1) * BUF;
        floatx4 acc0 = {0}, acc1 = {0};
#pragma unroll
        for (int st = 0; st < NS; st++) {
            const int cs = s * NS + st; // step within the chunk
            if constexpr (H16) {
                using M = Mma16<TS>;
                const typename M::V av = __builtin_bit_cast(typename M::V, b.a[st]);
#pragma unroll
                for (int t = 0; t < M::NB; t++) {
                    const typename M::V *bp = reinterpret_cast<const typename M::V *>(buf) +
                                              (size_t)(t * (RR_CHUNK / 8) + cs * (SC / 8) + g) * MFMA_MU;
                    acc0 = M::mma16(av, bp[r], acc0);
                    acc1 = M::mma16(av, bp[r + 16], acc1);
                }
            } else {
                const float4 *bp = reinterpret_cast<const float4 *>(buf) + (size_t)(cs * 4 + g) * MFMA_MU;
                const float4 b0 = bp[r], b1 = bp[r + 16];
                const float av[4] = {__uint_as_float(b.a[st].x), __uint_as_float(b.a[st].y),
                                     __uint_as_float(b.a[st].z), __uint_as_float(b.a[st].w)};
                const float bl[4] = {b0.x, b0.y, b0.z, b0.w}, bh[4] = {b1.x, b1.y, b1.z, b1.w};
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av[m], bl[m], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(av[m], bh[m], acc1, 0, 0, 0);
                }
            }
        }
        // chunk j's store came after every unit of chunk j - 1 and before
        // this one: inv_cur is chunk j's
        const double inv = inv_cur;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            accd[0][i] += (double)acc0[i] * inv;
            accd[1][i] += (double)acc1[i] * inv;
        }
    };
    rr_walk<UPC, Buf>(nchunk, load, run, xload, xstore);
    // lane (r, g) holds rows 4 g + i of the tile, right-hand sides r and r + 16;
    // t entries go to the RHS-interleaved z the row phase gathers from
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int row = 16 * w + 4 * g + i;
        if (row < K) {
            if (d.nparts > 1) {
                part[(size_t)(d.part_off + row) * MFMA_MU + r]      = accd[0][i];
                part[(size_t)(d.part_off + row) * MFMA_MU + r + 16] = accd[1][i];
            } else {
                const double f = tscale ? tscale[d.t_index + row] : 1.0;
                const float v0 = (float)(accd[0][i] * f), v1 = (float)(accd[1][i] * f);
                zil[(size_t)(d.t_index + row) * MFMA_MU + r]      = v0;
                zil[(size_t)(d.t_index + row) * MFMA_MU + r + 16] = v1;
                m = fmaxf(m, fmaxf(fabsf(v0), fabsf(v1)));
            }
        }
    }
    if (d.cls == 1 && d.nparts == 1) { // the class-1 scale of the row phase
        m = wave_max(m);
        if (lane == 0)
            atomicMax(tmax + (blockIdx.x & (TMAX_SLOTS - 1)), __float_as_uint(m));
    }
}

template <typename TS0, typename TS1>
__global__ void __launch_bounds__(WG) k_vx_rr(const VTile *items, const TS0 *varena0, const TS1 *varena1,
                                              const float *z, float *zil, const float *xmax, unsigned *tmax,
                                              double *part, const double *tscale) {
    constexpr int B0 = rr_stage_bytes<TS0>(), B1 = rr_stage_bytes<TS1>();
    __shared__ __align__(16) unsigned char lds[2 * (B0 > B1 ? B0 : B1)];
    const VTile d = items[blockIdx.x];
    if (d.k == 0) // padding entry of an XCD-ordered list
        return;
    if (d.cls == 0)
        rr_rows32<TS0>(d, varena0, z, zil, xmax, tmax, lds, part, tscale);
    else
        rr_rows32<TS1>(d, varena1, z, zil, xmax, tmax, lds, part, tscale);
}

// FP64 work items: x staged RHS-interleaved as in z ([column][32]), chunks of
// RR_CHUNK64 columns; units of four 8-column steps (v_mfma_f64_16x16x4f64:
// lane (r, g) loads columns 2 g, 2 g + 1 of the step, MFMA m takes column
// 2 g + m as k = g; result register i of lane (r, g) holds row 4 i + g). Sums
// are FP64 throughout; the walk is the FP32 one.
__global__ void __launch_bounds__(WG) k_vx_rr64(const VTile *items, const double *varena0, const double *varena1,
                                                double *z, double *part, const double *tscale) {
    constexpr int MU = 32, NS = 4, SC = 8, UPC = RR_CHUNK64 / (SC * NS), XPT = RR_CHUNK64 * MU / 2 / WG;
    __shared__ __align__(16) double s_x[2][RR_CHUNK64 * MU];
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    struct Buf {
        u4 a[NS];
    };
    const VTile d = items[blockIdx.x];
    if (d.k == 0) // padding entry of an XCD-ordered list
        return;
    const double *varena = d.cls == 0 ? varena0 : varena1;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 15, g = lane >> 4;
    const int K      = d.k;
    const int nstep  = (d.ncols + SC - 1) / SC;
    const int nchunk = (d.ncols + RR_CHUNK64 - 1) / RR_CHUNK64;
    const bool real  = 16 * w < K;
    const double *rp = varena + d.v_off + (size_t)min(16 * w + r, K - 1) * d.ld + 2 * g;
    auto load = [&](int u, Buf &b) {
#pragma unroll
        for (int st = 0; st < NS; st++)
            b.a[st] = nt_load16(rp + (size_t)min(u * NS + st, nstep - 1) * SC);
    };
    const double2 *xsrc = reinterpret_cast<const double2 *>(z + (size_t)d.x_off * MU);
    double2 xv[XPT];
    auto xload = [&](int j) {
#pragma unroll
        for (int i = 0; i < XPT; i++) {
            const int f = (int)threadIdx.x + WG * i; // column j * RR_CHUNK64 + f / 16 of the range, RHS pair f % 16
            xv[i] = xsrc[(size_t)min(j * RR_CHUNK64 + (f >> 4), d.ncols - 1) * (MU / 2) + (f & 15)];
        }
    };
    auto xstore = [&](int j) {
        double2 *buf = reinterpret_cast<double2 *>(s_x[j & 1]);
#pragma unroll
        for (int i = 0; i < XPT; i++) {
            const int f = (int)threadIdx.x + WG * i;
            buf[f] = j * RR_CHUNK64 + (f >> 4) < d.ncols ? xv[i] : make_double2(0.0, 0.0);
        }
    };
    doublex4 acc0 = {0}, acc1 = {0};
    auto run = [&](int u, const Buf &b) {
        if (!real)
            return;
        const int j = u / UPC, s = u % UPC;
        const double *buf = s_x[j & 1];
#pragma unroll
        for (int st = 0; st < NS; st++) {
            const int cs = s * NS + st;
            const double av[2] = {__hiloint2double((int)b.a[st].y, (int)b.a[st].x),
                                  __hiloint2double((int)b.a[st].w, (int)b.a[st].z)};
#pragma unroll
            for (int m = 0; m < 2; m++) {
                const double *bp = buf + (size_t)(cs * SC + 2 * g + m) * MU;
                acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m], bp[r], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m], bp[r + 16], acc1, 0, 0, 0);
            }
        }
    };
    rr_walk<UPC, Buf>(nchunk, load, run, xload, xstore);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int row = 16 * w + 4 * i + g;
        if (row < K) {
            if (d.nparts > 1) {
                part[(size_t)(d.part_off + row) * MU + r]      = acc0[i];
                part[(size_t)(d.part_off + row) * MU + r + 16] = acc1[i];
            } else {
                const double f = tscale ? tscale[d.t_index + row] : 1.0;
                z[(size_t)(d.t_index + row) * MU + r]      = acc0[i] * f;
                z[(size_t)(d.t_index + row) * MU + r + 16] = acc1[i] * f;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// 32-RHS row phase by row-block parts of at most RR_ROWS rows (wave w owns
// rows 16 w .. 16 w + 15). Per precision class the part's columns are walked
// in chunks as in the V phase (rr_walk): the chunk's vector entries are
// gathered through the column indices from the RHS-interleaved z (16 bytes =
// four right-hand sides of one entry per load) and staged in LDS in the form
// the class multiplies, double buffered; the matrix rows stream in pipelined
// units; both classes add into the same FP64 row sums, stored once to y in
// original order.
template <typename TS>
__device__ __forceinline__ void rb_class32(const RowClass &rc, const TS *arena, int rb, int roff, int R,
                                           const float *zil, float tsc, unsigned char *lds, double (&accd)[2][4]) {
    const int cb = rc.rb_col0[rb], C = rc.rb_col0[rb + 1] - cb;
    if (C == 0)
        return;
    constexpr bool H16 = Mma16<TS>::on;
    constexpr int VW   = Vec<TS>::W;
    constexpr int SC   = H16 ? 32 : 16;
    constexpr int NS   = 4;
    constexpr int UPC  = RR_CHUNK / (SC * NS);
    constexpr int BUF  = rr_stage_bytes<TS>();
    constexpr int XPT  = RR_CHUNK * MFMA_MU / 4 / WG; // gathered float4 per thread (4)
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    struct Buf {
        u4 a[NS];
    };
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 15, g = lane >> 4;
    const int Cs     = (C + rc.salign - 1) & ~(rc.salign - 1); // row stride
    const int nstep  = (C + SC - 1) / SC;
    const int nchunk = (C + RR_CHUNK - 1) / RR_CHUNK;
    const bool real  = 16 * w < R;
    const TS *rp = arena + rc.rb_aoff[rb] + (size_t)(roff + min(16 * w + r, R - 1)) * Cs + VW * g;
    auto load = [&](int u, Buf &b) {
#pragma unroll
        for (int st = 0; st < NS; st++)
            b.a[st] = nt_load16(rp + (size_t)min(u * NS + st, nstep - 1) * SC);
    };
    const float4 *zf = reinterpret_cast<const float4 *>(zil);
    const int *ci    = rc.colidx + cb;
    float4 xv[XPT];
    // column indices one chunk ahead: the walk reads chunks 0, 1, 2, ... in
    // order (the last one twice, unused), so xload(j) finds chunk j's indices
    // in registers, loaded with the previous chunk's entries, and issues the
    // vector loads at once rather than behind a wait on the index loads
    int idx[XPT];
    auto iload = [&](int j) {
#pragma unroll
        for (int i = 0; i < XPT; i++)
            idx[i] = ci[min(j * RR_CHUNK + (((int)threadIdx.x + WG * i) >> 3), C - 1)];
    };
    iload(0);
    auto xload = [&](int j) { // float4 f: column f / 8 of the chunk, right-hand sides 4 (f % 8) ..
#pragma unroll
        for (int i = 0; i < XPT; i++)
            xv[i] = zf[(size_t)idx[i] * (MFMA_MU / 4) + (((int)threadIdx.x + WG * i) & 7)];
        iload(j + 1);
    };
    auto xstore = [&](int j) {
        unsigned char *buf = lds + (j & 1) * BUF;
#pragma unroll
        for (int i = 0; i < XPT; i++) {
            const int f = (int)threadIdx.x + WG * i, cl = f >> 3, q = f & 7;
            const bool in = j * RR_CHUNK + cl < C;
            const float b[4] = {in ? xv[i].x : 0.0f, in ? xv[i].y : 0.0f, in ? xv[i].z : 0.0f, in ? xv[i].w : 0.0f};
            if constexpr (!H16) {
                float *bf = reinterpret_cast<float *>(buf);
#pragma unroll
                for (int e = 0; e < 4; e++)
                    bf[(cl >> 2) * (4 * MFMA_MU) + (4 * q + e) * 4 + (cl & 3)] = b[e];
            } else {
                using M  = Mma16<TS>;
                using EL = typename M::E;
                EL *be   = reinterpret_cast<EL *>(buf);
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    EL o[M::NB];
                    M::split1(b[e], tsc, o);
#pragma unroll
                    for (int u = 0; u < M::NB; u++)
                        be[((size_t)(u * (RR_CHUNK / 8) + (cl >> 3)) * MFMA_MU + 4 * q + e) * 8 + (cl & 7)] = o[u];
                }
            }
        }
    };
    const double inv = 1.0 / (double)tsc;
    auto run = [&](int u, const Buf &b) {
        if (!real)
            return;
        const int j = u / UPC, s = u % UPC;
        const unsigned char *buf = lds + (j & 1) * BUF;
        floatx4 acc0 = {0}, acc1 = {0};
#pragma unroll
        for (int st = 0; st < NS; st++) {
            const int cs = s * NS + st;
            if constexpr (H16) {
                using M = Mma16<TS>;
                const typename M::V av = __builtin_bit_cast(typename M::V, b.a[st]);
#pragma unroll
                for (int t = 0; t < M::NB; t++) {
                    const typename M::V *bp = reinterpret_cast<const typename M::V *>(buf) +
                                              (size_t)(t * (RR_CHUNK / 8) + cs * (SC / 8) + g) * MFMA_MU;
                    acc0 = M::mma16(av, bp[r], acc0);
                    acc1 = M::mma16(av, bp[r + 16], acc1);
                }
            } else {
                const float4 *bp = reinterpret_cast<const float4 *>(buf) + (size_t)(cs * 4 + g) * MFMA_MU;
                const float4 b0 = bp[r], b1 = bp[r + 16];
                const float av[4] = {__uint_as_float(b.a[st].x), __uint_as_float(b.a[st].y),
                                     __uint_as_float(b.a[st].z), __uint_as_float(b.a[st].w)};
                const float bl[4] = {b0.x, b0.y, b0.z, b0.w}, bh[4] = {b1.x, b1.y, b1.z, b1.w};
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av[m], bl[m], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(av[m], bh[m], acc1, 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            accd[0][i] += (double)acc0[i] * inv;
            accd[1][i] += (double)acc1[i] * inv;
        }
    };
    rr_walk<UPC, Buf>(nchunk, load, run, xload, xstore);
}

// items[b] = row block | part << 24; the class-1 (16-bit FP16) vector entries
// carry one power-of-two scale from tmax (the largest class-1 |t|, kept by the
// V phase).
template <typename TS0, typename TS1>
__global__ void __launch_bounds__(WG) k_rows_rr(const int *items, const int *rb_row0, RowClass c0, const TS0 *arena0,
                                                RowClass c1, const TS1 *arena1, const float *zil,
                                                const unsigned *tmax, const int *perm_t, int nr, int mu_base,
                                                double *y) {
    constexpr int B0 = rr_stage_bytes<TS0>(), B1 = rr_stage_bytes<TS1>();
    __shared__ __align__(16) unsigned char lds[2 * (B0 > B1 ? B0 : B1)];
    const int item = items[blockIdx.x], rb = item & 0xffffff, roff = RR_ROWS * (item >> 24);
    if (rb == 0xffffff) // padding entry of an XCD-ordered list
        return;
    const int r0 = rb_row0[rb] + roff, R = min(RR_ROWS, rb_row0[rb + 1] - r0);
    double accd[2][4] = {};
    // every wave reduces the class-1 scale's slots (lane l: slots l, l + 64,
    // ...), requested before the class-0 walk and reduced after it
    constexpr bool SC1 = Mma16<TS1>::on && Mma16<TS1>::scaled;
    unsigned tv[TMAX_SLOTS / 64] = {};
    if (SC1 && arena1)
#pragma unroll
        for (int i = 0; i < TMAX_SLOTS / 64; i++)
            tv[i] = tmax[(threadIdx.x & 63) + 64 * i];
    rb_class32<TS0>(c0, arena0, rb, roff, R, zil, 1.0f, lds, accd);
    if (arena1) {
        float tsc = 1.0f;
        if constexpr (SC1) {
            float m = 0.0f;
#pragma unroll
            for (int i = 0; i < TMAX_SLOTS / 64; i++)
                m = fmaxf(m, __uint_as_float(tv[i]));
            tsc = pass_scale(wave_max(m));
        }
        rb_class32<TS1>(c1, arena1, rb, roff, R, zil, tsc, lds, accd);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int i = 0; i < 4; i++) { // lane (r, g) holds rows 4 g + i, right-hand sides r and r + 16
        const int row = 16 * w + 4 * g + i;
        if (row < R) {
            double *yo = y + perm_t[r0 + row] + (size_t)nr * mu_base;
            yo[(size_t)nr * r]        = accd[0][i];
            yo[(size_t)nr * (r + 16)] = accd[1][i];
        }
    }
}

// FP64 row-block parts: entries gathered from the RHS-interleaved FP64 z
// (16 bytes = two right-hand sides per load), staged [column][32] as in the
// FP64 V phase; units of four 8-column steps; FP64 sums.
__global__ void __launch_bounds__(WG) k_rows_rr64(const int *items, const int *rb_row0, RowClass c0,
                                                  const double *arena0, RowClass c1, const double *arena1,
                                                  const double *z, const int *perm_t, int nr, int mu_base,
                                                  double *y) {
    constexpr int MU = 32, NS = 4, SC = 8, UPC = RR_CHUNK64 / (SC * NS), XPT = RR_CHUNK64 * MU / 2 / WG;
    __shared__ __align__(16) double s_x[2][RR_CHUNK64 * MU];
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    struct Buf {
        u4 a[NS];
    };
    const int item = items[blockIdx.x], rb = item & 0xffffff, roff = RR_ROWS * (item >> 24);
    if (rb == 0xffffff) // padding entry of an XCD-ordered list
        return;
    const int r0 = rb_row0[rb] + roff, R = min(RR_ROWS, rb_row0[rb + 1] - r0);
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 15, g = lane >> 4;
    const bool real = 16 * w < R;
    const double2 *z2 = reinterpret_cast<const double2 *>(z);
    doublex4 acc0 = {0}, acc1 = {0};
    for (int cls = 0; cls < 2; cls++) {
        const RowClass &rc = cls == 0 ? c0 : c1;
        const double *arena = cls == 0 ? arena0 : arena1;
        if (!arena)
            continue;
        const int cb = rc.rb_col0[rb], C = rc.rb_col0[rb + 1] - cb;
        if (C == 0)
            continue;
        const int Cs     = (C + rc.salign - 1) & ~(rc.salign - 1);
        const int nstep  = (C + SC - 1) / SC;
        const int nchunk = (C + RR_CHUNK64 - 1) / RR_CHUNK64;
        const double *rp = arena + rc.rb_aoff[rb] + (size_t)(roff + min(16 * w + r, R - 1)) * Cs + 2 * g;
        const int *ci    = rc.colidx + cb;
        auto load = [&](int u, Buf &b) {
#pragma unroll
            for (int st = 0; st < NS; st++)
                b.a[st] = nt_load16(rp + (size_t)min(u * NS + st, nstep - 1) * SC);
        };
        double2 xv[XPT];
        int idx[XPT]; // column indices one chunk ahead, as in rb_class32
        auto iload = [&](int j) {
#pragma unroll
            for (int i = 0; i < XPT; i++)
                idx[i] = ci[min(j * RR_CHUNK64 + (((int)threadIdx.x + WG * i) >> 4), C - 1)];
        };
        iload(0);
        auto xload = [&](int j) { // double2 f: column f / 16 of the chunk, right-hand sides 2 (f % 16) ..
#pragma unroll
            for (int i = 0; i < XPT; i++)
                xv[i] = z2[(size_t)idx[i] * (MU / 2) + (((int)threadIdx.x + WG * i) & 15)];
            iload(j + 1);
        };
        auto xstore = [&](int j) {
            double2 *buf = reinterpret_cast<double2 *>(s_x[j & 1]);
#pragma unroll
            for (int i = 0; i < XPT; i++) {
                const int f = (int)threadIdx.x + WG * i;
                buf[f] = j * RR_CHUNK64 + (f >> 4) < C ? xv[i] : make_double2(0.0, 0.0);
            }
        };
        auto run = [&](int u, const Buf &b) {
            if (!real)
                return;
            const int j = u / UPC, s = u % UPC;
            const double *buf = s_x[j & 1];
#pragma unroll
            for (int st = 0; st < NS; st++) {
                const int cs = s * NS + st;
                const double av[2] = {__hiloint2double((int)b.a[st].y, (int)b.a[st].x),
                                      __hiloint2double((int)b.a[st].w, (int)b.a[st].z)};
#pragma unroll
                for (int m = 0; m < 2; m++) {
                    const double *bp = buf + (size_t)(cs * SC + 2 * g + m) * MU;
                    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m], bp[r], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m], bp[r + 16], acc1, 0, 0, 0);
                }
            }
        };
        rr_walk<UPC, Buf>(nchunk, load, run, xload, xstore);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) { // result register i of lane (r, g): row 4 i + g
        const int row = 16 * w + 4 * i + g;
        if (row < R) {
            double *yo = y + perm_t[r0 + row] + (size_t)nr * mu_base;
            yo[(size_t)nr * r]        = acc0[i];
            yo[(size_t)nr * (r + 16)] = acc1[i];
        }
    }
}

// ---------------------------------------------------------------------------
// Complex 16-RHS products on the matrix cores. The 32 B columns of a tile are
// the 16 right-hand sides' real parts (slots 0-15) then imaginary parts (slots
// 16-31), in the blocked z layout (zblk(j, slot)). Each stored complex entry
// (ar, ai) feeds two MFMAs with the same B row: one accumulator for Re(A), one
// for Im(A). Slot c of the result is Re y_s = (Re A) x_re - (Im A) x_im for
// c = s < 16 and Im y_s = (Re A) x_im + (Im A) x_re for c = s + 16; the Im(A)
// accumulator's slots are exchanged by one lane shuffle (c xor 16). Entries
// per 16-byte vector: VE = 2 (FP32 pairs) or 4 (16-bit pairs).
constexpr int MFMA_MU_C   = 16; // complex right-hand sides per matrix-core pass
constexpr int RPASS_C     = 64; // row-block entries staged per pass

template <int VE>
__device__ __forceinline__ void load_b_slots(const float *p, float (&b)[VE]) {
    if constexpr (VE == 2) {
        const float2 v = *reinterpret_cast<const float2 *>(p);
        b[0] = v.x, b[1] = v.y;
    } else {
        const float4 v = *reinterpret_cast<const float4 *>(p);
        b[0] = v.x, b[1] = v.y, b[2] = v.z, b[3] = v.w;
    }
}

// Combines the Re(A) and Im(A) accumulators of a 32x32 tile into slot values.
__device__ __forceinline__ void combine32(const floatx16 &ar, const floatx16 &ai, int lane, double (&accd)[16],
                                          double inv = 1.0) {
    const bool im_slot = (lane & 31) >= 16;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const float t = __shfl_xor(ai[i], 16, 64);
        accd[i] += (double)(im_slot ? ar[i] + t : ar[i] - t) * inv;
    }
}

template <typename TS, int S>
__device__ __forceinline__ void rows_class_mfma_c(const TS *arena, const RowClass &rc, int rb, int R, const float *z,
                                                  float *s_z, int t, int q, bool active, double (&accd)[16]) {
    constexpr int SL   = 32; // B slots
    constexpr int VW   = Vec<TS>::W;
    constexpr int VE   = VW / 2;
    constexpr int WCOL = RPASS_C / S; // entries per wave per pass
    constexpr int NIT  = WCOL / (2 * VE);
    constexpr int ZPT  = RPASS_C * SL / WG;
    constexpr bool PREFETCH_A = S == 4;
    const int cb = rc.rb_col0[rb], C = rc.rb_col0[rb + 1] - cb;
    const int Cp = (C + rc.align - 1) & ~(rc.align - 1);    // columns read (vector width)
    const int Cs = (C + rc.salign - 1) & ~(rc.salign - 1);  // row stride (128-byte rows)
    if (Cp == 0)
        return;
    const TS *A    = arena + 2 * rc.rb_aoff[rb];
    const int lane = threadIdx.x & 63, half = lane >> 5;
    const int row  = 32 * t + (lane & 31);
    const TS *rowp = A + 2 * ((size_t)(row < R ? row : 0) * Cs + q * WCOL);
    const int zs = threadIdx.x & (SL - 1), zc = threadIdx.x / SL;
    float v[ZPT];
    TS a[NIT][VW], an[PREFETCH_A ? NIT : 1][VW];
    auto load_z = [&](int cc) {
        int idx[ZPT];
#pragma unroll
        for (int i = 0; i < ZPT; i++) {
            const int c = cc + zc + (WG / SL) * i;
            idx[i]      = rc.colidx[cb + (c < C ? c : 0)];
        }
#pragma unroll
        for (int i = 0; i < ZPT; i++)
            v[i] = z[zblk(idx[i], zs)];
    };
    auto load_a = [&](int cc, auto &dst) {
        const int nz = min(RPASS_C, Cp - cc);
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int kk = q * WCOL + 2 * VE * it + VE * half;
            Vec<TS>::load(rowp + 2 * (cc + (kk < nz ? kk - q * WCOL : -q * WCOL)), dst[it]);
        }
    };
    load_z(0);
    if (active)
        load_a(0, a);
    for (int cc = 0; cc < Cp; cc += RPASS_C) {
        const int nz = min(RPASS_C, Cp - cc);
        float zmax = 0.0f; // largest staged magnitude (16-bit FP16 products: the pass scale)
#pragma unroll
        for (int i = 0; i < ZPT; i++) {
            const int c      = zc + (WG / SL) * i;
            const float vz   = cc + c < C ? v[i] : 0.0f;
            s_z[zblk(c, zs)] = vz;
            zmax             = fmaxf(zmax, fabsf(vz));
        }
        // the kernel's LDS holds the staged pass and then four floats for the waves' maxima
        float *s_mx = s_z + RPASS_C * SL;
        if constexpr (Mma16<TS>::on && Mma16<TS>::scaled) {
            zmax = wave_max(zmax);
            if ((threadIdx.x & 63) == 0)
                s_mx[threadIdx.x >> 6] = zmax;
        }
        __syncthreads();
        const bool more = cc + RPASS_C < Cp;
        if (more) {
            load_z(cc + RPASS_C);
            if (PREFETCH_A && active)
                load_a(cc + RPASS_C, an);
        }
        if (active) {
            floatx16 ar = {0}, ai = {0};
            double inv  = 1.0;
            if constexpr (Mma16<TS>::on) {
                // two loads (eight complex entries) per lane and step: one
                // 32x32x16 MFMA per 16-bit term for Re(A) and one for Im(A)
                using M  = Mma16<TS>;
                float sc = 1.0f;
                if constexpr (M::scaled) {
                    sc  = pass_scale(fmaxf(fmaxf(s_mx[0], s_mx[1]), fmaxf(s_mx[2], s_mx[3])));
                    inv = 1.0 / (double)sc;
                }
                static_assert(NIT % 2 == 0, "steps come in pairs");
#pragma unroll
                for (int it = 0; it < NIT; it += 2) {
                    const int k0 = q * WCOL + 2 * VE * it;
                    if (k0 < nz) {
                        const int kk = k0 + VE * half;
                        float bb[8];
                        {
                            const float4 b0 = *reinterpret_cast<const float4 *>(s_z + zblk(kk, lane & 31));
                            const float4 b1 = *reinterpret_cast<const float4 *>(s_z + zblk(kk + 2 * VE, lane & 31));
                            bb[0] = b0.x, bb[1] = b0.y, bb[2] = b0.z, bb[3] = b0.w;
                            bb[4] = b1.x, bb[5] = b1.y, bb[6] = b1.z, bb[7] = b1.w;
                        }
                        typename M::V bs[M::NB], are, aim;
                        M::split(bb, sc, bs);
                        deinterleave16<TS>(a[it], a[it + 1], true, true, are, aim);
#pragma unroll
                        for (int u = 0; u < M::NB; u++) {
                            ar = M::mma32(are, bs[u], ar);
                            ai = M::mma32(aim, bs[u], ai);
                        }
                    }
                }
            } else {
#pragma unroll
                for (int it = 0; it < NIT; it++) {
                    const int k0 = q * WCOL + 2 * VE * it;
                    if (k0 < nz) {
                        const int kk = k0 + VE * half;
                        float b[VE];
                        load_b_slots<VE>(s_z + zblk(kk, lane & 31), b);
#pragma unroll
                        for (int e = 0; e < VE; e++) {
                            ar = __builtin_amdgcn_mfma_f32_32x32x2f32(cvt<float>(a[it][2 * e]), b[e], ar, 0, 0, 0);
                            ai = __builtin_amdgcn_mfma_f32_32x32x2f32(cvt<float>(a[it][2 * e + 1]), b[e], ai, 0, 0,
                                                                      0);
                        }
                    }
                }
            }
            combine32(ar, ai, lane, accd, inv);
        }
        __syncthreads();
        if (more && active) {
            if constexpr (PREFETCH_A) {
#pragma unroll
                for (int it = 0; it < NIT; it++)
#pragma unroll
                    for (int u = 0; u < VW; u++)
                        a[it][u] = an[it][u];
            } else {
                load_a(cc + RPASS_C, a);
            }
        }
    }
}

template <typename TS0, typename TS1, int S>
__device__ __forceinline__ void rows_mfma_split_c(const int rb, const int R, const RowClass &c0, const TS0 *arena0,
                                                  const RowClass &c1, const TS1 *arena1, const float *z, float *s_z,
                                                  double (&accd)[16], int &q, bool &active) {
    const int w = threadIdx.x >> 6, T = (R + 31) / 32;
    const int t = w / S;
    q           = w % S;
    active      = t < T;
    rows_class_mfma_c<TS0, S>(arena0, c0, rb, R, z, s_z, t, q, active, accd);
    if (arena1)
        rows_class_mfma_c<TS1, S>(arena1, c1, rb, R, z, s_z, t, q, active, accd);
}

template <typename TS0, typename TS1>
__global__ void __launch_bounds__(WG) k_rows_mfma_c(const int *rb_order, const int *rb_row0, RowClass c0,
                                                    const TS0 *arena0, RowClass c1, const TS1 *arena1,
                                                    const float *z, const int *perm_t, int nr, int mu_base,
                                                    double *y) {
    __shared__ __align__(16) double smem[3 * 16 * 64];
    static_assert(RPASS_C * 32 * sizeof(float) + 4 * sizeof(float) <= sizeof(smem),
                  "staged pass and the waves' maxima fit the workgroup's LDS");
    float *s_z   = reinterpret_cast<float *>(smem);
    const int rb = rb_order[blockIdx.x];
    const int r0 = rb_row0[rb], R = rb_row0[rb + 1] - r0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int T = (R + 31) / 32;
    const int S = T == 1 ? 4 : (T == 2 ? 2 : 1);
    double accd[16];
#pragma unroll
    for (int i = 0; i < 16; i++)
        accd[i] = 0;
    int q       = 0;
    bool active = false;
    if (S == 4)
        rows_mfma_split_c<TS0, TS1, 4>(rb, R, c0, arena0, c1, arena1, z, s_z, accd, q, active);
    else if (S == 2)
        rows_mfma_split_c<TS0, TS1, 2>(rb, R, c0, arena0, c1, arena1, z, s_z, accd, q, active);
    else
        rows_mfma_split_c<TS0, TS1, 1>(rb, R, c0, arena0, c1, arena1, z, s_z, accd, q, active);
    const int t = w / S;
    if (S > 1) {
        if (active && q > 0)
#pragma unroll
            for (int i = 0; i < 16; i++)
                smem[((t * (S - 1) + q - 1) * 16 + i) * 64 + lane] = accd[i];
        __syncthreads();
        if (active && q == 0)
            for (int qq = 1; qq < S; qq++)
#pragma unroll
                for (int i = 0; i < 16; i++)
                    accd[i] += smem[((t * (S - 1) + qq - 1) * 16 + i) * 64 + lane];
    }
    if (active && q == 0) {
        const int c = lane & 31, sr = c & 15, part = c >> 4;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int orow = 32 * t + mfma32_row(i, lane);
            if (orow < R)
                y[2 * ((size_t)perm_t[r0 + orow] + (size_t)nr * (mu_base + sr)) + part] = accd[i];
        }
    }
}

// Stores slot c of row r of a chunk's V x (scaled when the leaf is one chunk).
__device__ __forceinline__ void vx_store_slot(const VTile &d, int r, int c, double v, float *z, double *part,
                                              const double *tscale) {
    if (d.nparts > 1)
        part[(size_t)(d.part_off + r) * 32 + c] = v;
    else
        z[zblk(d.t_index + r, c)] = (float)(tscale ? v * tscale[d.t_index + r] : v);
}

// Chunks of at most 16 rows: v_mfma_f32_16x16x4f32, lane (g, r) = (l / 16, l % 16)
// takes row r and entry 4 VE it + VE g + e; four accumulators (Re/Im A x slots
// 0-15 / 16-31), so lane r holds both parts of right-hand side r.
template <typename TS>
__device__ __forceinline__ void vx_chunk_mfma16_c(const VTile &d, const TS *varena, float *z, double *part,
                                                  const double *tscale, int lane) {
    constexpr int VW  = Vec<TS>::W;
    constexpr int VE  = VW / 2;
    constexpr int NIT = VPASS16 / (4 * VE);
    const int g = lane >> 4, r = lane & 15;
    const TS *rowp = varena + 2 * (d.v_off + (size_t)(r < d.k ? r : 0) * d.ld);
    double re[4] = {}, im[4] = {};
    for (int c0 = 0; c0 < d.ncols; c0 += VPASS16) {
        const int nc = min(VPASS16, d.ncols - c0);
        TS a[NIT][VW];
        float b0[NIT][VE], b1[NIT][VE];
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int kk = 4 * VE * it + VE * g;
            const int kc = kk < nc ? kk : 0;
            Vec<TS>::load(rowp + 2 * (c0 + kc), a[it]);
            const size_t j = (size_t)d.x_off + c0 + kc;
            load_b_slots<VE>(z + zblk(j, r), b0[it]);
            load_b_slots<VE>(z + zblk(j, r + 16), b1[it]);
        }
        floatx4 rr = {0}, ri = {0}, ir = {0}, ii = {0}; // (Re|Im A) x (Re|Im x)
        double inv = 1.0;
        if constexpr (Mma16<TS>::on) {
            // steps in pairs: eight complex entries per lane, one 16x16x32 MFMA
            // per 16-bit term and (Re|Im A, Re|Im x) pair
            using M  = Mma16<TS>;
            float sc = 1.0f;
            if constexpr (M::scaled) {
                float m = 0.0f;
#pragma unroll
                for (int it = 0; it < NIT; it++)
#pragma unroll
                    for (int e = 0; e < VE; e++)
                        m = fmaxf(m, fmaxf(fabsf(b0[it][e]), fabsf(b1[it][e])));
                sc  = pass_scale(wave_max(m));
                inv = 1.0 / (double)sc;
            }
            static_assert(NIT % 2 == 0 && VE == 4, "steps come in pairs of four entries");
#pragma unroll
            for (int it = 0; it < NIT; it += 2) {
                if (4 * VE * it < nc) {
                    const bool in0 = 4 * VE * it + VE * g < nc, in1 = 4 * VE * (it + 1) + VE * g < nc;
                    const float x0[8] = {b0[it][0], b0[it][1], b0[it][2], b0[it][3],
                                         b0[it + 1][0], b0[it + 1][1], b0[it + 1][2], b0[it + 1][3]};
                    const float x1[8] = {b1[it][0], b1[it][1], b1[it][2], b1[it][3],
                                         b1[it + 1][0], b1[it + 1][1], b1[it + 1][2], b1[it + 1][3]};
                    typename M::V s0[M::NB], s1[M::NB], are, aim;
                    M::split(x0, sc, s0);
                    M::split(x1, sc, s1);
                    deinterleave16<TS>(a[it], a[it + 1], in0, in1, are, aim);
#pragma unroll
                    for (int u = 0; u < M::NB; u++) {
                        rr = M::mma16(are, s0[u], rr);
                        ri = M::mma16(are, s1[u], ri);
                        ir = M::mma16(aim, s0[u], ir);
                        ii = M::mma16(aim, s1[u], ii);
                    }
                }
            }
        } else {
#pragma unroll
            for (int it = 0; it < NIT; it++) {
                if (4 * VE * it < nc) {
                    const bool in = 4 * VE * it + VE * g < nc;
#pragma unroll
                    for (int e = 0; e < VE; e++) {
                        const float xr = cvt<float>(a[it][2 * e]), xi = cvt<float>(a[it][2 * e + 1]);
                        const float arv = in ? xr : 0.0f, aiv = in ? xi : 0.0f;
                        rr = __builtin_amdgcn_mfma_f32_16x16x4f32(arv, b0[it][e], rr, 0, 0, 0);
                        ri = __builtin_amdgcn_mfma_f32_16x16x4f32(arv, b1[it][e], ri, 0, 0, 0);
                        ir = __builtin_amdgcn_mfma_f32_16x16x4f32(aiv, b0[it][e], ir, 0, 0, 0);
                        ii = __builtin_amdgcn_mfma_f32_16x16x4f32(aiv, b1[it][e], ii, 0, 0, 0);
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            re[i] += (double)(rr[i] - ii[i]) * inv;
            im[i] += (double)(ri[i] + ir[i]) * inv;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int row = 4 * g + i;
        if (row < d.k) {
            vx_store_slot(d, row, r, re[i], z, part, tscale);
            vx_store_slot(d, row, r + 16, im[i], z, part, tscale);
        }
    }
}

// Chunks of more than 16 rows: v_mfma_f32_32x32x2f32 over 32-row tiles.
template <typename TS>
__device__ __forceinline__ void vx_chunk_mfma32_c(const VTile &d, const TS *varena, float *z, double *part,
                                                  const double *tscale, int lane) {
    constexpr int VW  = Vec<TS>::W;
    constexpr int VE  = VW / 2;
    constexpr int NIT = VPASS32 / (2 * VE);
    const int half = lane >> 5, c = lane & 31;
    for (int rt = 0; rt < d.k; rt += 32) {
        const int row  = rt + c;
        const TS *rowp = varena + 2 * (d.v_off + (size_t)(row < d.k ? row : 0) * d.ld);
        double accd[16];
#pragma unroll
        for (int i = 0; i < 16; i++)
            accd[i] = 0;
        for (int c0 = 0; c0 < d.ncols; c0 += VPASS32) {
            const int nc = min(VPASS32, d.ncols - c0);
            TS a[NIT][VW];
            float b[NIT][VE];
#pragma unroll
            for (int it = 0; it < NIT; it++) {
                const int kk = 2 * VE * it + VE * half;
                const int kc = kk < nc ? kk : 0;
                Vec<TS>::load(rowp + 2 * (c0 + kc), a[it]);
                load_b_slots<VE>(z + zblk((size_t)d.x_off + c0 + kc, c), b[it]);
            }
            floatx16 ar = {0}, ai = {0};
            double inv  = 1.0;
            if constexpr (Mma16<TS>::on) {
                using M  = Mma16<TS>;
                float sc = 1.0f;
                if constexpr (M::scaled) {
                    float m = 0.0f;
#pragma unroll
                    for (int it = 0; it < NIT; it++)
#pragma unroll
                        for (int e = 0; e < VE; e++)
                            m = fmaxf(m, fabsf(b[it][e]));
                    sc  = pass_scale(wave_max(m));
                    inv = 1.0 / (double)sc;
                }
                static_assert(NIT % 2 == 0 && VE == 4, "steps come in pairs of four entries");
#pragma unroll
                for (int it = 0; it < NIT; it += 2) {
                    if (2 * VE * it < nc) {
                        const bool in0 = 2 * VE * it + VE * half < nc, in1 = 2 * VE * (it + 1) + VE * half < nc;
                        const float xb[8] = {b[it][0], b[it][1], b[it][2], b[it][3],
                                             b[it + 1][0], b[it + 1][1], b[it + 1][2], b[it + 1][3]};
                        typename M::V bs[M::NB], are, aim;
                        M::split(xb, sc, bs);
                        deinterleave16<TS>(a[it], a[it + 1], in0, in1, are, aim);
#pragma unroll
                        for (int u = 0; u < M::NB; u++) {
                            ar = M::mma32(are, bs[u], ar);
                            ai = M::mma32(aim, bs[u], ai);
                        }
                    }
                }
            } else {
#pragma unroll
                for (int it = 0; it < NIT; it++) {
                    if (2 * VE * it < nc) {
                        const bool in = 2 * VE * it + VE * half < nc;
#pragma unroll
                        for (int e = 0; e < VE; e++) {
                            ar = __builtin_amdgcn_mfma_f32_32x32x2f32(in ? cvt<float>(a[it][2 * e]) : 0.0f, b[it][e],
                                                                      ar, 0, 0, 0);
                            ai = __builtin_amdgcn_mfma_f32_32x32x2f32(in ? cvt<float>(a[it][2 * e + 1]) : 0.0f,
                                                                      b[it][e], ai, 0, 0, 0);
                        }
                    }
                }
            }
            combine32(ar, ai, lane, accd, inv);
        }
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int r = rt + mfma32_row(i, lane);
            if (r < d.k)
                vx_store_slot(d, r, c, accd[i], z, part, tscale);
        }
    }
}

template <typename TS0, typename TS1>
__global__ void __launch_bounds__(WG) k_vx_mfma_c(const int *wv_off, int nwaves, const VTile *tiles,
                                                  const TS0 *varena0, const TS1 *varena1, float *z, double *part,
                                                  const double *tscale) {
    const int wv = blockIdx.x * NWAVE + (threadIdx.x >> 6);
    if (wv >= nwaves)
        return;
    const int lane = threadIdx.x & 63;
    for (int ti = wv_off[wv]; ti < wv_off[wv + 1]; ti++) {
        const VTile d = tiles[ti];
        if (d.cls == 0) {
            if (d.k <= 16)
                vx_chunk_mfma16_c<TS0>(d, varena0, z, part, tscale, lane);
            else
                vx_chunk_mfma32_c<TS0>(d, varena0, z, part, tscale, lane);
        } else {
            if (d.k <= 16)
                vx_chunk_mfma16_c<TS1>(d, varena1, z, part, tscale, lane);
            else
                vx_chunk_mfma32_c<TS1>(d, varena1, z, part, tscale, lane);
        }
    }
}

// Complex x (nc x 16, original order) into the blocked slots: Re in slot s,
// Im in slot s + 16; zeros from nc on.
__global__ void k_gather_xc_blk(const double *x, const int *perm_s, int nc, int tpad, int mu_base, float *z) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= tpad * MFMA_MU_C)
        return;
    const int j = e / MFMA_MU_C, s = e % MFMA_MU_C;
    float re = 0, im = 0;
    if (j < nc) {
        const double *src = x + 2 * (perm_s[j] + (size_t)nc * (mu_base + s));
        re                = (float)src[0];
        im                = (float)src[1];
    }
    z[zblk(j, s)]      = re;
    z[zblk(j, s + 16)] = im;
}

// The same in two coalesced steps (as for the real 32-wide passes): complex RHS
// s0..s0+7 of original entry i to slots s0.. (Re) and 16 + s0.. (Im) of row
// iperm[i] of an interleaved staging copy, then k_block_x32 into the blocks.
__global__ void k_scatter_xc32(const double *x, const int *iperm_s, int nc, int mu_base, float *stage) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc)
        return;
    const int s0 = 8 * blockIdx.y;
    float re[8], im[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const double2 v = *reinterpret_cast<const double2 *>(x + 2 * (i + (size_t)nc * (mu_base + s0 + u)));
        re[u]           = (float)v.x;
        im[u]           = (float)v.y;
    }
    float4 *o = reinterpret_cast<float4 *>(stage + (size_t)iperm_s[i] * 32 + s0);
    o[0]      = make_float4(re[0], re[1], re[2], re[3]);
    o[1]      = make_float4(re[4], re[5], re[6], re[7]);
    o[4]      = make_float4(im[0], im[1], im[2], im[3]); // slot 16 + s0 is 16 floats on
    o[5]      = make_float4(im[4], im[5], im[6], im[7]);
}

template <typename TS>
__global__ void k_convert(const double *src, TS *dst, uint64_t n) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x)
        dst[e] = (TS)src[e];
}

// ---------------------------------------------------------------------------
// Per-term scaling of low-rank leaves (real or complex), one workgroup per
// leaf: unit U columns and V rows, the product of the norms kept as the term's
// scale. Lets 16-bit storage hold every term near unit magnitude.
__device__ __forceinline__ double sq_mag(double v) { return v * v; }
__device__ __forceinline__ double sq_mag(cdbl v) { return v.re * v.re + v.im * v.im; }
__device__ __forceinline__ double scale_s(double v, double f) { return v * f; }
__device__ __forceinline__ cdbl scale_s(cdbl v, double f) { return cdbl{v.re * f, v.im * f}; }

template <typename E>
__global__ void __launch_bounds__(WG) k_term_scale(const ScaleLeaf *leaves, E *store, double *sig) {
    __shared__ double s_red[NWAVE];
    const ScaleLeaf L = leaves[blockIdx.x];
    for (int j = 0; j < L.k; j++) {
        E *u = store + L.src_u + (size_t)j * L.M;
        E *v = store + L.src_v + (size_t)j * L.N;
        double su = 0, sv = 0;
        for (int a = threadIdx.x; a < L.M; a += WG)
            su += sq_mag(u[a]);
        for (int b = threadIdx.x; b < L.N; b += WG)
            sv += sq_mag(v[b]);
        const double nu = sqrt(block_sum(su, s_red)), nv = sqrt(block_sum(sv, s_red));
        if (nu > 0 && nv > 0) {
            const double fu = 1.0 / nu, fv = 1.0 / nv;
            for (int a = threadIdx.x; a < L.M; a += WG)
                u[a] = scale_s(u[a], fu);
            for (int b = threadIdx.x; b < L.N; b += WG)
                v[b] = scale_s(v[b], fv);
        }
        if (threadIdx.x == 0)
            sig[L.sig_off + j] = nu * nv;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Complex products. Each stored entry is a (re, im) pair of the storage type
// in the real product's layout (the same offsets, counted in complex entries);
// z holds (re, im) pairs of R, RHS-interleaved; arithmetic in R (float, or
// double in FP64 mode); lane-group sums, chunk partial sums and the output in
// FP64.
template <typename R>
struct cpair {
    R re, im;
};

// FP64 complex MU-wide passes: z[(iperm[i] * MU + s) * 2 + {re, im}] = RHS s
// of original entry i, zeros for the padding entries [nc, tpad). A workgroup
// covers WG / (MU / 8) consecutive entries whole: thread (e, sg) reads RHS
// 8 sg .. 8 sg + 7 of its entry (consecutive entries: coalesced) and writes
// them as 128 contiguous bytes, so each entry's row is written by one
// workgroup in full cache lines.
template <int MU>
__global__ void __launch_bounds__(WG) k_scatter_xc64(const double *x, const int *iperm_s, int nc, int tpad,
                                                     int mu_base, double *z) {
    constexpr int E = WG / (MU / 8); // entries per workgroup
    const int i = blockIdx.x * E + (int)threadIdx.x % E, s0 = 8 * ((int)threadIdx.x / E);
    if (i >= tpad)
        return;
    double2 v[8];
#pragma unroll
    for (int u = 0; u < 8; u++)
        v[u] = i < nc ? *reinterpret_cast<const double2 *>(x + 2 * (i + (size_t)nc * (mu_base + s0 + u)))
                      : make_double2(0.0, 0.0);
    double2 *o = reinterpret_cast<double2 *>(z) + (size_t)(i < nc ? iperm_s[i] : i) * MU + s0;
#pragma unroll
    for (int u = 0; u < 8; u++)
        o[u] = v[u];
}

template <typename R, int MU>
__global__ void k_gather_xc(const double *x, const int *perm_s, int nc, int tpad, int mu_base, R *z) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= tpad * MU)
        return;
    const int j = e / MU, s = e % MU;
    R re = 0, im = 0;
    if (j < nc) {
        const double *src = x + 2 * (perm_s[j] + (size_t)nc * (mu_base + s));
        re                = (R)src[0];
        im                = (R)src[1];
    }
    z[2 * (size_t)e]     = re;
    z[2 * (size_t)e + 1] = im;
}

// W stored values of a row (W/2 complex entries) times z entries; adds into acc.
template <typename TS, typename R, int MU>
__device__ __forceinline__ void cmac_vec(const TS *rowp, const R *zc, cpair<R> (&acc)[MU]) {
    constexpr int W = Vec<TS>::W;
    TS a[W];
    Vec<TS>::load(rowp, a);
#pragma unroll
    for (int u = 0; u < W / 2; u++) {
        const R ar = cvt<R>(a[2 * u]), ai = cvt<R>(a[2 * u + 1]);
        const R *zz = zc + (size_t)u * MU * 2;
#pragma unroll
        for (int s = 0; s < MU; s++) {
            const R zr = zz[2 * s], zi = zz[2 * s + 1];
            acc[s].re += ar * zr - ai * zi;
            acc[s].im += ar * zi + ai * zr;
        }
    }
}

template <int MU, typename R>
__device__ __forceinline__ void group_reduce(const cpair<R> (&acc)[MU], int lpr, cpair<double> (&v)[MU]) {
#pragma unroll
    for (int s = 0; s < MU; s++)
        v[s] = cpair<double>{(double)acc[s].re, (double)acc[s].im};
    for (int o = lpr >> 1; o > 0; o >>= 1) {
#pragma unroll
        for (int s = 0; s < MU; s++) {
            v[s].re += __shfl_xor(v[s].re, o, 64);
            v[s].im += __shfl_xor(v[s].im, o, 64);
        }
    }
}

template <typename TS, typename R, int MU>
__device__ __forceinline__ void vx_chunk_c(const VTile &d, const TS *varena, R *z, double *part, const double *tscale,
                                           int lane) {
    constexpr int W = Vec<TS>::W; // stored values per 16-byte vector: W/2 complex entries
    const int nvec  = d.ncols * 2 / W;
    const int lpr   = lanes_per_row(nvec);
    const int rpw   = 64 / lpr;
    const int sl = lane & (lpr - 1), sub = lane / lpr;
    const R *xv = z + (size_t)d.x_off * MU * 2;
    for (int rbase = 0; rbase < d.k; rbase += rpw) {
        const int r = rbase + sub;
        cpair<R> acc[MU];
#pragma unroll
        for (int s = 0; s < MU; s++)
            acc[s] = cpair<R>{0, 0};
        if (r < d.k) {
            const TS *row = varena + 2 * (d.v_off + (size_t)r * d.ld);
            for (int q = sl; q < nvec; q += lpr)
                cmac_vec<TS, R, MU>(row + (size_t)q * W, xv + (size_t)q * (W / 2) * MU * 2, acc);
        }
        cpair<double> v[MU];
        group_reduce<MU>(acc, lpr, v);
        if (sl == 0 && r < d.k) {
            if (d.nparts > 1) {
                double *o = part + 2 * (size_t)(d.part_off + r) * MU;
#pragma unroll
                for (int s = 0; s < MU; s++)
                    o[2 * s] = v[s].re, o[2 * s + 1] = v[s].im;
            } else {
                const double sc = tscale ? tscale[d.t_index + r] : 1.0;
                R *o            = z + 2 * (size_t)(d.t_index + r) * MU;
#pragma unroll
                for (int s = 0; s < MU; s++)
                    o[2 * s] = (R)(v[s].re * sc), o[2 * s + 1] = (R)(v[s].im * sc);
            }
        }
    }
}

template <typename TS0, typename TS1, typename R, int MU>
__global__ void __launch_bounds__(WG) k_vx_c(const int *wv_off, int nwaves, const VTile *tiles, const TS0 *varena0,
                                             const TS1 *varena1, R *z, double *part, const double *tscale) {
    const int wv = blockIdx.x * NWAVE + (threadIdx.x >> 6);
    if (wv >= nwaves)
        return;
    const int lane = threadIdx.x & 63;
    for (int ti = wv_off[wv]; ti < wv_off[wv + 1]; ti++) {
        const VTile d = tiles[ti];
        if (d.cls == 0)
            vx_chunk_c<TS0, R, MU>(d, varena0, z, part, tscale, lane);
        else
            vx_chunk_c<TS1, R, MU>(d, varena1, z, part, tscale, lane);
    }
}

template <typename R, int MU>
__global__ void __launch_bounds__(WG) k_vx_sum_c(const VTile *first, int nmulti,
                                                 const double *part, R *z, const double *tscale,
                                                 unsigned *tmax = nullptr) {
    constexpr int GPB = vx_sum_groups_per_block<MU>();
    constexpr int TPG = WG / GPB;
    const int gi      = blockIdx.x * GPB + (int)threadIdx.x / TPG;
    if (gi >= nmulti)
        return;
    const VTile d    = first[gi];
    const int km     = d.k * MU;
    const double *pb = part + 2 * (size_t)d.part_base * MU;
    float m = 0.0f;
    // gridDim.y workgroups share one group's entries
    for (int e = threadIdx.x % TPG + TPG * blockIdx.y; e < km; e += TPG * gridDim.y) {
        double v[2] = {0, 0};
        sum_parts<2>(pb + 2 * (size_t)e, 2 * (size_t)km, d.nparts, v);
        const double sc = tscale ? tscale[d.t_index + e / MU] : 1.0;
        R *o            = z + 2 * ((size_t)d.t_index * MU + e);
        o[0]            = (R)(v[0] * sc);
        o[1]            = (R)(v[1] * sc);
        m = fmaxf(m, fmaxf(fabsf((float)o[0]), fabsf((float)o[1])));
    }
    if (tmax && d.cls == 1) { // complex 32-RHS FP32 passes: one group per workgroup
        m = wave_max(m);
        if ((threadIdx.x & 63) == 0)
            atomicMax(tmax + ((blockIdx.x + blockIdx.y) & (TMAX_SLOTS - 1)), __float_as_uint(m));
    }
}

constexpr int ZCAP_C = 1024; // complex vector entries staged in LDS per pass

template <typename TS, typename R, int MU>
__device__ __forceinline__ void rows_class_c(const TS *arena, const RowClass &rc, int rb, int R_, const R *z, R *s_z,
                                             cpair<double> *s_acc) {
    constexpr int W     = Vec<TS>::W;
    constexpr int ZCOLS = ZCAP_C / MU;
    const int c0 = rc.rb_col0[rb], C = rc.rb_col0[rb + 1] - c0;
    const int Cp = (C + rc.align - 1) & ~(rc.align - 1);    // columns read (vector width)
    const int Cs = (C + rc.salign - 1) & ~(rc.salign - 1);  // row stride (128-byte rows)
    const TS *A  = arena + 2 * rc.rb_aoff[rb];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int cc = 0; cc < Cp; cc += ZCOLS) {
        const int nz = min(ZCOLS, Cp - cc);
        for (int e = threadIdx.x; e < nz * MU; e += WG) {
            const int c = cc + e / MU;
            R re = 0, im = 0;
            if (c < C) {
                const R *src = z + 2 * ((size_t)rc.colidx[c0 + c] * MU + e % MU);
                re           = src[0];
                im           = src[1];
            }
            s_z[2 * e]     = re;
            s_z[2 * e + 1] = im;
        }
        __syncthreads();
        const int nvec = nz * 2 / W;
        const int lpr  = lanes_per_row(nvec);
        const int rpw  = 64 / lpr;
        const int sl = lane & (lpr - 1), sub = lane / lpr;
        for (int rbase = w * rpw; rbase < R_; rbase += NWAVE * rpw) {
            const int r = rbase + sub;
            cpair<R> acc[MU];
#pragma unroll
            for (int s = 0; s < MU; s++)
                acc[s] = cpair<R>{0, 0};
            if (r < R_) {
                const TS *row = A + 2 * ((size_t)r * Cs + cc);
                for (int q = sl; q < nvec; q += lpr)
                    cmac_vec<TS, R, MU>(row + (size_t)q * W, s_z + (size_t)q * (W / 2) * MU * 2, acc);
            }
            cpair<double> v[MU];
            group_reduce<MU>(acc, lpr, v);
            if (sl == 0 && r < R_)
#pragma unroll
                for (int s = 0; s < MU; s++) {
                    s_acc[r * MU + s].re += v[s].re;
                    s_acc[r * MU + s].im += v[s].im;
                }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Complex FP64 16-RHS products on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64). A complex entry a = ar + i ai is one 16-byte
// load; lane (g, r) = (l / 16, l % 16) takes row r of a 16-row tile and, in
// step it, complex column 4 it + g. Four real MFMAs per step:
// re += ar br - ai bi, im += ar bi + ai br, where (br, bi) is the vector entry
// of that column for right-hand side r. Result register i of lane (g, r) is
// row 4 i + g, right-hand side r. z, the partial sums and y keep the layout of
// the other complex paths (re, im interleaved, RHS-interleaved).
__device__ __forceinline__ void cmfma64(double ar, double ai, double br, double bi, doublex4 &re, doublex4 &im) {
    re = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, re, 0, 0, 0);
    re = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi, re, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, im, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, im, 0, 0, 0);
}

constexpr int MU_C64   = 16; // complex right-hand sides of the LDS-staged FP64 kernels
constexpr int RPASS_C64 = 32;                // complex row-block columns staged per pass
constexpr int ZST_C64   = 2 * MU_C64 + 2;    // doubles per staged column (padded against bank conflicts)


__global__ void __launch_bounds__(WG) k_rows_mfma64c16(const int *rb_order, const int *rb_row0, RowClass c0,
                                                     const double *arena0, RowClass c1, const double *arena1,
                                                     const double *z, const int *perm_t, int nr, int mu_base,
                                                     double *y) {
    constexpr int MU  = MU_C64;
    constexpr int NIT = RPASS_C64 / 4;
    constexpr int ZPT = RPASS_C64 * MU / WG; // staged complex entries per thread
    __shared__ double s_z[RPASS_C64 * ZST_C64];
    const int rb = rb_order[blockIdx.x];
    const int r0 = rb_row0[rb], R = rb_row0[rb + 1] - r0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int T16 = (R + 15) / 16; // <= 8 tiles; wave w takes tiles w and w + 4
    doublex4 acc_re[2] = {}, acc_im[2] = {};
    const int zs = threadIdx.x & (MU - 1), zc = threadIdx.x / MU;
    for (int cls = 0; cls < 2; cls++) {
        const RowClass &rc   = cls == 0 ? c0 : c1;
        const double *arena = cls == 0 ? arena0 : arena1;
        if (!arena)
            continue;
        const int cb = rc.rb_col0[rb], C = rc.rb_col0[rb + 1] - cb;
        const int Cp = (C + rc.align - 1) & ~(rc.align - 1);    // columns read (vector width)
        const int Cs = (C + rc.salign - 1) & ~(rc.salign - 1);  // row stride (128-byte rows)
        const double *A = arena + 2 * rc.rb_aoff[rb];
        for (int cc = 0; cc < Cp; cc += RPASS_C64) {
            const int nz = min(RPASS_C64, Cp - cc);
            {
                int idx[ZPT];
#pragma unroll
                for (int i = 0; i < ZPT; i++) {
                    const int c = cc + zc + (WG / MU) * i;
                    idx[i]      = rc.colidx[cb + (c < C ? c : 0)];
                }
#pragma unroll
                for (int i = 0; i < ZPT; i++) {
                    const int c = zc + (WG / MU) * i;
                    double v[2];
                    Vec<double>::load(z + 2 * ((size_t)idx[i] * MU + zs), v);
                    const bool ok             = cc + c < C;
                    s_z[c * ZST_C64 + 2 * zs]     = ok ? v[0] : 0.0;
                    s_z[c * ZST_C64 + 2 * zs + 1] = ok ? v[1] : 0.0;
                }
            }
            __syncthreads();
#pragma unroll
            for (int tt = 0; tt < 2; tt++) {
                const int t = w + 4 * tt;
                if (t < T16) {
                    const int row    = 16 * t + r16;
                    const double *rp = A + 2 * ((size_t)(row < R ? row : 0) * Cs + cc);
                    double a[NIT][2];
#pragma unroll
                    for (int it = 0; it < NIT; it++) {
                        const int kk = 4 * it + g;
                        Vec<double>::load(rp + 2 * (kk < nz ? kk : 0), a[it]);
                    }
#pragma unroll
                    for (int it = 0; it < NIT; it++) {
                        if (4 * it < nz) {
                            const int kk     = 4 * it + g;
                            const bool in    = kk < nz;
                            const double *bz = s_z + (in ? kk : 0) * ZST_C64 + 2 * r16;
                            cmfma64(in ? a[it][0] : 0.0, in ? a[it][1] : 0.0, bz[0], bz[1], acc_re[tt], acc_im[tt]);
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int tt = 0; tt < 2; tt++) {
        const int t = w + 4 * tt;
        if (t < T16) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int orow = 16 * t + 4 * i + g;
                if (orow < R) {
                    double *o = y + 2 * ((size_t)perm_t[r0 + orow] + (size_t)nr * (mu_base + r16));
                    o[0]      = acc_re[tt][i];
                    o[1]      = acc_im[tt][i];
                }
            }
        }
    }
}

__global__ void __launch_bounds__(WG) k_vx_mfma64c16(const int *wv_off, int nwaves, const VTile *tiles,
                                                   const double *varena0, const double *varena1, double *z,
                                                   double *part, const double *tscale) {
    constexpr int MU  = MU_C64;
    constexpr int NIT = 8; // 32 complex columns per pass
    const int wv      = blockIdx.x * NWAVE + (threadIdx.x >> 6);
    if (wv >= nwaves)
        return;
    const int lane = threadIdx.x & 63, g = lane >> 4, r16 = lane & 15;
    for (int ti = wv_off[wv]; ti < wv_off[wv + 1]; ti++) {
        const VTile d       = tiles[ti];
        const double *vare  = d.cls == 0 ? varena0 : varena1;
        const double *xv    = z + 2 * (size_t)d.x_off * MU;
        for (int rt = 0; rt < d.k; rt += 16) {
            const int row    = rt + r16;
            const double *rp = vare + 2 * (d.v_off + (size_t)(row < d.k ? row : 0) * d.ld);
            doublex4 acc_re = {0}, acc_im = {0};
            for (int c0 = 0; c0 < d.ncols; c0 += 4 * NIT) {
                const int nc = min(4 * NIT, d.ncols - c0);
                double a[NIT][2];
#pragma unroll
                for (int it = 0; it < NIT; it++) {
                    const int kk = 4 * it + g;
                    Vec<double>::load(rp + 2 * (c0 + (kk < nc ? kk : 0)), a[it]);
                }
#pragma unroll
                for (int it = 0; it < NIT; it++) {
                    if (4 * it < nc) {
                        const int kk  = 4 * it + g;
                        const bool in = kk < nc;
                        double b[2];
                        Vec<double>::load(xv + 2 * ((size_t)(c0 + (in ? kk : 0)) * MU + r16), b);
                        cmfma64(in ? a[it][0] : 0.0, in ? a[it][1] : 0.0, b[0], b[1], acc_re, acc_im);
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = rt + 4 * i + g;
                if (r < d.k) {
                    if (d.nparts > 1) {
                        double *o = part + 2 * ((size_t)(d.part_off + r) * MU + r16);
                        o[0]      = acc_re[i];
                        o[1]      = acc_im[i];
                    } else {
                        const double sc = tscale ? tscale[d.t_index + r] : 1.0;
                        double *o       = z + 2 * ((size_t)(d.t_index + r) * MU + r16);
                        o[0]            = acc_re[i] * sc;
                        o[1]            = acc_im[i] * sc;
                    }
                }
            }
        }
    }
}


// 32 complex right-hand sides per pass (FP64), without LDS: each wave works
// on its own 16-row tile with the vector entries read from L2, so no barrier
// ever stalls the matrix cores; each A entry a lane loads feeds 6 MFMAs (two
// 16-wide RHS tiles, three MFMAs per complex product). The z block of 32
// complex RHS is far larger than the caches, which caps this form below the
// matrix-core roofline; at 16 RHS the LDS-staged kernels above are faster.
// ---------------------------------------------------------------------------
// Complex FP64 32-RHS products by row ranges: the real FP64 walk (rr_walk)
// with (re, im) entries. A chunk is RRC_CHUNK columns of the vector entries
// for 32 complex right-hand sides (512 bytes per column), staged in LDS as in
// z ([column][32][re, im]); a step is four columns (lane (r, g) loads the
// complex entry of row r, column 4 step + g, as one 16-byte load) and feeds
// three real MFMAs per 16 right-hand sides (Gauss: K1 += br (ar + ai),
// K2 += ar (bi - br), K3 += ai (br + bi); re = K1 - K3, im = K1 + K2).
constexpr int RRC_CHUNK = 32;
typedef unsigned u4x __attribute__((ext_vector_type(4)));
struct RrcUnit { // one unit's raw 16-byte loads (four steps)
    u4x a[4];
};

template <typename Load, typename XLoad, typename XStore>
__device__ __forceinline__ void rrc_walk(int nchunk, int nstep, bool real, const double *s_x, Load load, XLoad xload,
                                         XStore xstore, doublex4 (&acc)[2][3]) {
    constexpr int MU = 32, NS = 4, UPC = RRC_CHUNK / (4 * NS);
    using Buf = RrcUnit;
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    auto run = [&](int u, const Buf &b) {
        if (!real)
            return;
        const int j = u / UPC, s = u % UPC;
        const double2 *buf = reinterpret_cast<const double2 *>(s_x + (size_t)(j & 1) * RRC_CHUNK * MU * 2);
#pragma unroll
        for (int st = 0; st < NS; st++) {
            const int cs = s * NS + st; // step within the chunk
            const double ar = __hiloint2double((int)b.a[st].y, (int)b.a[st].x);
            const double ai = __hiloint2double((int)b.a[st].w, (int)b.a[st].z);
            const double as = ar + ai;
            const double2 *bp = buf + (size_t)(4 * cs + g) * MU;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const double2 x = bp[16 * h + r];
                acc[h][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(as, x.x, acc[h][0], 0, 0, 0);
                acc[h][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, x.y - x.x, acc[h][1], 0, 0, 0);
                acc[h][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, x.x + x.y, acc[h][2], 0, 0, 0);
            }
        }
    };
    rr_walk<UPC, Buf>(nchunk, load, run, xload, xstore);
}

// V phase: at most RR_ROWS stacked complex V rows of one panel class over a
// column range (VTile), one workgroup each; x staged from the interleaved z.
__global__ void __launch_bounds__(WG) k_vx_rrc64(const VTile *items, const double *varena0, const double *varena1,
                                                 double *z, double *part, const double *tscale) {
    constexpr int MU = 32, NS = 4, SC = 4, XPT = RRC_CHUNK * MU / WG; // complex entries staged per thread (4)
    __shared__ __align__(16) double s_x[2 * RRC_CHUNK * MU * 2];
    const VTile d = items[blockIdx.x];
    if (d.k == 0) // padding entry of an XCD-ordered list
        return;
    const double *varena = d.cls == 0 ? varena0 : varena1;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 15, g = lane >> 4;
    const int K      = d.k;
    const int nstep  = (d.ncols + SC - 1) / SC;
    const int nchunk = (d.ncols + RRC_CHUNK - 1) / RRC_CHUNK;
    const double *rp = varena + 2 * (d.v_off + (size_t)min(16 * w + r, K - 1) * d.ld + g);
    auto load = [&](int u, RrcUnit &b) {
#pragma unroll
        for (int st = 0; st < NS; st++)
            b.a[st] = nt_load16(rp + 2 * (size_t)min(u * NS + st, nstep - 1) * SC);
    };
    const double2 *xsrc = reinterpret_cast<const double2 *>(z) + (size_t)d.x_off * MU;
    double2 xv[XPT];
    auto xload = [&](int j) { // complex entry f: column f / 32 of the chunk, RHS f % 32
#pragma unroll
        for (int i = 0; i < XPT; i++) {
            const int f = (int)threadIdx.x + WG * i;
            xv[i] = xsrc[(size_t)min(j * RRC_CHUNK + (f >> 5), d.ncols - 1) * MU + (f & 31)];
        }
    };
    auto xstore = [&](int j) {
        double2 *buf = reinterpret_cast<double2 *>(s_x) + (size_t)(j & 1) * RRC_CHUNK * MU;
#pragma unroll
        for (int i = 0; i < XPT; i++) {
            const int f = (int)threadIdx.x + WG * i;
            buf[f] = j * RRC_CHUNK + (f >> 5) < d.ncols ? xv[i] : make_double2(0.0, 0.0);
        }
    };
    doublex4 acc[2][3] = {};
    rrc_walk(nchunk, nstep, 16 * w < K, s_x, load, xload, xstore, acc);
#pragma unroll
    for (int i = 0; i < 4; i++) { // result register i of lane (r, g): row 4 i + g, RHS 16 h + r
        const int row = 16 * w + 4 * i + g;
        if (row < K) {
            const double f = (tscale && d.nparts == 1) ? tscale[d.t_index + row] : 1.0;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                double2 *o = reinterpret_cast<double2 *>(d.nparts > 1 ? part : z) +
                             (size_t)((d.nparts > 1 ? d.part_off : d.t_index) + row) * MU + 16 * h + r;
                *o = make_double2((acc[h][0][i] - acc[h][2][i]) * f, (acc[h][0][i] + acc[h][1][i]) * f);
            }
        }
    }
}

// Row phase: row-block parts of at most RR_ROWS rows (items as in the real
// 32-RHS rows), vector entries gathered through the column indices from the
// interleaved z, both classes into the same sums.
__global__ void __launch_bounds__(WG) k_rows_rrc64(const int *items, const int *rb_row0, RowClass c0,
                                                   const double *arena0, RowClass c1, const double *arena1,
                                                   const double *z, const int *perm_t, int nr, int mu_base,
                                                   double *y) {
    constexpr int MU = 32, NS = 4, SC = 4, XPT = RRC_CHUNK * MU / WG;
    __shared__ __align__(16) double s_x[2 * RRC_CHUNK * MU * 2];
    const int item = items[blockIdx.x], rb = item & 0xffffff, roff = RR_ROWS * (item >> 24);
    if (rb == 0xffffff) // padding entry of an XCD-ordered list
        return;
    const int r0 = rb_row0[rb] + roff, R = min(RR_ROWS, rb_row0[rb + 1] - r0);
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 15, g = lane >> 4;
    const double2 *z2 = reinterpret_cast<const double2 *>(z);
    doublex4 acc[2][3] = {};
    for (int cls = 0; cls < 2; cls++) {
        const RowClass &rc  = cls == 0 ? c0 : c1;
        const double *arena = cls == 0 ? arena0 : arena1;
        if (!arena)
            continue;
        const int cb = rc.rb_col0[rb], C = rc.rb_col0[rb + 1] - cb;
        if (C == 0)
            continue;
        const int Cs     = (C + rc.salign - 1) & ~(rc.salign - 1);
        const int nstep  = (C + SC - 1) / SC;
        const int nchunk = (C + RRC_CHUNK - 1) / RRC_CHUNK;
        const double *rp = arena + 2 * (rc.rb_aoff[rb] + (size_t)(roff + min(16 * w + r, R - 1)) * Cs + g);
        const int *ci    = rc.colidx + cb;
        auto load = [&](int u, RrcUnit &b) {
#pragma unroll
            for (int st = 0; st < NS; st++)
                b.a[st] = nt_load16(rp + 2 * (size_t)min(u * NS + st, nstep - 1) * SC);
        };
        double2 xv[XPT];
        int idx[XPT]; // column indices one chunk ahead, as in rb_class32
        auto iload = [&](int j) {
#pragma unroll
            for (int i = 0; i < XPT; i++)
                idx[i] = ci[min(j * RRC_CHUNK + (((int)threadIdx.x + WG * i) >> 5), C - 1)];
        };
        iload(0);
        auto xload = [&](int j) {
#pragma unroll
            for (int i = 0; i < XPT; i++)
                xv[i] = z2[(size_t)idx[i] * MU + (((int)threadIdx.x + WG * i) & 31)];
            iload(j + 1);
        };
        auto xstore = [&](int j) {
            double2 *buf = reinterpret_cast<double2 *>(s_x) + (size_t)(j & 1) * RRC_CHUNK * MU;
#pragma unroll
            for (int i = 0; i < XPT; i++) {
                const int f = (int)threadIdx.x + WG * i;
                buf[f] = j * RRC_CHUNK + (f >> 5) < C ? xv[i] : make_double2(0.0, 0.0);
            }
        };
        rrc_walk(nchunk, nstep, 16 * w < R, s_x, load, xload, xstore, acc);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int row = 16 * w + 4 * i + g;
        if (row < R) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                double *o = y + 2 * ((size_t)perm_t[r0 + row] + (size_t)nr * (mu_base + 16 * h + r));
                o[0]      = acc[h][0][i] - acc[h][2][i];
                o[1]      = acc[h][0][i] + acc[h][1][i];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Complex products in FP32 arithmetic by row ranges, MUC = 32 or 64 complex
// right-hand sides per pass (the FP64 walk above with FP32 tiles). The vector
// entries are RHS-interleaved (re, im) FP32 pairs, 8 MUC bytes per entry (zc:
// x from the gather, t from the V phase). A chunk is rrc_chunk<MUC> columns
// (64 for 32 RHS, 32 for 64: 16 KB of FP32 x either way). FP32 entries: a
// lane's 16-byte load is two complex entries (columns 2 g, 2 g + 1 of an
// 8-column step, MFMA m taking column 2 g + m as k = g), three
// v_mfma_f32_16x16x4f32 per 16 right-hand sides (Gauss, as in FP64), x staged
// as in zc. 16-bit entries: two loads give eight complex entries (columns
// 8 g .. 8 g + 7 of a 32-column step), deinterleaved into Re and Im operands of
// v_mfma_f32_16x16x32, x staged split into 16-bit terms per part ([term][re,
// im][column / 8][RHS][column % 8]); four products per term (Re Re, Im Im,
// Re Im, Im Re: no sum of two 16-bit entries is formed, so their rounding
// stays the storage's). Unit sums FP32, then FP64. 64 right-hand sides read
// every stored entry once per 64 instead of per 32.
template <int MUC>
constexpr int rrc_chunk() {
    return MUC == 32 ? 64 : 32;
}

template <typename TS, int MUC>
constexpr int rrc32_stage_bytes() {
    if constexpr (Mma16<TS>::on)
        return Mma16<TS>::NB * 2 * rrc_chunk<MUC>() * MUC * 2;
    else
        return rrc_chunk<MUC>() * MUC * 8;
}

// x of a chunk in registers between its loads and its store to LDS: FP32
// entries (float4 = two right-hand sides of a column), 16-bit entries (float2 =
// one right-hand side of a column).
struct RrcX32 {
    float4 v[4];
};
struct RrcX16 {
    float2 v[8];
};

// The staged chunk, FP32 entries: v[i] = the (re, im) pairs of right-hand
// sides 2 q, 2 q + 1 of chunk column f / (MUC / 2) (q = f % (MUC / 2), f =
// thread + WG i), zero past the columns, stored as loaded.
template <int MUC>
__device__ __forceinline__ void rrc32_store(unsigned char *buf, const float4 (&v)[4], int j, int ncols) {
    constexpr int CH = rrc_chunk<MUC>(), QP = MUC / 2;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int f = (int)threadIdx.x + WG * i;
        reinterpret_cast<float4 *>(buf)[f] = j * CH + f / QP < ncols ? v[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// The staged chunk, 16-bit entries: thread (s, cg) = (thread % MUC, thread /
// MUC) holds v[c] = the (re, im) pair of right-hand side s at chunk column
// 8 cg + c (zero past the columns: rrc16_load) and writes each 16-bit term of
// the eight re and of the eight im parts (FP16: times sc) as one 16-byte
// vector, in the layout [term][re, im][column / 8][RHS][column % 8]: a wave
// writes 1 KB contiguous.
template <typename TS, int MUC>
__device__ __forceinline__ void rrc16_store(unsigned char *buf, const float2 (&v)[8], float sc) {
    using M            = Mma16<TS>;
    constexpr int CH   = rrc_chunk<MUC>();
    const int s = (int)threadIdx.x % MUC, cg = (int)threadIdx.x / MUC;
    float re[8], im[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        re[c] = v[c].x;
        im[c] = v[c].y;
    }
    typename M::V ore[M::NB], oim[M::NB];
    M::split(re, sc, ore);
    M::split(im, sc, oim);
    typename M::V *bv = reinterpret_cast<typename M::V *>(buf);
#pragma unroll
    for (int t = 0; t < M::NB; t++) {
        bv[((size_t)(t * 2 + 0) * (CH / 8) + cg) * MUC + s] = ore[t];
        bv[((size_t)(t * 2 + 1) * (CH / 8) + cg) * MUC + s] = oim[t];
    }
}

// Thread (s, cg)'s x of chunk j for rrc16_store: v[c] = entry col(8 cg + c)
// of right-hand side s (z2: (re, im) pairs, MUC per entry), zero for chunk
// columns at or past ncols; col(i) = the z entry of chunk column i, or, for an
// int col, col + i.
template <int MUC, typename Col>
__device__ __forceinline__ void rrc16_load(float2 (&v)[8], const float2 *z2, int j, int ncols, Col col) {
    constexpr int CH = rrc_chunk<MUC>();
    // 64 RHS: the column group is the wave's (made wave-uniform, so the column
    // indices are scalar loads)
    const int cg = MUC == 64 ? __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6) : (int)threadIdx.x / MUC;
    const int s = (int)threadIdx.x % MUC, c0 = j * CH + 8 * cg;
    if (c0 + 8 <= ncols) {
        if constexpr (std::is_same<Col, int>::value) { // entries col + i: one address, immediate offsets
            const float2 *p = z2 + ((size_t)col + c0) * MUC + s;
#pragma unroll
            for (int c = 0; c < 8; c++)
                v[c] = p[c * MUC];
        } else {
            int e[8];
#pragma unroll
            for (int c = 0; c < 8; c++)
                e[c] = col(c0 + c);
#pragma unroll
            for (int c = 0; c < 8; c++)
                v[c] = z2[(size_t)e[c] * MUC + s];
        }
    } else if constexpr (std::is_same<Col, int>::value) {
#pragma unroll
        for (int c = 0; c < 8; c++)
            v[c] = c0 + c < ncols ? z2[((size_t)col + c0 + c) * MUC + s] : make_float2(0.0f, 0.0f);
    } else {
#pragma unroll
        for (int c = 0; c < 8; c++)
            v[c] = c0 + c < ncols ? z2[(size_t)col(c0 + c) * MUC + s] : make_float2(0.0f, 0.0f);
    }
}

// One work item's rows (wave w: rows 16 w + r) over the columns walked by
// rr_walk / rr_walk2; acc: [RHS sixteen][re, im][row] FP64 sums, fed by FP32
// sums over one unit (FP32 entries) or two (16-bit entries; ACC32: over the
// whole walk, for 16-bit entries whose own rounding dwarfs that of FP32 sums,
// which keeps the FP64 sums out of the loop's registers).
template <typename TS, int MUC, typename XS, bool ACC32, typename Load, typename XLoad, typename XStore>
__device__ __forceinline__ void rrc32_walk(int nchunk, bool real, const unsigned char *lds, float sc, Load load,
                                           XLoad xload, XStore xstore, double (&accd)[MUC / 16][2][4]) {
    constexpr bool H16 = Mma16<TS>::on;
    constexpr int CH   = rrc_chunk<MUC>();
    constexpr int NS   = H16 ? (MUC == 32 ? 2 : 1) : 4; // steps per unit
    constexpr int SC   = H16 ? 32 : 8;                  // columns per step
    constexpr int UPC  = CH / (SC * NS);                // units per chunk (1 or 2)
    constexpr int BUF  = rrc32_stage_bytes<TS, MUC>();
    constexpr int NH   = MUC / 16;
    static_assert(UPC >= 1 && UPC * SC * NS == CH, "whole units per chunk");
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const int U = nchunk * UPC;
    const double inv = 1.0 / (double)sc;
    floatx4 P16[NH][2] = {}; // 16-bit entries: [sixteen][re, im] FP32 sums of the current two units
    auto run = [&](int u, const RrcUnit &b) {
        if (!real)
            return;
        const int j = u / UPC, s = u % UPC;
        const unsigned char *buf = lds + (j & 1) * BUF;
        if constexpr (!H16) {
            floatx4 K[NH][3] = {};
#pragma unroll
            for (int st = 0; st < NS; st++) {
                const int cs = s * NS + st;
                const float a[4] = {__uint_as_float(b.a[st].x), __uint_as_float(b.a[st].y),
                                    __uint_as_float(b.a[st].z), __uint_as_float(b.a[st].w)};
#pragma unroll
                for (int m = 0; m < 2; m++) {
                    const float ar = a[2 * m], ai = a[2 * m + 1], as = ar + ai;
                    const float2 *bp = reinterpret_cast<const float2 *>(buf) + (size_t)(cs * 8 + 2 * g + m) * MUC;
#pragma unroll
                    for (int h = 0; h < NH; h++) {
                        const float2 x = bp[16 * h + r];
                        K[h][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(as, x.x, K[h][0], 0, 0, 0);
                        K[h][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, x.y - x.x, K[h][1], 0, 0, 0);
                        K[h][2] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, x.x + x.y, K[h][2], 0, 0, 0);
                    }
                }
            }
            // re = K0 - K2, im = K0 + K1 formed in FP32: their rounding (2^-24 of the
            // result) is below that of the FP32 sums K themselves
#pragma unroll
            for (int h = 0; h < NH; h++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    accd[h][0][i] += (double)(K[h][0][i] - K[h][2][i]);
                    accd[h][1][i] += (double)(K[h][0][i] + K[h][1][i]);
                }
        } else {
            using M = Mma16<TS>;
            typename M::V are[NS], aim[NS], aimn[NS]; // aimn = -aim (sign bits flipped)
#pragma unroll
            for (int st = 0; st < NS; st++) {
                TS p0[8], p1[8];
                const u4x w0 = b.a[2 * st], w1 = b.a[2 * st + 1];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const unsigned u0 = w0[e], u1 = w1[e];
                    p0[2 * e]     = __builtin_bit_cast(TS, (unsigned short)(u0 & 0xffffu));
                    p0[2 * e + 1] = __builtin_bit_cast(TS, (unsigned short)(u0 >> 16));
                    p1[2 * e]     = __builtin_bit_cast(TS, (unsigned short)(u1 & 0xffffu));
                    p1[2 * e + 1] = __builtin_bit_cast(TS, (unsigned short)(u1 >> 16));
                }
                deinterleave16<TS>(p0, p1, true, true, are[st], aim[st]);
                aimn[st] = __builtin_bit_cast(typename M::V, __builtin_bit_cast(u4x, aim[st]) ^ 0x80008000u);
            }
            const typename M::V *bp = reinterpret_cast<const typename M::V *>(buf);
            // re += Re a Re x - Im a Im x, im += Re a Im x + Im a Re x (two MFMAs each)
#pragma unroll
            for (int h = 0; h < NH; h++)
#pragma unroll
                for (int st = 0; st < NS; st++) {
                    const int cs = s * NS + st;
#pragma unroll
                    for (int t = 0; t < M::NB; t++) {
                        const typename M::V xr = bp[((size_t)(t * 2 + 0) * (CH / 8) + cs * 4 + g) * MUC + 16 * h + r];
                        const typename M::V xi = bp[((size_t)(t * 2 + 1) * (CH / 8) + cs * 4 + g) * MUC + 16 * h + r];
                        P16[h][0] = M::mma16(are[st], xr, P16[h][0]);
                        P16[h][0] = M::mma16(aimn[st], xi, P16[h][0]);
                        P16[h][1] = M::mma16(are[st], xi, P16[h][1]);
                        P16[h][1] = M::mma16(aim[st], xr, P16[h][1]);
                    }
                }
            // FP32 sums over two units, then FP64 (ACC32: over the whole walk)
            if constexpr (!ACC32) {
                if ((u & 1) || u == U - 1) {
#pragma unroll
                    for (int h = 0; h < NH; h++)
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            accd[h][0][i] += (double)P16[h][0][i] * inv;
                            accd[h][1][i] += (double)P16[h][1][i] * inv;
                            P16[h][0][i] = 0.0f;
                            P16[h][1][i] = 0.0f;
                        }
                }
            }
        }
    };
    if constexpr (UPC == 1) {
        rr_walk2<RrcUnit, XS>(nchunk, load, run, xload, xstore);
    } else {
        XS x;
        rr_walk<UPC, RrcUnit>(
            nchunk, load, run, [&](int j) { xload(j, x); }, [&](int j) { xstore(j, x); });
    }
    if constexpr (H16 && ACC32) {
        if (real) {
#pragma unroll
            for (int h = 0; h < NH; h++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    accd[h][0][i] += (double)P16[h][0][i] * inv;
                    accd[h][1][i] += (double)P16[h][1][i] * inv;
                }
        }
    }
}

// The largest |re|, |im| of the x entries of each 32-column block of zc (the
// FP16 scale of the complex V phase): one workgroup per block.
template <int MUC>
__global__ void __launch_bounds__(WG) k_colmax_c32(const float *zc, int nc, float *xmax) {
    constexpr int FPC = MUC / 2; // float4 per column
    __shared__ float s_m[NWAVE];
    const int b = blockIdx.x;
    const float4 *src = reinterpret_cast<const float4 *>(zc) + (size_t)b * 32 * FPC;
    float m = 0.0f;
    for (int f = threadIdx.x; f < 32 * FPC; f += WG)
        if (b * 32 + f / FPC < nc) {
            const float4 v = src[f];
            m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
        }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0)
        s_m[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        xmax[b] = fmaxf(fmaxf(s_m[0], s_m[1]), fmaxf(s_m[2], s_m[3]));
}

// zc[(iperm[i] * MUC + s) * 2 + {re, im}] = RHS s of original entry i as FP32.
// A workgroup covers WG / (MUC / 8) consecutive entries whole: thread (e, sg)
// reads RHS 8 sg .. 8 sg + 7 of its entry (consecutive entries: coalesced) and
// writes them as 64 contiguous bytes, so each entry's row is written by one
// workgroup in full cache lines. Workgroup 0 also clears the TMAX_SLOTS maxima
// of the class-1 t entries.
template <int MUC>
__global__ void __launch_bounds__(WG) k_scatter_xc32w(const double *x, const int *iperm_s, int nc, int mu_base,
                                                      float *zc, unsigned *tmax) {
    static_assert(TMAX_SLOTS == WG, "one slot per thread of workgroup 0");
    constexpr int E = WG / (MUC / 8); // entries per workgroup
    if (blockIdx.x == 0)
        tmax[threadIdx.x] = 0u;
    const int i = blockIdx.x * E + (int)threadIdx.x % E, s0 = 8 * ((int)threadIdx.x / E);
    if (i >= nc)
        return;
    float4 *o = reinterpret_cast<float4 *>(zc + ((size_t)iperm_s[i] * MUC + s0) * 2);
#pragma unroll
    for (int u = 0; u < 8; u += 2) {
        const double2 a = *reinterpret_cast<const double2 *>(x + 2 * (i + (size_t)nc * (mu_base + s0 + u)));
        const double2 c = *reinterpret_cast<const double2 *>(x + 2 * (i + (size_t)nc * (mu_base + s0 + u + 1)));
        o[u / 2] = make_float4((float)a.x, (float)a.y, (float)c.x, (float)c.y);
    }
}

template <typename TS, int MUC, bool ACC32>
__global__ void __launch_bounds__(WG) k_vx_rrc32(const VTile *items, const TS *varena0, const TS *varena1,
                                                 float *zc, const float *xmax, unsigned *tmax, double *part,
                                                 const double *tscale) {
    constexpr int CH = rrc_chunk<MUC>(), QP = MUC / 2, NH = MUC / 16;
    constexpr bool H16 = Mma16<TS>::on;
    constexpr int BUF = rrc32_stage_bytes<TS, MUC>();
    __shared__ __align__(16) unsigned char lds[2 * BUF];
    const VTile d = items[blockIdx.x];
    if (d.k == 0) // padding entry of an XCD-ordered list
        return;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 15, g = lane >> 4;
    const int K = d.k, nchunk = (d.ncols + CH - 1) / CH;
    const TS *varena = d.cls == 0 ? varena0 : varena1;
    double accd[NH][2][4] = {};
    constexpr int NS = H16 ? (MUC == 32 ? 2 : 1) : 4, SC = H16 ? 32 : 8, LPS = H16 ? 2 : 1; // loads per step
    const int nstep = (d.ncols + SC - 1) / SC;
    // lane (r, g): row 16 w + r, columns (2 or 8) g .. of each step
    const TS *rp = varena + 2 * (d.v_off + (size_t)min(16 * w + r, K - 1) * d.ld) + (H16 ? 16 : 4) * g;
    auto load = [&](int u, RrcUnit &b) {
#pragma unroll
        for (int st = 0; st < NS; st++)
#pragma unroll
            for (int l = 0; l < LPS; l++)
                b.a[st * LPS + l] = nt_load16(rp + 2 * (size_t)min(u * NS + st, nstep - 1) * SC + 8 * l);
    };
    // FP16: one scale for the item (the largest over its chunks), so the unit sums share it
    float sc = 1.0f;
    if constexpr (H16 && Mma16<TS>::scaled) {
        const int c0 = d.x_off, c1 = d.x_off + d.ncols - 1;
        float m = 0.0f;
        for (int b = c0 >> 5; b <= (c1 >> 5); b++)
            m = fmaxf(m, xmax[b]);
        sc = pass_scale(m);
    }
    if constexpr (H16) {
        // thread (s, cg): right-hand side s of chunk columns 8 cg .. 8 cg + 7
        const float2 *z2 = reinterpret_cast<const float2 *>(zc);
        auto xload16 = [&](int j, RrcX16 &x) { rrc16_load<MUC>(x.v, z2, j, d.ncols, d.x_off); };
        auto xstore16 = [&](int j, const RrcX16 &x) { rrc16_store<TS, MUC>(lds + (j & 1) * BUF, x.v, sc); };
        rrc32_walk<TS, MUC, RrcX16, ACC32>(nchunk, 16 * w < K, lds, sc, load, xload16, xstore16, accd);
    } else {
        const float4 *xsrc = reinterpret_cast<const float4 *>(zc) + (size_t)d.x_off * QP;
        auto xload = [&](int j, RrcX32 &x) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int f = (int)threadIdx.x + WG * i;
                x.v[i] = xsrc[(size_t)min(j * CH + f / QP, d.ncols - 1) * QP + f % QP];
            }
        };
        auto xstore = [&](int j, const RrcX32 &x) { rrc32_store<MUC>(lds + (j & 1) * BUF, x.v, j, d.ncols); };
        rrc32_walk<TS, MUC, RrcX32, false>(nchunk, 16 * w < K, lds, sc, load, xload, xstore, accd);
    }
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; i++) { // lane (r, g): rows 4 g + i, RHS 16 h + r
        const int row = 16 * w + 4 * g + i;
        if (row < K) {
            const double f = (tscale && d.nparts == 1) ? tscale[d.t_index + row] : 1.0;
#pragma unroll
            for (int h = 0; h < NH; h++) {
                if (d.nparts > 1) {
                    reinterpret_cast<double2 *>(part)[(size_t)(d.part_off + row) * MUC + 16 * h + r] =
                        make_double2(accd[h][0][i], accd[h][1][i]);
                } else {
                    const float re = (float)(accd[h][0][i] * f), im = (float)(accd[h][1][i] * f);
                    reinterpret_cast<float2 *>(zc)[(size_t)(d.t_index + row) * MUC + 16 * h + r] = make_float2(re, im);
                    m = fmaxf(m, fmaxf(fabsf(re), fabsf(im)));
                }
            }
        }
    }
    if (d.cls == 1 && d.nparts == 1) {
        m = wave_max(m);
        if (lane == 0)
            atomicMax(tmax + (blockIdx.x & (TMAX_SLOTS - 1)), __float_as_uint(m));
    }
}

template <typename TS0, typename TS1, int MUC>
__global__ void __launch_bounds__(WG) k_rows_rrc32(const int *items, const int *rb_row0, RowClass c0,
                                                   const TS0 *arena0, RowClass c1, const TS1 *arena1, const float *zc,
                                                   const unsigned *tmax, const int *perm_t, int nr, int mu_base,
                                                   double *y) {
    constexpr int CH = rrc_chunk<MUC>(), QP = MUC / 2, NH = MUC / 16;
    constexpr int B0 = rrc32_stage_bytes<TS0, MUC>(), B1 = rrc32_stage_bytes<TS1, MUC>();
    __shared__ __align__(16) unsigned char lds[2 * (B0 > B1 ? B0 : B1)];
    const int item = items[blockIdx.x], rb = item & 0xffffff, roff = RR_ROWS * (item >> 24);
    if (rb == 0xffffff) // padding entry of an XCD-ordered list
        return;
    const int r0 = rb_row0[rb] + roff, R = min(RR_ROWS, rb_row0[rb + 1] - r0);
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 15, g = lane >> 4;
    const float4 *zf = reinterpret_cast<const float4 *>(zc);
    double accd[NH][2][4] = {};
    auto go = [&](auto tag, const RowClass &rc, const auto *arena, float sc) {
        using TS = decltype(tag);
        constexpr bool H16 = Mma16<TS>::on;
        constexpr int NS = H16 ? (MUC == 32 ? 2 : 1) : 4, SC = H16 ? 32 : 8, LPS = H16 ? 2 : 1;
        const int cb = rc.rb_col0[rb], C = rc.rb_col0[rb + 1] - cb;
        if (C == 0)
            return;
        const int Cs     = (C + rc.salign - 1) & ~(rc.salign - 1);
        const int nstep  = (C + SC - 1) / SC;
        const int nchunk = (C + CH - 1) / CH;
        const TS *rp = arena + 2 * (rc.rb_aoff[rb] + (size_t)(roff + min(16 * w + r, R - 1)) * Cs) + (H16 ? 16 : 4) * g;
        const int *ci = rc.colidx + cb;
        auto load = [&](int u, RrcUnit &b) {
#pragma unroll
            for (int st = 0; st < NS; st++)
#pragma unroll
                for (int l = 0; l < LPS; l++)
                    b.a[st * LPS + l] = nt_load16(rp + 2 * (size_t)min(u * NS + st, nstep - 1) * SC + 8 * l);
        };
        if constexpr (H16) {
            // thread (s, cg): right-hand side s of chunk columns 8 cg .. 8 cg + 7
            const float2 *z2 = reinterpret_cast<const float2 *>(zc);
            // column indices one chunk ahead (the walk reads chunks in order,
            // the last one again unused), as in rb_class32; chunk column
            // 8 cg + c is entry e[c]
            const int cg = MUC == 64 ? __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6) : (int)threadIdx.x / MUC;
            int e[8];
            auto eload = [&](int j) {
#pragma unroll
                for (int c = 0; c < 8; c++)
                    e[c] = ci[min(j * CH + 8 * cg + c, C - 1)];
            };
            eload(0);
            auto xload16 = [&](int j, RrcX16 &x) {
                rrc16_load<MUC>(x.v, z2, j, C, [&](int i) { return e[i & 7]; });
                eload(j + 1);
            };
            auto xstore16 = [&](int j, const RrcX16 &x) {
                rrc16_store<TS, MUC>(lds + (j & 1) * rrc32_stage_bytes<TS, MUC>(), x.v, sc);
            };
            rrc32_walk<TS, MUC, RrcX16, true>(nchunk, 16 * w < R, lds, sc, load, xload16, xstore16, accd);
        } else {
            int idx[4]; // column indices one chunk ahead, as above
            auto iload = [&](int j) {
#pragma unroll
                for (int i = 0; i < 4; i++)
                    idx[i] = ci[min(j * CH + ((int)threadIdx.x + WG * i) / QP, C - 1)];
            };
            iload(0);
            auto xload = [&](int j, RrcX32 &x) {
#pragma unroll
                for (int i = 0; i < 4; i++)
                    x.v[i] = zf[(size_t)idx[i] * QP + ((int)threadIdx.x + WG * i) % QP];
                iload(j + 1);
            };
            auto xstore = [&](int j, const RrcX32 &x) {
                rrc32_store<MUC>(lds + (j & 1) * rrc32_stage_bytes<TS, MUC>(), x.v, j, C);
            };
            rrc32_walk<TS, MUC, RrcX32, false>(nchunk, 16 * w < R, lds, sc, load, xload, xstore, accd);
        }
    };
    // the class-1 scale's slots (lane l: slots l, l + 64, ...) are requested
    // before the class-0 walk and reduced after it
    constexpr bool SC1 = Mma16<TS1>::on && Mma16<TS1>::scaled;
    unsigned tv[TMAX_SLOTS / 64] = {};
    if (SC1 && arena1)
#pragma unroll
        for (int i = 0; i < TMAX_SLOTS / 64; i++)
            tv[i] = tmax[(threadIdx.x & 63) + 64 * i];
    go(TS0{}, c0, arena0, 1.0f);
    if (arena1) {
        float sc = 1.0f;
        if constexpr (SC1) {
            float m = 0.0f;
#pragma unroll
            for (int i = 0; i < TMAX_SLOTS / 64; i++)
                m = fmaxf(m, __uint_as_float(tv[i]));
            sc = pass_scale(wave_max(m));
        }
        go(TS1{}, c1, arena1, sc);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int row = 16 * w + 4 * g + i;
        if (row < R) {
#pragma unroll
            for (int h = 0; h < NH; h++) {
                double *o = y + 2 * ((size_t)perm_t[r0 + row] + (size_t)nr * (mu_base + 16 * h + r));
                o[0]      = accd[h][0][i];
                o[1]      = accd[h][1][i];
            }
        }
    }
}

template <typename TS0, typename TS1, typename R, int MU>
__global__ void __launch_bounds__(WG) k_rows_c(const int *rb_order, const int *rb_row0, RowClass c0, const TS0 *arena0,
                                               RowClass c1, const TS1 *arena1, const R *z, const int *perm_t, int nr,
                                               int mu_base, double *y) {
    __shared__ R s_z[2 * ZCAP_C];
    __shared__ cpair<double> s_acc[ROWMAX * MU];
    const int rb = rb_order[blockIdx.x];
    const int r0 = rb_row0[rb], R_ = rb_row0[rb + 1] - r0;
    for (int e = threadIdx.x; e < R_ * MU; e += WG)
        s_acc[e] = cpair<double>{0, 0};
    __syncthreads();
    rows_class_c<TS0, R, MU>(arena0, c0, rb, R_, z, s_z, s_acc);
    if (arena1)
        rows_class_c<TS1, R, MU>(arena1, c1, rb, R_, z, s_z, s_acc);
    __syncthreads();
    for (int e = threadIdx.x; e < R_ * MU; e += WG) {
        const int r = e / MU, s = e % MU;
        double *o   = y + 2 * ((size_t)perm_t[r0 + r] + (size_t)nr * (mu_base + s));
        o[0]        = s_acc[e].re;
        o[1]        = s_acc[e].im;
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
        hipLaunchKernelGGL(k_aca<double>, dim3(nblocks), dim3(WG), 0, st, k, T, S, blocks, eps, reqrank, ws, vis_ws,
                           vis_off, ranks);
}

void launch_aca_c(const KernelDesc &k, DevPoints T, DevPoints S, const AcaBlock *blocks, int nblocks, double eps,
                  int reqrank, cdbl *ws, unsigned char *vis_ws, const uint64_t *vis_off, int *ranks, hipStream_t st) {
    if (nblocks > 0)
        hipLaunchKernelGGL(k_aca<cdbl>, dim3(nblocks), dim3(WG), 0, st, k, T, S, blocks, eps, reqrank, ws, vis_ws,
                           vis_off, ranks);
}

void launch_acaw_init(const KernelDesc &k, DevPoints T, DevPoints S, const AcaBlock *blocks, AcaWide *st,
                      int nblocks, int reqrank, double eps, unsigned char *vis_ws, const uint64_t *vis_off,
                      hipStream_t s) {
    (void)k;
    if (nblocks > 0)
        hipLaunchKernelGGL(k_acaw_init, dim3(nblocks), dim3(WG), 0, s, blocks, st, reqrank, eps, T, S, vis_ws,
                           vis_off);
}

template <typename E>
static void acaw_term(const KernelDesc &k, DevPoints T, DevPoints S, const AcaBlock *blocks, AcaWide *st,
                      int nblocks, const AcawTask *tasks1, int ntasks1, const AcawTask *tasks2, int ntasks2,
                      int reqrank, double eps, E *ws, unsigned char *vis_ws, const uint64_t *vis_off, double *pv,
                      int *pi, double *dd, E *xs, hipStream_t s) {
    hipLaunchKernelGGL((k_acaw_vec<E, true>), dim3(ntasks2), dim3(WG), 0, s, k, T, S, blocks, st, tasks2, 0, ws,
                       vis_ws, vis_off, pv, pi, dd, xs, (const E *)nullptr, (const uint64_t *)nullptr);
    hipLaunchKernelGGL(k_acaw_piv<E>, dim3((nblocks + WG - 1) / WG), dim3(WG), 0, s, blocks, st, nblocks, ws, vis_ws,
                       vis_off, pv, pi);
    hipLaunchKernelGGL((k_acaw_vec<E, false>), dim3(ntasks1), dim3(WG), 0, s, k, T, S, blocks, st, tasks1, ntasks2, ws,
                       vis_ws, vis_off, pv, pi, dd, xs, (const E *)nullptr, (const uint64_t *)nullptr);
    hipLaunchKernelGGL(k_acaw_fin<E>, dim3(nblocks), dim3(WG), 0, s, blocks, st, reqrank, eps, vis_ws, vis_off, pv, pi,
                       dd, xs);
}

// The two halves of a cross term with generator values from the host (a user
// generator): the residual row and pivot column from gv, then the scaled
// residual column, next pivot row and stopping test from gv.
template <typename E>
static void acaw_half(bool row, DevPoints T, DevPoints S, const AcaBlock *blocks, AcaWide *st, int nblocks,
                      const AcawTask *tasks, int ntasks, int slot0, int reqrank, double eps, E *ws,
                      unsigned char *vis_ws, const uint64_t *vis_off, double *pv, int *pi, double *dd, E *xs,
                      const E *gv, const uint64_t *gv_off, hipStream_t s) {
    const KernelDesc k{};
    if (row) {
        hipLaunchKernelGGL((k_acaw_vec<E, true>), dim3(ntasks), dim3(WG), 0, s, k, T, S, blocks, st, tasks, slot0, ws,
                           vis_ws, vis_off, pv, pi, dd, xs, gv, gv_off);
        hipLaunchKernelGGL(k_acaw_piv<E>, dim3((nblocks + WG - 1) / WG), dim3(WG), 0, s, blocks, st, nblocks, ws,
                           vis_ws, vis_off, pv, pi);
    } else {
        hipLaunchKernelGGL((k_acaw_vec<E, false>), dim3(ntasks), dim3(WG), 0, s, k, T, S, blocks, st, tasks, slot0,
                           ws, vis_ws, vis_off, pv, pi, dd, xs, gv, gv_off);
        hipLaunchKernelGGL(k_acaw_fin<E>, dim3(nblocks), dim3(WG), 0, s, blocks, st, reqrank, eps, vis_ws, vis_off, pv,
                           pi, dd, xs);
    }
}

void launch_acaw_half(bool row, DevPoints T, DevPoints S, const AcaBlock *blocks, AcaWide *st, int nblocks,
                      const AcawTask *tasks, int ntasks, int slot0, int reqrank, double eps, void *ws, bool cplx,
                      unsigned char *vis_ws, const uint64_t *vis_off, double *pv, int *pi, double *dd, void *xs,
                      const void *gv, const uint64_t *gv_off, hipStream_t s) {
    if (nblocks <= 0)
        return;
    if (cplx)
        acaw_half<cdbl>(row, T, S, blocks, st, nblocks, tasks, ntasks, slot0, reqrank, eps, static_cast<cdbl *>(ws),
                        vis_ws, vis_off, pv, pi, dd, static_cast<cdbl *>(xs), static_cast<const cdbl *>(gv), gv_off, s);
    else
        acaw_half<double>(row, T, S, blocks, st, nblocks, tasks, ntasks, slot0, reqrank, eps,
                          static_cast<double *>(ws), vis_ws, vis_off, pv, pi, dd, static_cast<double *>(xs),
                          static_cast<const double *>(gv), gv_off, s);
}

void launch_acaw_term(const KernelDesc &k, DevPoints T, DevPoints S, const AcaBlock *blocks, AcaWide *st,
                      int nblocks, const AcawTask *tasks1, int ntasks1, const AcawTask *tasks2, int ntasks2,
                      int reqrank, double eps, void *ws, bool cplx, unsigned char *vis_ws, const uint64_t *vis_off,
                      double *pv, int *pi, double *dd, void *xs, hipStream_t s) {
    if (nblocks <= 0)
        return;
    if (cplx)
        acaw_term<cdbl>(k, T, S, blocks, st, nblocks, tasks1, ntasks1, tasks2, ntasks2, reqrank, eps,
                        static_cast<cdbl *>(ws), vis_ws, vis_off, pv, pi, dd, static_cast<cdbl *>(xs), s);
    else
        acaw_term<double>(k, T, S, blocks, st, nblocks, tasks1, ntasks1, tasks2, ntasks2, reqrank, eps,
                          static_cast<double *>(ws), vis_ws, vis_off, pv, pi, dd, static_cast<double *>(xs), s);
}

void launch_lr_pack(const LrPack *packs, int npacks, const UTileDst *utiles, const double *ws, double *arena,
                    double *varena, hipStream_t st) {
    if (npacks > 0)
        hipLaunchKernelGGL(k_lr_pack<double>, dim3(npacks), dim3(WG), 0, st, packs, utiles, ws, arena, varena);
}

void launch_lr_pack_c(const LrPack *packs, int npacks, const UTileDst *utiles, const cdbl *ws, cdbl *arena,
                      cdbl *varena, hipStream_t st) {
    if (npacks > 0)
        hipLaunchKernelGGL(k_lr_pack<cdbl>, dim3(npacks), dim3(WG), 0, st, packs, utiles, ws, arena, varena);
}

void launch_dense_fill_c(const KernelDesc &k, DevPoints T, DevPoints S, const FillTile *tiles, int ntiles, cdbl *arena,
                         hipStream_t st) {
    if (ntiles > 0)
        hipLaunchKernelGGL(k_dense_fill_c, dim3(ntiles), dim3(WG), 0, st, k, T, S, tiles, arena);
}

void launch_term_scale(const ScaleLeaf *leaves, int nleaves, double *store, double *sig, hipStream_t st) {
    if (nleaves > 0)
        hipLaunchKernelGGL(k_term_scale<double>, dim3(nleaves), dim3(WG), 0, st, leaves, store, sig);
}

void launch_term_scale_c(const ScaleLeaf *leaves, int nleaves, cdbl *store, double *sig, hipStream_t st) {
    if (nleaves > 0)
        hipLaunchKernelGGL(k_term_scale<cdbl>, dim3(nleaves), dim3(WG), 0, st, leaves, store, sig);
}

void launch_recompress(const RcLeaf *leaves, int nleaves, double *store, double tol, int *rank, int *ok,
                       double *sigma, hipStream_t st) {
    if (nleaves > 0)
        hipLaunchKernelGGL(k_recompress, dim3(nleaves), dim3(WG), 0, st, leaves, store, tol, rank, ok, sigma);
}

void launch_recompress_c(const RcLeaf *leaves, int nleaves, cdbl *store, double tol, int *rank, int *ok,
                         double *sigma, hipStream_t st) {
    if (nleaves > 0)
        hipLaunchKernelGGL(k_recompress_c, dim3(nleaves), dim3(WG), 0, st, leaves, store, tol, rank, ok, sigma);
}

// Layout changes between the cluster ordering with interleaved right-hand sides
// (entry (j, s) at j * mu + s, the layout of mymvprod_local_to_local,
// hmatrix.hpp:1117-1131) and the original ordering, column-major (entry
// (perm[j], s) at perm[j] + n * s); e doubles per entry.
__global__ void k_to_original(const double *in, const int *perm, int n, int mu, int e, double *out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * mu)
        return;
    const int j = (int)(t / mu), s = (int)(t % mu);
    for (int u = 0; u < e; u++)
        out[e * ((size_t)perm[j] + (size_t)n * s) + u] = in[e * t + u];
}
__global__ void k_to_cluster(const double *in, const int *perm, int n, int mu, int e, double *out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * mu)
        return;
    const int j = (int)(t / mu), s = (int)(t % mu);
    for (int u = 0; u < e; u++)
        out[e * t + u] = in[e * ((size_t)perm[j] + (size_t)n * s) + u];
}
void launch_to_original(const double *in, const int *perm, int n, int mu, int e, double *out, hipStream_t st) {
    const size_t tot = (size_t)n * mu;
    if (tot)
        hipLaunchKernelGGL(k_to_original, dim3((unsigned)((tot + WG - 1) / WG)), dim3(WG), 0, st, in, perm, n, mu, e,
                           out);
}
void launch_to_cluster(const double *in, const int *perm, int n, int mu, int e, double *out, hipStream_t st) {
    const size_t tot = (size_t)n * mu;
    if (tot)
        hipLaunchKernelGGL(k_to_cluster, dim3((unsigned)((tot + WG - 1) / WG)), dim3(WG), 0, st, in, perm, n, mu, e,
                           out);
}

template <typename TS>
void launch_convert(const double *src, TS *dst, uint64_t n, hipStream_t st) {
    if (n == 0)
        return;
    uint64_t blocks = (n + WG - 1) / WG;
    if (blocks > 65536)
        blocks = 65536;
    hipLaunchKernelGGL(k_convert<TS>, dim3((unsigned)blocks), dim3(WG), 0, st, src, dst, n);
}
template void launch_convert<float>(const double *, float *, uint64_t, hipStream_t);
template void launch_convert<__half>(const double *, __half *, uint64_t, hipStream_t);
template void launch_convert<__hip_bfloat16>(const double *, __hip_bfloat16 *, uint64_t, hipStream_t);

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

template <typename TC>
void launch_gather_x(const double *x, const int *perm_s, int nc, int tpad, int mu_base, int MU, TC *z,
                     hipStream_t st) {
    const int n = tpad * MU;
    const dim3 grid((n + WG - 1) / WG);
    if constexpr (std::is_same<TC, float>::value) {
        if (MU == MFMA_MU) {
            hipLaunchKernelGGL((k_gather_x<TC, MFMA_MU, true>), grid, dim3(WG), 0, st, x, perm_s, nc, tpad, mu_base,
                               z);
            return;
        }
    }
#define HTG_GATHER(M) \
    hipLaunchKernelGGL((k_gather_x<TC, M, false>), grid, dim3(WG), 0, st, x, perm_s, nc, tpad, mu_base, z)
    HTG_MU_SWITCH(MU, HTG_GATHER)
#undef HTG_GATHER
}

void launch_gather_x64(const double *x, const int *iperm_s, int nc, int tpad, int mu_base, double *z, hipStream_t st) {
    hipLaunchKernelGGL(k_scatter_x64, dim3((tpad + WG - 1) / WG, MFMA_MU / 8), dim3(WG), 0, st, x, iperm_s, nc, tpad,
                       mu_base, z);
}

void launch_gather_x32(const double *x, const int *iperm_s, int nc, int tpad, int mu_base, float *stage, float *z,
                       float *xmax, unsigned *tmax, hipStream_t st) {
    hipLaunchKernelGGL(k_scatter_x32, dim3((nc + WG - 1) / WG, MFMA_MU / 8), dim3(WG), 0, st, x, iperm_s, nc, mu_base,
                       stage);
    hipLaunchKernelGGL(k_block_x32, dim3((tpad * 8 + WG - 1) / WG), dim3(WG), 0, st, stage, nc, tpad, z, xmax, tmax);
}

template <typename TS0, typename TS1, typename TC>
void launch_vx(const int *vg_off, int ngroups, const VTile *tiles, const VTile *multi_tiles, int nmulti,
               const TS0 *varena0, const TS1 *varena1, TC *z, double *part, const double *tscale, int MU,
               hipStream_t st) {
    if (ngroups <= 0)
        return;
#define HTG_VX(M)                                                                                          \
    do {                                                                                                   \
        hipLaunchKernelGGL((k_vx<TS0, TS1, TC, M>), dim3((ngroups + NWAVE - 1) / NWAVE), dim3(WG), 0, st, vg_off, \
                           ngroups, tiles, varena0, varena1, z, part, tscale);                             \
        if (nmulti > 0)                                                                                    \
            hipLaunchKernelGGL((k_vx_sum<TC, M>), dim3((nmulti + vx_sum_groups_per_block<M>() - 1) /         \
                                                     vx_sum_groups_per_block<M>()),                         \
                               dim3(WG), 0, st, multi_tiles, nmulti, part, z, tscale);              \
    } while (0)
    HTG_MU_SWITCH(MU, HTG_VX)
#undef HTG_VX
}

template <typename TS0, typename TS1, typename TC>
void launch_vx_rr(const VTile *items, int nitems, const VTile *multi_tiles, int nmulti, const TS0 *varena0,
                  const TS1 *varena1, TC *z, float *zil, const float *xmax, unsigned *tmax, double *part,
                  const double *tscale, hipStream_t st) {
    if (nitems <= 0)
        return;
    if constexpr (std::is_same<TC, double>::value) {
        static_assert(std::is_same<TS0, double>::value && std::is_same<TS1, double>::value, "FP64 items");
        hipLaunchKernelGGL(k_vx_rr64, dim3(nitems), dim3(WG), 0, st, items, varena0, varena1, z, part, tscale);
        if (nmulti > 0) // a range's rows x 32 sums (<= 4096) over four workgroups
            hipLaunchKernelGGL((k_vx_sum<TC, MFMA_MU>), dim3(nmulti, 4), dim3(WG), 0, st, multi_tiles, nmulti, part,
                               z, tscale);
    } else {
        hipLaunchKernelGGL((k_vx_rr<TS0, TS1>), dim3(nitems), dim3(WG), 0, st, items, varena0, varena1,
                           (const float *)z, zil, xmax, tmax, part, tscale);
        if (nmulti > 0)
            hipLaunchKernelGGL((k_vx_sum<float, MFMA_MU, false>), dim3(nmulti, 4), dim3(WG), 0, st, multi_tiles,
                               nmulti, part, zil, tscale, tmax);
    }
}

template <typename TS0, typename TS1, typename TC>
void launch_rows_rr(const int *items, int nitems, const int *rb_row0, RowClass c0, const TS0 *arena0, RowClass c1,
                    const TS1 *arena1, const TC *z, const float *zil, const unsigned *tmax, const int *perm_t, int nr,
                    int mu_base, double *y, hipStream_t st) {
    if (nitems <= 0)
        return;
    if constexpr (std::is_same<TC, double>::value) {
        static_assert(std::is_same<TS0, double>::value && std::is_same<TS1, double>::value, "FP64 parts");
        hipLaunchKernelGGL(k_rows_rr64, dim3(nitems), dim3(WG), 0, st, items, rb_row0, c0, arena0, c1, arena1, z,
                           perm_t, nr, mu_base, y);
    } else {
        hipLaunchKernelGGL((k_rows_rr<TS0, TS1>), dim3(nitems), dim3(WG), 0, st, items, rb_row0, c0, arena0, c1,
                           arena1, zil, tmax, perm_t, nr, mu_base, y);
    }
}

template <typename TS0, typename TS1, typename TC>
void launch_rows(const int *rb_order, const int *rb_row0, int nrb, RowClass c0, const TS0 *arena0, RowClass c1,
                 const TS1 *arena1, const TC *z, const int *perm_t, int nr, int mu_base, int MU, double *y,
                 hipStream_t st) {
    if (nrb <= 0)
        return;
#define HTG_ROWS(M)                                                                                          \
    hipLaunchKernelGGL((k_rows<TS0, TS1, TC, M>), dim3(nrb), dim3(WG), 0, st, rb_order, rb_row0, c0, arena0, c1, \
                       arena1, z, perm_t, nr, mu_base, y)
    HTG_MU_SWITCH(MU, HTG_ROWS)
#undef HTG_ROWS
}

#define HTG_MU_SWITCH_C(MUV, CALL) \
    switch (MUV) {                 \
    case 1: CALL(1); break;        \
    case 2: CALL(2); break;        \
    case 4: CALL(4); break;        \
    case 8: CALL(8); break;        \
    default: CALL(16); break;      \
    }

template <typename R>
void launch_gather_xc(const double *x, const int *perm_s, int nc, int tpad, int mu_base, int MU, R *z,
                      hipStream_t st) {
    const dim3 grid((tpad * MU + WG - 1) / WG);
    if constexpr (std::is_same<R, float>::value) {
        if (MU == MFMA_MU_C) {
            hipLaunchKernelGGL(k_gather_xc_blk, grid, dim3(WG), 0, st, x, perm_s, nc, tpad, mu_base, z);
            return;
        }
    }
#define HTG_GATHER_C(M) hipLaunchKernelGGL((k_gather_xc<R, M>), grid, dim3(WG), 0, st, x, perm_s, nc, tpad, mu_base, z)
    HTG_MU_SWITCH_C(MU, HTG_GATHER_C)
#undef HTG_GATHER_C
}

void launch_gather_xc32(const double *x, const int *iperm_s, int nc, int tpad, int mu_base, float *stage, float *z,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_scatter_xc32, dim3((nc + WG - 1) / WG, MFMA_MU_C / 8), dim3(WG), 0, st, x, iperm_s, nc,
                       mu_base, stage);
    hipLaunchKernelGGL(k_block_x32, dim3((tpad * 8 + WG - 1) / WG), dim3(WG), 0, st, stage, nc, tpad, z, nullptr,
                       nullptr);
}

void launch_gather_xc64(const double *x, const int *iperm_s, int nc, int tpad, int mu_base, int MU, double *z,
                        hipStream_t st) {
    if (MU == 32)
        hipLaunchKernelGGL(k_scatter_xc64<32>, dim3((tpad + WG / 4 - 1) / (WG / 4)), dim3(WG), 0, st, x, iperm_s, nc, tpad,
                           mu_base, z);
    else
        hipLaunchKernelGGL(k_scatter_xc64<16>, dim3((tpad + WG / 2 - 1) / (WG / 2)), dim3(WG), 0, st, x, iperm_s, nc, tpad,
                           mu_base, z);
}

void launch_rrc64(const VTile *items, int nitems, const VTile *multi_tiles, int nmulti, const int *rb_items,
                  int nrb_items, const int *rb_row0, RowClass c0, const double *arena0, RowClass c1,
                  const double *arena1, const double *varena0, const double *varena1, double *z, double *part,
                  const double *tscale, const int *perm_t, int nr, int mu_base, double *y, hipStream_t st,
                  hipEvent_t ev_vx_done) {
    if (nitems > 0) {
        hipLaunchKernelGGL(k_vx_rrc64, dim3(nitems), dim3(WG), 0, st, items, varena0, varena1, z, part, tscale);
        if (nmulti > 0) // a range's rows x 32 sums (<= 2048) over four workgroups
            hipLaunchKernelGGL((k_vx_sum_c<double, 32>), dim3(nmulti, 4), dim3(WG), 0, st, multi_tiles, nmulti, part,
                               z, tscale);
    }
    if (ev_vx_done)
        (void)hipEventRecord(ev_vx_done, st);
    if (nrb_items > 0)
        hipLaunchKernelGGL(k_rows_rrc64, dim3(nrb_items), dim3(WG), 0, st, rb_items, rb_row0, c0, arena0, c1, arena1,
                           z, perm_t, nr, mu_base, y);
}

template <typename TS0, typename TS1>
void launch_rrc32(int muc, const double *x, const int *iperm_s, int nc, int tpad, int mu_base, float *zc,
                  float *xmax, unsigned *tmax, const VTile *items, int nitems, int nitems0, const VTile *multi_tiles,
                  int nmulti,
                  const int *rb_items, int nrb_items, const int *rb_row0, RowClass c0, const TS0 *arena0, RowClass c1,
                  const TS1 *arena1, const TS0 *varena0, const TS1 *varena1, double *part, const double *tscale,
                  const int *perm_t, int nr, double *y, hipStream_t st, hipEvent_t ev_gather, hipEvent_t ev_vx) {
    auto run = [&](auto muc_tag) {
        constexpr int MUC = decltype(muc_tag)::value;
        hipLaunchKernelGGL(k_scatter_xc32w<MUC>, dim3((nc + WG / (MUC / 8) - 1) / (WG / (MUC / 8))), dim3(WG), 0, st,
                           x, iperm_s, nc, mu_base, zc, tmax);
        if (Mma16<TS0>::scaled || Mma16<TS1>::scaled)
            hipLaunchKernelGGL(k_colmax_c32<MUC>, dim3((tpad + 31) / 32), dim3(WG), 0, st, zc, nc, xmax);
        if (ev_gather)
            (void)hipEventRecord(ev_gather, st);
        if (nitems > 0) {
            if constexpr (std::is_same<TS0, TS1>::value) {
                hipLaunchKernelGGL((k_vx_rrc32<TS0, MUC, false>), dim3(nitems), dim3(WG), 0, st, items, varena0,
                                   varena1, zc, xmax, tmax, part, tscale);
            } else { // items [0, nitems0) are class 0, the rest class 1
                if (nitems0 > 0)
                    hipLaunchKernelGGL((k_vx_rrc32<TS0, MUC, false>), dim3(nitems0), dim3(WG), 0, st, items, varena0,
                                       varena0, zc, xmax, tmax, part, tscale);
                if (nitems > nitems0)
                    hipLaunchKernelGGL((k_vx_rrc32<TS1, MUC, true>), dim3(nitems - nitems0), dim3(WG), 0, st,
                                       items + nitems0, varena1, varena1, zc, xmax, tmax, part, tscale);
            }
            if (nmulti > 0)
                hipLaunchKernelGGL((k_vx_sum_c<float, MUC>), dim3(nmulti, 4), dim3(WG), 0, st, multi_tiles, nmulti,
                                   part, zc, tscale, tmax);
        }
        if (ev_vx)
            (void)hipEventRecord(ev_vx, st);
        if (nrb_items > 0)
            hipLaunchKernelGGL((k_rows_rrc32<TS0, TS1, MUC>), dim3(nrb_items), dim3(WG), 0, st, rb_items, rb_row0, c0,
                               arena0, c1, arena1, zc, tmax, perm_t, nr, mu_base, y);
    };
    if (muc == 64)
        run(std::integral_constant<int, 64>{});
    else
        run(std::integral_constant<int, 32>{});
}
#define HTG_RRC32_INST(TS0, TS1)                                                                                   \
    template void launch_rrc32<TS0, TS1>(int, const double *, const int *, int, int, int, float *, float *,         \
                                         unsigned *, const VTile *, int, int, const VTile *, int, const int *, int, \
                                         const int *, RowClass, const TS0 *, RowClass, const TS1 *, const TS0 *,    \
                                         const TS1 *, double *, const double *, const int *, int, double *,         \
                                         hipStream_t, hipEvent_t, hipEvent_t);
HTG_RRC32_INST(float, float)
HTG_RRC32_INST(float, __half)
HTG_RRC32_INST(float, __hip_bfloat16)
#undef HTG_RRC32_INST

template <typename TS0, typename TS1, typename R>
void launch_vx_c(const int *vg_off, int ngroups, const VTile *tiles, const VTile *multi_tiles, int nmulti,
                 const TS0 *varena0, const TS1 *varena1, R *z, double *part, const double *tscale, int MU,
                 hipStream_t st) {
    if (ngroups <= 0)
        return;
    if constexpr (std::is_same<R, double>::value && std::is_same<TS0, double>::value &&
                  std::is_same<TS1, double>::value) {
        if (MU == MU_C64) {
            hipLaunchKernelGGL(k_vx_mfma64c16, dim3((ngroups + NWAVE - 1) / NWAVE), dim3(WG), 0, st, vg_off, ngroups,
                               tiles, varena0, varena1, z, part, tscale);
            if (nmulti > 0)
                hipLaunchKernelGGL((k_vx_sum_c<R, 16>), dim3(nmulti), dim3(WG), 0, st, multi_tiles, nmulti, part, z,
                                   tscale);
            return;
        }
    }
    if constexpr (std::is_same<R, float>::value) {
        if (MU == MFMA_MU_C) {
            hipLaunchKernelGGL((k_vx_mfma_c<TS0, TS1>), dim3((ngroups + NWAVE - 1) / NWAVE), dim3(WG), 0, st, vg_off,
                               ngroups, tiles, varena0, varena1, z, part, tscale);
            if (nmulti > 0) // slots sum independently: the real 32-slot reduction applies
                hipLaunchKernelGGL((k_vx_sum<float, 32, true>), dim3(nmulti), dim3(WG), 0, st, multi_tiles,
                                   nmulti, part, z, tscale);
            return;
        }
    }
#define HTG_VXC(M)                                                                                                  \
    do {                                                                                                            \
        hipLaunchKernelGGL((k_vx_c<TS0, TS1, R, M>), dim3((ngroups + NWAVE - 1) / NWAVE), dim3(WG), 0, st, vg_off, \
                           ngroups, tiles, varena0, varena1, z, part, tscale);                                      \
        if (nmulti > 0)                                                                                             \
            hipLaunchKernelGGL((k_vx_sum_c<R, M>), dim3((nmulti + vx_sum_groups_per_block<M>() - 1) /            \
                                                       vx_sum_groups_per_block<M>()),                            \
                               dim3(WG), 0, st, multi_tiles, nmulti, part, z, tscale);                       \
    } while (0)
    HTG_MU_SWITCH_C(MU, HTG_VXC)
#undef HTG_VXC
}

template <typename TS0, typename TS1, typename R>
void launch_rows_c(const int *rb_order, const int *rb_row0, int nrb, RowClass c0, const TS0 *arena0, RowClass c1,
                   const TS1 *arena1, const R *z, const int *perm_t, int nr, int mu_base, int MU, double *y,
                   hipStream_t st) {
    if (nrb <= 0)
        return;
    if constexpr (std::is_same<R, double>::value && std::is_same<TS0, double>::value &&
                  std::is_same<TS1, double>::value) {
        if (MU == MU_C64) {
            hipLaunchKernelGGL(k_rows_mfma64c16, dim3(nrb), dim3(WG), 0, st, rb_order, rb_row0, c0, arena0, c1,
                               arena1, z, perm_t, nr, mu_base, y);
            return;
        }
    }
    if constexpr (std::is_same<R, float>::value) {
        if (MU == MFMA_MU_C) {
            hipLaunchKernelGGL((k_rows_mfma_c<TS0, TS1>), dim3(nrb), dim3(WG), 0, st, rb_order, rb_row0, c0, arena0,
                               c1, arena1, z, perm_t, nr, mu_base, y);
            return;
        }
    }
#define HTG_ROWSC(M)                                                                                          \
    hipLaunchKernelGGL((k_rows_c<TS0, TS1, R, M>), dim3(nrb), dim3(WG), 0, st, rb_order, rb_row0, c0, arena0, c1, \
                       arena1, z, perm_t, nr, mu_base, y)
    HTG_MU_SWITCH_C(MU, HTG_ROWSC)
#undef HTG_ROWSC
}

template void launch_gather_xc<double>(const double *, const int *, int, int, int, int, double *, hipStream_t);
template void launch_gather_xc<float>(const double *, const int *, int, int, int, int, float *, hipStream_t);
#define HTG_INST_C(TS0, TS1, R)                                                                                       \
    template void launch_vx_c<TS0, TS1, R>(const int *, int, const VTile *, const VTile *, int, const TS0 *,          \
                                           const TS1 *, R *, double *, const double *, int, hipStream_t);           \
    template void launch_rows_c<TS0, TS1, R>(const int *, const int *, int, RowClass, const TS0 *, RowClass,        \
                                             const TS1 *, const R *, const int *, int, int, int, double *, hipStream_t);
HTG_INST_C(double, double, double)
HTG_INST_C(float, float, float)
HTG_INST_C(float, __half, float)
HTG_INST_C(float, __hip_bfloat16, float)
#undef HTG_INST_C

template void launch_gather_x<double>(const double *, const int *, int, int, int, int, double *, hipStream_t);
template void launch_gather_x<float>(const double *, const int *, int, int, int, int, float *, hipStream_t);
template void launch_vx<double, double, double>(const int *, int, const VTile *, const VTile *, int, const double *,
                                                const double *, double *, double *, const double *, int, hipStream_t);
template void launch_vx<float, float, float>(const int *, int, const VTile *, const VTile *, int, const float *,
                                             const float *, float *, double *, const double *, int, hipStream_t);
template void launch_vx<float, __half, float>(const int *, int, const VTile *, const VTile *, int, const float *,
                                              const __half *, float *, double *, const double *, int, hipStream_t);
template void launch_vx<float, __hip_bfloat16, float>(const int *, int, const VTile *, const VTile *, int,
                                                      const float *, const __hip_bfloat16 *, float *, double *,
                                                      const double *, int, hipStream_t);
#define HTG_RR_INST(TS0, TS1, TC)                                                                                \
    template void launch_vx_rr<TS0, TS1, TC>(const VTile *, int, const VTile *, int, const TS0 *, const TS1 *, TC *, \
                                             float *, const float *, unsigned *, double *, const double *,           \
                                             hipStream_t);                                                            \
    template void launch_rows_rr<TS0, TS1, TC>(const int *, int, const int *, RowClass, const TS0 *, RowClass,      \
                                               const TS1 *, const TC *, const float *, const unsigned *, const int *, \
                                               int, int, double *, hipStream_t);
HTG_RR_INST(double, double, double)
HTG_RR_INST(float, float, float)
HTG_RR_INST(float, __half, float)
HTG_RR_INST(float, __hip_bfloat16, float)
#undef HTG_RR_INST
template void launch_rows<double, double, double>(const int *, const int *, int, RowClass, const double *, RowClass,
                                                  const double *, const double *, const int *, int, int, int,
                                                  double *, hipStream_t);
template void launch_rows<float, float, float>(const int *, const int *, int, RowClass, const float *, RowClass,
                                               const float *, const float *, const int *, int, int, int, double *,
                                               hipStream_t);
template void launch_rows<float, __half, float>(const int *, const int *, int, RowClass, const float *, RowClass,
                                                const __half *, const float *, const int *, int, int, int, double *,
                                                hipStream_t);
template void launch_rows<float, __hip_bfloat16, float>(const int *, const int *, int, RowClass, const float *,
                                                        RowClass, const __hip_bfloat16 *, const float *,
                                                        const int *, int, int, int, double *, hipStream_t);

} // namespace htg
