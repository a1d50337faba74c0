// Batched long dot products: R[a + b ldr] = sum_k A[k + a lda] B[k + b ldb]
// for a < na, b < nb over k < K, FP64.
//
// The QML contractions <D_b, M_q> (pipeline.qml: q_b + nb_b = 1/2 y^T dC_b y
// and 1/2 Tr[C^-1 N C^-1 dC_b], spectra.py:1124-1186) have K = ma mb up to
// 2.6e8, na = bins and nb = simulations + 1: a GEMM would compute tiles
// padded to 96 x 96 of which ~1% is wanted.  This is an HBM-bound pass
// instead: a block takes a 4096-element chunk of K, holds B's chunk (up to
// four vectors) in registers, streams every A vector's chunk once
// (coalesced 8-byte loads) and reduces each of the na x nb products inside
// its waves; the per-wave partials are summed in a fixed order by a second
// kernel, so results are deterministic.
#include "cfmi_internal.h"

namespace {

constexpr int CT = 256;          // threads per block
constexpr int CJ = 16;           // elements of K per thread
constexpr int CCH = CT * CJ;     // K elements per block
constexpr int CMAXB = 4;         // B vectors per pass

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int NB>
__global__ __launch_bounds__(CT) void dot_partial_kernel(int64_t K, int64_t na, const double* __restrict__ A,
                                                         int64_t lda, const double* __restrict__ B, int64_t ldb,
                                                         double* __restrict__ part)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t k0 = (int64_t)blockIdx.x * CCH;
    double bv[CJ][NB];
#pragma unroll
    for (int j = 0; j < CJ; ++j) {
        const int64_t k = k0 + tid + (int64_t)j * CT;
#pragma unroll
        for (int q = 0; q < NB; ++q) bv[j][q] = k < K ? B[k + q * ldb] : 0.0;
    }
    // part layout: [block][wave][a][q]
    double* out = part + ((int64_t)blockIdx.x * (CT / 64) + wave) * na * NB;
    for (int64_t a = 0; a < na; ++a) {
        const double* Aa = A + a * lda;
        double s[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) s[q] = 0.0;
#pragma unroll
        for (int j = 0; j < CJ; ++j) {
            const int64_t k = k0 + tid + (int64_t)j * CT;
            const double x = k < K ? Aa[k] : 0.0;
#pragma unroll
            for (int q = 0; q < NB; ++q) s[q] = fma(x, bv[j][q], s[q]);
        }
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const double t = wave_sum(s[q]);
            if (lane == 0) out[a * NB + q] = t;
        }
    }
}

// One block per output: sum of its nparts partials (stride na nb) in a
// fixed order (thread-strided, then a fixed tree).
__global__ __launch_bounds__(CT) void dot_reduce_kernel(int64_t nparts, int64_t na, int nb, const double* part,
                                                        double* R, int64_t ldr, int b0)
{
    const int64_t o = blockIdx.x;          // a * nb + q
    const int64_t a = o / nb, q = o % nb;
    double s = 0.0;
    for (int64_t p = threadIdx.x; p < nparts; p += CT) s += part[p * na * nb + o];
    __shared__ double red[CT / 64];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) R[a + (b0 + q) * ldr] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Off-diagonal nonzeros of an n x n column-major matrix: column j = blockIdx.y,
// rows from blockIdx.x; one atomic add per wave.
__global__ __launch_bounds__(CT) void offdiag_nnz_kernel(int64_t n, const double* A, int64_t lda,
                                                         unsigned long long* count)
{
    const int64_t j = blockIdx.y, i = (int64_t)blockIdx.x * CT + threadIdx.x;
    const unsigned long long v = (i < n && i != j && A[i + j * lda] != 0.0) ? 1ull : 0ull;
    unsigned long long s = v;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(count, s);
}

}  // namespace

extern "C" {

int cfmi_offdiag_nnz(cfmi_ctx* c, int64_t n, const double* A, int64_t lda, int64_t* out)
{
    if (!c || !out) return CFMI_EINVAL;
    *out = 0;
    if (n <= 1) return CFMI_OK;
    if (lda < n) return cfmi_fail(c, CFMI_EINVAL, "offdiag_nnz: lda < n");
    if (n > 65535) return cfmi_fail(c, CFMI_EINVAL, "offdiag_nnz: n > 65535 (one grid row per column)");
    // slot 8200 of the reduction staging (potrf uses [0, 4096), logdet 8192)
    unsigned long long* dev = reinterpret_cast<unsigned long long*>(c->d_red + 8200);
    CFMI_HIP(c, hipMemsetAsync(dev, 0, sizeof(unsigned long long), c->stream));
    dim3 grid((unsigned)((n + CT - 1) / CT), (unsigned)n);
    hipLaunchKernelGGL(offdiag_nnz_kernel, grid, dim3(CT), 0, c->stream, n, A, lda, dev);
    CFMI_LAUNCH_CHECK(c);
    unsigned long long h = 0;
    CFMI_HIP(c, hipMemcpyAsync(&h, dev, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    CFMI_HIP(c, hipStreamSynchronize(c->stream));
    *out = (int64_t)h;
    return CFMI_OK;
}

int cfmi_dot_batched(cfmi_ctx* c, int64_t K, int64_t na, const double* A, int64_t lda, int64_t nb,
                     const double* B, int64_t ldb, double* R, int64_t ldr)
{
    if (!c) return CFMI_EINVAL;
    if (K < 0 || na < 0 || nb < 0) return cfmi_fail(c, CFMI_EINVAL, "dot_batched: negative size");
    if (na == 0 || nb == 0) return CFMI_OK;
    if (ldr < na) return cfmi_fail(c, CFMI_EINVAL, "dot_batched: ldr < na");
    if ((na > 1 && lda < K) || (nb > 1 && ldb < K))
        return cfmi_fail(c, CFMI_EINVAL, "dot_batched: leading dimension below K");
    if (K == 0) {
        for (int64_t b = 0; b < nb; ++b) {
            int rc = cfmi_fill(c, CFMI_F64, na, 1, R + b * ldr, na, 0.0);
            if (rc) return rc;
        }
        return CFMI_OK;
    }
    const int64_t nblk = (K + CCH - 1) / CCH;
    if (nblk > 0x7fffffffLL) return cfmi_fail(c, CFMI_EINVAL, "dot_batched: K too large");
    const int64_t nparts = nblk * (CT / 64);
    double* part = (double*)cfmi_workspace(c, (size_t)nparts * na * CMAXB * sizeof(double));
    if (!part) return cfmi_fail(c, CFMI_ENOMEM, "dot_batched: workspace allocation failed");
    for (int64_t b0 = 0; b0 < nb; b0 += CMAXB) {
        const int g = (int)(nb - b0 < CMAXB ? nb - b0 : CMAXB);
        const double* Bg = B + b0 * ldb;
        switch (g) {
        case 1: hipLaunchKernelGGL((dot_partial_kernel<1>), dim3((unsigned)nblk), dim3(CT), 0, c->stream, K, na, A, lda, Bg, ldb, part); break;
        case 2: hipLaunchKernelGGL((dot_partial_kernel<2>), dim3((unsigned)nblk), dim3(CT), 0, c->stream, K, na, A, lda, Bg, ldb, part); break;
        case 3: hipLaunchKernelGGL((dot_partial_kernel<3>), dim3((unsigned)nblk), dim3(CT), 0, c->stream, K, na, A, lda, Bg, ldb, part); break;
        default: hipLaunchKernelGGL((dot_partial_kernel<4>), dim3((unsigned)nblk), dim3(CT), 0, c->stream, K, na, A, lda, Bg, ldb, part); break;
        }
        CFMI_LAUNCH_CHECK(c);
        hipLaunchKernelGGL(dot_reduce_kernel, dim3((unsigned)(na * g)), dim3(CT), 0, c->stream, nparts, na, g, part,
                           R, ldr, (int)b0);
        CFMI_LAUNCH_CHECK(c);
    }
    return CFMI_OK;
}

}  // extern "C"
