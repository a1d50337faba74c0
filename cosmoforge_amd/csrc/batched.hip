// Batched exact likelihood: many covariance matrices of uneven size, each a
// linear combination C_j = sum_k c_jk M_pk of per-patch FP64 component
// matrices, factorised and solved in one launch per stage.
//
// Reference: PICSLike._compute_likelihood_point (picslike/picslike.py:568-664):
// chi2 = d^T C^-1 d and ln|C| per parameter point, lnL = -(chi2 + ln|C|)/2,
// with C^-1 from the Jacobi-scaled dpotrf/dpotri (basics/linalg.py:152-219)
// and ln|C| from a second dpotrf (basics/linalg.py:430-474).
//
// Design (gfx950):
//   * one persistent kernel: a grid of (CUs x k) workgroups pulls jobs from
//     an atomic queue, largest matrices first; every workgroup leaves when
//     the queue is exhausted;
//   * one workgroup per matrix, right-looking blocked Cholesky in the job's
//     workspace (FP32 in mixed mode, FP64 otherwise): the NB x NB diagonal
//     block is factored in LDS in FP64 (its log-diagonal accumulates ln|C|
//     in FP64) and inverted there, the panel is multiplied by that inverse
//     on FP64 MFMA, and the trailing
//     lower triangle is updated by 32 x 32 tiles of 16x16x4 MFMA per wave;
//   * C is Jacobi-scaled (D C D, unit diagonal) as the reference scales it;
//   * solves are blocked: a wave solves each diagonal block with lane
//     shuffles while the workgroup applies the off-diagonal blocks;
//   * refinement against FP64 residuals from the components, without
//     re-reading them per job: the first kernel stores each job's y as a
//     column of its patch's Y_p, one FP64 GEMM per patch forms
//     T_p = [M_1 .. M_c]^T Y_p (every component read once for the patch's
//     points instead of once per point), and a second persistent kernel
//     takes r = D (x - sum_k c_k T_p[k]) per job, one more pair of solves
//     and chi2 = (D x)^T y in FP64.
#include "cfmi_internal.h"
#include "chol64.h"

#include <algorithm>
#include <numeric>
#include <vector>

namespace {

constexpr int BT = 256;     // threads per workgroup
constexpr int TT = 32;      // trailing-update tile edge (2 x 2 MFMA tiles per wave)
constexpr int VMAX = 2048;  // largest matrix (rows) a job may have (LDS vectors)
constexpr int MAXC = 4;     // largest number of component matrices
constexpr int RPT = VMAX / BT;

struct BatchArgs {
    int njobs;
    int ncomp;
    int iters;                 // refinement iterations (kernel 2)
    const int* order;          // processing order of the jobs
    const int* job_patch;
    const int64_t* patch_n;
    const int64_t* patch_comp;  // element offset of the patch's first component
    const int64_t* patch_x;     // element offset of the patch's data vector
    const double* comps;
    const double* xs;
    const double* coef;         // [njobs][ncomp]
    void* work;                 // factors
    const int64_t* job_work;    // element offset of the job's n x n factor
    const int64_t* job_vec;     // element offset of the job's n-vectors
    double* dsc;                // Jacobi scaling per job
    double* logdet;
    double* chi2;
    int* info;
    unsigned* queue;            // [0]: job counter
    // refinement (iters > 0)
    int last;                   // batch_refine_kernel: the last step (writes chi2)
    double* ys;                 // scaled y per job (job_vec offsets)
    double* Y;                  // per patch n x J_p, column r = the unscaled y of the patch's r-th job
    const int64_t* job_ycol;    // element offset of the job's column of Y
    const double* T;            // per patch (ncomp n) x J_p: T[(k n + i) + r (ncomp n)] = (M_k Y_p)[i, r]
    const int64_t* job_tcol;    // element offset of the job's column of T
};

template <typename T> struct Mf;
template <> struct Mf<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 run(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <> struct Mf<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 run(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

__device__ __forceinline__ int next_job(unsigned* q, int* s_job)
{
    if (threadIdx.x == 0) *s_job = (int)atomicAdd(q, 1u);
    __syncthreads();
    int v = *s_job;
    __syncthreads();
    return v;
}

// Lower tiles of a t x t tile grid, enumerated row by row.
__device__ __forceinline__ void tri_tile(int k, int& ti, int& tj)
{
    int r = (int)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= k) ++r;
    while (r * (r + 1) / 2 > k) --r;
    ti = r;
    tj = k - r * (r + 1) / 2;
}

// One wave: the 32 x 32 tile A[i0:i0+32, j0:j0+32] -= P_i P_j^T on and
// below the diagonal, P = A[:, kc:kc+kb] (kb <= NB).  TS: storage type, TC:
// MFMA type.  All NB/4 K-steps of the operands are loaded before the MFMA
// chain so that the tile's loads are in flight together; the product
// accumulates from zero and is added to the stored tile once.
template <typename TS, typename TC, int NB>
__device__ __forceinline__ void update_tile32(int n, TS* A, int i0, int j0, int kc, int kb)
{
    typedef typename Mf<TC>::v4 v4;
    constexpr bool kF64 = sizeof(TC) == 8;
    constexpr int KS = NB / 4;                  // every K-step (of 4) loaded together
    const int lane = threadIdx.x & 63;
    const int lm = lane & 15, lk = lane >> 4;
    // the stored tile, fetched with the operands (its latency overlaps theirs)
    TS cv[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = i0 + 16 * a + lm;
                const int col = j0 + 16 * b + (kF64 ? lk + 4 * r : 4 * lk + r);
                cv[a][b][r] = load_or_zero((const TS*)A, row + (int64_t)col * n, row < n && col < n && row >= col);
            }
    v4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = v4{0, 0, 0, 0};
    for (int k4 = 0; k4 < NB; k4 += 4 * KS) {
        TC av[KS][2], bv[KS][2];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int kk = k4 + 4 * s + lk;
            const bool kin = kk < kb;
            const TS* col = A + (int64_t)(kc + (kin ? kk : 0)) * n;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int ri = i0 + 16 * h + lm, rj = j0 + 16 * h + lm;
                av[s][h] = -(TC)load_or_zero(col, ri, kin && ri < n);
                bv[s][h] = (TC)load_or_zero(col, rj, kin && rj < n);
            }
        }
        // D = P_j P_i^T (the transposed tile): a lane's outputs are rows
        // i0 + 16 a + lm (contiguous in the column-major store) of columns
        // j0 + 16 b + {4 lk + r | lk + 4 r}.
#pragma unroll
        for (int s = 0; s < KS; ++s)
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) acc[a][b] = Mf<TC>::run(bv[s][b], av[s][a], acc[a][b]);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int row = i0 + 16 * a + lm;
                int col = j0 + 16 * b + (kF64 ? lk + 4 * r : 4 * lk + r);
                if (row < n && col < n && row >= col) A[row + (int64_t)col * n] = (TS)(cv[a][b][r] + (TS)acc[a][b][r]);
            }
}

// A[r0:, r0:] -= P P^T on the lower triangle, P = A[r0:, kc:kc+kb]; each
// wave owns 32 x 32 output tiles.
template <typename TS, typename TC, int NB>
__device__ void trailing_update(int n, TS* A, int r0, int kc, int kb)
{
    const int m = n - r0;
    if (m <= 0) return;
    const int t = (m + TT - 1) / TT;
    const int ntiles = t * (t + 1) / 2;
    const int wave = threadIdx.x >> 6;
    for (int tile = wave; tile < ntiles; tile += BT / 64) {
        int ti, tj;
        tri_tile(tile, ti, tj);
        update_tile32<TS, TC, NB>(n, A, r0 + ti * TT, r0 + tj * TT, kc, kb);
    }
}

// Stage the kb x kb diagonal block at k0 (lower part) into LDS as FP64.
template <typename T, int NB>
__device__ __forceinline__ void load_diag(int n, const T* A, int k0, int kb, double (*sd)[NB + 1])
{
#pragma unroll
    for (int s = 0; s < (NB * NB + BT - 1) / BT; ++s) {
        const int idx = threadIdx.x + s * BT;
        if (idx < NB * NB) {
            int r = idx % NB, q = idx / NB;
            sd[r][q] = (double)load_or_zero(A, (int64_t)(k0 + r) + (int64_t)(k0 + q) * n, r < kb && q < kb && r >= q);
        }
    }
}

// In-place Cholesky of the scaled matrix.  Returns 0 or the 1-based pivot
// of the first non-positive diagonal; *logdet_s gets 2 sum log L_ii (FP64).
// P <- P L_kk^-T for the panel rows below the diagonal block, FP64 on MFMA:
// each wave owns 32-row tiles; a tile's rows are read completely before the
// tile is overwritten.
template <typename TS, int NB>
__device__ void panel_solve(int n, TS* A, int k0, int kb, const double (*si)[NB + 1])
{
    typedef Mf<double>::v4 v4;
    constexpr int KS = NB / 4, NT = NB / 16;
    const int r0 = k0 + kb;
    const int m = n - r0;
    if (m <= 0) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lm = lane & 15, lk = lane >> 4;
    const int ntile = (m + TT - 1) / TT;
    for (int tile = wave; tile < ntile; tile += BT / 64) {
        const int i0 = r0 + tile * TT;
        double av[KS][2];
        double outv[NT][2][4];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int kk = 4 * s + lk;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int ri = i0 + 16 * h + lm;
                av[s][h] = (double)load_or_zero((const TS*)A, ri + (int64_t)(k0 + kk) * n, kk < kb && ri < n);
            }
        }
        for (int b = 0; b < NT; ++b) {      // 16 output columns at a time
            v4 acc[2] = {v4{0, 0, 0, 0}, v4{0, 0, 0, 0}};
            const int cc = 16 * b + lm;
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const int kk = 4 * s + lk;
                // op(B)[k][c] = L^-T[k][c] = L^-1[c][k]
                const double bv = (kk < kb && cc < kb) ? si[cc][kk] : 0.0;
#pragma unroll
                for (int a = 0; a < 2; ++a) acc[a] = Mf<double>::run(av[s][a], bv, acc[a]);
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = i0 + 16 * a + lk + 4 * r;
                    if (row < n && cc < kb) outv[b][a][r] = acc[a][r];
                }
        }
#pragma unroll
        for (int b = 0; b < NT; ++b)
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = i0 + 16 * a + lk + 4 * r;
                    const int col = 16 * b + lm;
                    if (row < n && col < kb) A[row + (k0 + col) * n] = (TS)outv[b][a][r];
                }
    }
}

template <typename TS, typename TC, int NB>
__device__ int factor(int n, TS* A, double (*sd)[NB + 1], double (*si)[NB + 1], int* s_bad,
                      double* logdet_s)
{
    const int tid = threadIdx.x;
    double ld = 0.0;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = (int)(n - k0 < NB ? n - k0 : NB);
        load_diag<TS, NB>(n, A, k0, kb, sd);
        if (tid == 0) *s_bad = 0;
        __syncthreads();
        for (int q = 0; q < kb; ++q) {
            if (tid == 0) {
                double v = sd[q][q];
                if (!(v > 0.0)) *s_bad = (int)(k0 + q + 1);
                else sd[q][q] = sqrt(v);
            }
            __syncthreads();
            if (*s_bad) break;
            const double piv = sd[q][q];
            for (int r = q + 1 + tid; r < kb; r += BT) sd[r][q] /= piv;
            __syncthreads();
            // rank-1 update of the remaining lower part (columns q+1..kb-1)
            const int w = kb - q - 1;
            for (int idx = tid; idx < w * w; idx += BT) {
                int r = q + 1 + idx % w, pc = q + 1 + idx / w;
                if (r >= pc) sd[r][pc] -= sd[r][q] * sd[pc][q];
            }
            __syncthreads();
        }
        const int bad = *s_bad;
        if (bad) return bad;
        if (tid == 0)
            for (int q = 0; q < kb; ++q) ld += 2.0 * log(sd[q][q]);
        for (int idx = tid; idx < NB * NB; idx += BT) {
            int r = idx % NB, q = idx / NB;
            if (r < kb && q < kb && r >= q) A[(k0 + r) + (k0 + q) * n] = (TS)sd[r][q];
        }
        // L_kk^-1 (FP64, lower) by columns, one thread per column.
        if (tid < kb) {
            const int j = tid;
            for (int i = 0; i < NB; ++i) si[i][j] = 0.0;
            for (int i = j; i < kb; ++i) {
                double acc = i == j ? 1.0 : 0.0;
                for (int k = j; k < i; ++k) acc -= sd[i][k] * si[k][j];
                si[i][j] = acc / sd[i][i];
            }
        }
        __syncthreads();
        panel_solve<TS, NB>(n, A, k0, kb, si);
        __syncthreads();
        trailing_update<TS, TC, NB>(n, A, k0 + kb, k0, kb);
        __syncthreads();
    }
    if (tid == 0) *logdet_s = ld;
    __syncthreads();
    return 0;
}

// v <- L^-1 v (forward) then L^-T v (backward); v in LDS (FP64, length n).
// red needs (BT / 64) * NB doubles.
template <typename T, int NB>
__device__ void solve_inplace(int n, const T* A, double* v, double (*sd)[NB + 1], double* red)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nblk = (n + NB - 1) / NB;
    for (int bk = 0; bk < nblk; ++bk) {
        const int k0 = bk * NB;
        const int kb = (int)(n - k0 < NB ? n - k0 : NB);
        load_diag<T, NB>(n, A, k0, kb, sd);
        __syncthreads();
        if (wave == 0) {
            double x = lane < kb ? v[k0 + lane] : 0.0;
            for (int q = 0; q < kb; ++q) {
                double xq = __shfl(x, q) / sd[q][q];
                if (lane == q) x = xq;
                else if (lane > q && lane < kb) x -= sd[lane][q] * xq;
            }
            if (lane < kb) v[k0 + lane] = x;
        }
        __syncthreads();
        for (int i = k0 + kb + tid; i < n; i += BT) {
            double s = 0.0;
            for (int q0 = 0; q0 < NB; q0 += 16) {
#pragma unroll
                for (int q = q0; q < q0 + 16; ++q)
                    if (q < kb) s += (double)A[i + (k0 + q) * n] * v[k0 + q];
            }
            v[i] -= s;
        }
        __syncthreads();
    }
    for (int bk = nblk - 1; bk >= 0; --bk) {
        const int k0 = bk * NB;
        const int kb = (int)(n - k0 < NB ? n - k0 : NB);
        // t_c = sum_{i >= k0 + kb} L[i, k0 + c] v_i: per-thread partials over
        // its rows (loads of a row in flight together), then a wave reduction.
        for (int c0 = 0; c0 < NB; c0 += 16) {
            double part[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) part[c] = 0.0;
            for (int i = k0 + kb + tid; i < n; i += BT) {
                const double vi = v[i];
#pragma unroll
                for (int c = 0; c < 16; ++c)
                    if (c0 + c < kb) part[c] += (double)A[i + (k0 + c0 + c) * n] * vi;
            }
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                double s = part[c];
                for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
                if (lane == 0) red[wave * NB + c0 + c] = s;
            }
        }
        load_diag<T, NB>(n, A, k0, kb, sd);
        __syncthreads();
        if (wave == 0) {
            double x = 0.0;
            if (lane < kb) {
                double t = 0.0;
                for (int w = 0; w < BT / 64; ++w) t += red[w * NB + lane];
                x = v[k0 + lane] - t;
            }
            for (int q = kb - 1; q >= 0; --q) {
                double yq = __shfl(x, q) / sd[q][q];
                if (lane == q) x = yq;
                else if (lane < q) x -= sd[q][lane] * yq;
            }
            if (lane < kb) v[k0 + lane] = x;
        }
        __syncthreads();
    }
}

__device__ double block_sum(double s, double* red)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < BT / 64; ++w) t += red[w];
    __syncthreads();
    return t;  // valid on thread 0
}

// Per job: assemble D C D, factor, ln|C|, y = (D C D)^-1 D x refined with
// FP64 residuals from the components, chi2 = (D x)^T y.
template <typename TS, typename TC, int NB>
__global__ __launch_bounds__(BT, 2) void batch_factor_kernel(BatchArgs a)
{
    __shared__ double sd[NB][NB + 1];
    __shared__ double si[NB][NB + 1];
    __shared__ double sv[VMAX];
    __shared__ double sy[VMAX];
    __shared__ double red[(BT / 64) * NB];
    __shared__ int s_job, s_bad;
    __shared__ double s_ld;
    const int tid = threadIdx.x;
    for (;;) {
        const int q = next_job(&a.queue[0], &s_job);
        if (q >= a.njobs) break;
        const int job = a.order[q];
        const int p = a.job_patch[job];
        const int n = (int)a.patch_n[p];
        const double* M = a.comps + a.patch_comp[p];
        const double* x = a.xs + a.patch_x[p];
        TS* A = (TS*)a.work + a.job_work[job];
        double* d = a.dsc + a.job_vec[job];
        double c[MAXC];
#pragma unroll
        for (int k = 0; k < MAXC; ++k) c[k] = k < a.ncomp ? a.coef[(int64_t)job * a.ncomp + k] : 0.0;
        const int nn = n * n;
        // Jacobi scaling from the FP64 diagonal
        double ldd = 0.0;
        for (int i = tid; i < n; i += BT) {
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < MAXC; ++k)
                if (k < a.ncomp) v += c[k] * M[k * nn + i + i * n];
            double di = v > 0.0 ? 1.0 / sqrt(v) : 1.0;
            d[i] = di;
            sv[i] = di;
            ldd += log(di);
        }
        const double sum_logd = block_sum(ldd, red);
        // lower triangle of D C D, column-major, flat over all n^2 entries
        constexpr int U = 4;
        for (int base = tid; base < nn; base += BT * U) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int idx = base + u * BT;
                if (idx < nn) {
                    const int j = idx / n, i = idx - j * n;
                    if (i >= j) {
                        double v = 0.0;
#pragma unroll
                        for (int k = 0; k < MAXC; ++k)
                            if (k < a.ncomp) v += c[k] * M[k * nn + idx];
                        A[idx] = (TS)(v * sv[i] * sv[j]);
                    }
                }
            }
        }
        __syncthreads();
        const int bad = factor<TS, TC, NB>(n, A, sd, si, &s_bad, &s_ld);
        if (bad) {
            if (a.iters > 0) {
                double* yc = a.Y + a.job_ycol[job];
                for (int i = tid; i < n; i += BT) yc[i] = 0.0;
            }
            if (tid == 0) {
                a.info[job] = bad;
                a.logdet[job] = __builtin_nan("");
                a.chi2[job] = __builtin_nan("");
            }
            __syncthreads();
            continue;
        }
        // b = D x, then y = (DCD)^-1 b
        for (int i = tid; i < n; i += BT) sv[i] = d[i] * x[i];
        __syncthreads();
        solve_inplace<TS, NB>(n, A, sv, sd, red);
        if (a.iters > 0) {
            // refinement follows (per-patch residual products, batch_refine_kernel)
            double* ys = a.ys + a.job_vec[job];
            double* yc = a.Y + a.job_ycol[job];
            for (int i = tid; i < n; i += BT) {
                ys[i] = sv[i];
                yc[i] = d[i] * sv[i];
            }
            if (tid == 0) {
                a.info[job] = 0;
                a.logdet[job] = s_ld - 2.0 * sum_logd;
            }
            __syncthreads();
            continue;
        }
        double part = 0.0;
        for (int i = tid; i < n; i += BT) part += d[i] * x[i] * sv[i];
        const double chi = block_sum(part, red);
        if (tid == 0) {
            a.info[job] = 0;
            a.logdet[job] = s_ld - 2.0 * sum_logd;
            a.chi2[job] = chi;
        }
        __syncthreads();
    }
}

// One refinement step per job: r = D (x - sum_k c_k T[k]) with T = M Y from
// the patch product, y += (D C D)^-1 r; the last step writes chi2.
template <typename TS, int NB>
__global__ __launch_bounds__(BT, 2) void batch_refine_kernel(BatchArgs a)
{
    __shared__ double sd[NB][NB + 1];
    __shared__ double sy[VMAX];
    __shared__ double red[(BT / 64) * NB];
    __shared__ int s_job;
    const int tid = threadIdx.x;
    for (;;) {
        const int q = next_job(&a.queue[0], &s_job);
        if (q >= a.njobs) break;
        const int job = a.order[q];
        if (a.info[job] != 0) continue;              // not positive definite: reported by the factor kernel
        const int p = a.job_patch[job];
        const int n = (int)a.patch_n[p];
        const double* x = a.xs + a.patch_x[p];
        const TS* A = (const TS*)a.work + a.job_work[job];
        const double* d = a.dsc + a.job_vec[job];
        double* ys = a.ys + a.job_vec[job];
        const double* t = a.T + a.job_tcol[job];
        double c[MAXC];
#pragma unroll
        for (int k = 0; k < MAXC; ++k) c[k] = k < a.ncomp ? a.coef[(int64_t)job * a.ncomp + k] : 0.0;
        for (int i = tid; i < n; i += BT) {
            double cy = 0.0;
#pragma unroll
            for (int k = 0; k < MAXC; ++k)
                if (k < a.ncomp) cy += c[k] * t[(int64_t)k * n + i];
            sy[i] = d[i] * (x[i] - cy);
        }
        __syncthreads();
        solve_inplace<TS, NB>(n, A, sy, sd, red);
        if (!a.last) {
            double* yc = a.Y + a.job_ycol[job];
            for (int i = tid; i < n; i += BT) {
                const double v = ys[i] + sy[i];
                ys[i] = v;
                yc[i] = d[i] * v;
            }
            __syncthreads();
            continue;
        }
        double part = 0.0;
        for (int i = tid; i < n; i += BT) part += d[i] * x[i] * (ys[i] + sy[i]);
        const double chi = block_sum(part, red);
        if (tid == 0) a.chi2[job] = chi;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Step-synchronous engine (nb = 64): instead of one workgroup walking one
// matrix, every job's step k is one launch per stage over all jobs still
// active at column k0 = 64 k (jobs are ordered largest n first, so they are
// a prefix of the order):
//   sw_diag_kernel    one wave per job: the 64 x 64 diagonal block factored
//                     and inverted in FP64 registers (chol64.h), ln|C|
//                     accumulated in FP64;
//   sw_panel_kernel   one workgroup per (job, 64-row tile) of the panel
//                     below: P <- P L_kk^-T on FP64 MFMA;
//   sw_update_kernel  one workgroup per (job, 64 x 64 lower tile) of the
//                     trailing matrix: A22 -= P_i P_j^T on MFMA.
// Each launch exposes thousands of independent tiles, which is what hides
// the memory latency that a single workgroup per matrix cannot.  The
// assembly reads each component tile once per 32 jobs of its patch.  The
// per-step grids are fixed by the batch's sizes (plans built once), so the
// whole sequence is captured as one graph.
// ---------------------------------------------------------------------------
constexpr int SNB = 64;     // panel width and tile edge of the step-synchronous engine
constexpr int SJC = 32;     // jobs per assembly workgroup

struct StepArgs {
    const int* order;
    const int* job_patch;
    const int64_t* patch_n;
    const int64_t* patch_comp;
    const double* comps;
    const double* coef;
    int ncomp;
    void* work;
    const int64_t* job_work;
    const int64_t* job_vec;
    double* dsc;
    double* linv;               // per job its ceil(n / 64) 64 x 64 FP64 diagonal-block inverses
    const int64_t* job_inv;     // element offset of the job's first block inverse
    double* ldacc;              // per job 2 sum log L_ii (FP64)
    int* info;
    double* rv;                 // per job (job_vec offsets) the right-hand side, solved in place
    int njobs;
    int k0;                     // the step's first column
    // segments of equal n in the order (this step's): first position,
    // per-job tile counts, tile offsets (ascending)
    int nseg;
    const int64_t* seg_q0;
    const int64_t* seg_ppj;
    const int64_t* seg_upj;
    const int64_t* seg_poff;
    const int64_t* seg_uoff;
    // assembly: per patch its lower 64-tile count, chunk count and WG offset, and its jobs
    int npatch;
    const int64_t* asm_tiles;
    const int64_t* asm_chunks;
    const int64_t* asm_off;
    const int* pjobs;
    const int64_t* pjoff;
};

// Last index s < cnt with off[s] <= x (off ascending, off[0] = 0).
__device__ __forceinline__ int find_last_le(const int64_t* off, int cnt, int64_t x)
{
    int lo = 0, hi = cnt - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Jacobi scales d_i = 1 / sqrt(C_ii) from the FP64 diagonals; clears the
// job's ln|C| accumulator and info.
__global__ __launch_bounds__(BT) void sw_scale_kernel(StepArgs a)
{
    const int job = blockIdx.x;
    const int p = a.job_patch[job];
    const int n = (int)a.patch_n[p];
    const double* M = a.comps + a.patch_comp[p];
    double c[MAXC];
#pragma unroll
    for (int k = 0; k < MAXC; ++k) c[k] = k < a.ncomp ? a.coef[(int64_t)job * a.ncomp + k] : 0.0;
    double* d = a.dsc + a.job_vec[job];
    const int64_t nn = (int64_t)n * n;
    for (int i = threadIdx.x; i < n; i += BT) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < MAXC; ++k)
            if (k < a.ncomp) v += c[k] * M[k * nn + i + (int64_t)i * n];
        d[i] = v > 0.0 ? 1.0 / sqrt(v) : 1.0;
    }
    if (threadIdx.x == 0) {
        a.ldacc[job] = 0.0;
        a.info[job] = 0;
    }
}

// Lower 64 x 64 tile (ti, tj) of D C D for up to SJC jobs of one patch: the
// component tiles are read once into registers (thread: row tid % 64,
// columns tid / 64 + 4 u), then each job's combination is stored.
template <typename TS>
__global__ __launch_bounds__(BT) void sw_assemble_kernel(StepArgs a)
{
    const int p = find_last_le(a.asm_off, a.npatch, (int64_t)blockIdx.x);
    const int64_t local = (int64_t)blockIdx.x - a.asm_off[p];
    const int64_t nch = a.asm_chunks[p];
    const int tile = (int)(local / nch), chunk = (int)(local - tile * nch);
    int ti, tj;
    tri_tile(tile, ti, tj);
    const int n = (int)a.patch_n[p];
    const int64_t nn = (int64_t)n * n;
    const double* M = a.comps + a.patch_comp[p];
    const int i = ti * SNB + (threadIdx.x & 63);
    const int jb = tj * SNB + (threadIdx.x >> 6);
    double m[16][MAXC];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int j = jb + 4 * u;
        const bool in = i < n && j < n && i >= j;
#pragma unroll
        for (int k = 0; k < MAXC; ++k) m[u][k] = load_or_zero(M, k * nn + i + (int64_t)j * n, in && k < a.ncomp);
    }
    const int64_t j0 = a.pjoff[p] + (int64_t)chunk * SJC;
    const int64_t j1 = min(a.pjoff[p + 1], j0 + SJC);
    for (int64_t q = j0; q < j1; ++q) {
        const int job = a.pjobs[q];
        double c[MAXC];
#pragma unroll
        for (int k = 0; k < MAXC; ++k) c[k] = k < a.ncomp ? a.coef[(int64_t)job * a.ncomp + k] : 0.0;
        const double* d = a.dsc + a.job_vec[job];
        TS* A = (TS*)a.work + a.job_work[job];
        const double di = i < n ? d[i] : 0.0;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int j = jb + 4 * u;
            if (i < n && j < n && i >= j) {
                double v = 0.0;
#pragma unroll
                for (int k = 0; k < MAXC; ++k) v += c[k] * m[u][k];
                A[i + (int64_t)j * n] = (TS)(v * di * d[j]);
            }
        }
    }
}

template <typename TS>
__global__ __launch_bounds__(64) void sw_diag_kernel(StepArgs a)
{
    __shared__ double lt[SNB][SNB + 1];
    __shared__ double rdiag[SNB];
    const int job = a.order[blockIdx.x];
    if (a.info[job] != 0) return;
    const int n = (int)a.patch_n[a.job_patch[job]];
    const int k0 = a.k0;
    const int kb = min(SNB, n - k0);
    TS* A = (TS*)a.work + a.job_work[job] + k0 + (int64_t)k0 * n;
    const int r = threadIdx.x;
    TS raw[SNB];
#pragma unroll
    for (int c = 0; c < SNB; ++c) raw[c] = A[min(r, kb - 1) + (int64_t)min(c, kb - 1) * n];
    // Consume the stored values before converting and masking them: a load
    // whose address or result select waits on a per-column condition is
    // issued one at a time, and the compiler would otherwise sink each load
    // into the factor step that first reads it.
#pragma unroll
    for (int c = 0; c < SNB; ++c) asm volatile("" ::"v"(raw[c]));
    double row[SNB];
#pragma unroll
    for (int c = 0; c < SNB; ++c) row[c] = (r < kb && c <= r) ? (double)raw[c] : 0.0;
    const int bad = wave_chol64<double>(row, kb, lt, rdiag);
    if (bad) {
        if (r == 0) a.info[job] = k0 + bad;
        return;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < SNB; ++c)
        if (r < kb && c <= r) A[r + (int64_t)c * n] = (TS)lt[c][r];   // L[r][c]
    double ld = r < kb ? 2.0 * log(lt[r][r]) : 0.0;    // lt[q][i] = L[i][q]
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) ld += __shfl_xor(ld, off, 64);
    if (r == 0) a.ldacc[job] += ld;
    wave_inv64_to_lds<double>(lt, rdiag);
    double* li = a.linv + a.job_inv[job] + (int64_t)(k0 / SNB) * SNB * SNB;
#pragma unroll
    for (int cc = 0; cc < SNB; ++cc) li[r + cc * SNB] = lt[cc][r];
}

// Workgroup x of a grid of g runs on XCD x % 8 (round-robin dispatch); it
// takes item x / 8 of that XCD's contiguous share of the grid, so that the
// tiles of one job (which share its panel) meet in one XCD's L2.
__device__ __forceinline__ int64_t xcd_item(int64_t x, int64_t g)
{
    const int64_t q = g >> 3, r = g & 7, xcd = x & 7;
    return xcd * q + (xcd < r ? xcd : r) + (x >> 3);
}

// The (job, tile) of a flat step grid: seg_* tables, per-job counts.
__device__ __forceinline__ bool step_item(const StepArgs& a, const int64_t* off, const int64_t* per, int& job,
                                          int64_t& item)
{
    const int64_t v = xcd_item(blockIdx.x, gridDim.x);
    const int s = find_last_le(off, a.nseg, v);
    const int64_t local = v - off[s];
    const int64_t pj = per[s];
    if (pj <= 0) return false;
    const int64_t q = a.seg_q0[s] + local / pj;
    item = local - (local / pj) * pj;
    job = a.order[q];
    return a.info[job] == 0;
}

// P (SPR rows of the panel below the diagonal block) <- P L_kk^-T, FP64
// MFMA; each wave takes 32-row tiles and reads a tile completely before
// overwriting it.  One load of L_kk^-1 into LDS serves SPR rows.
constexpr int SPR = 256;    // panel rows per workgroup
// Panels per trailing-matrix update (sw_update_kernel).  3 measured 4% faster
// in mixed mode (164 -> 158 ms per disc batch) but the deeper FP32 sums moved
// the mixed -2 ln L from 4.6e-4 to 6.7e-4 of the FP64 batch (target 1e-3).
constexpr int SLA = 2;

template <typename TS>
__global__ __launch_bounds__(256) void sw_panel_kernel(StepArgs a)
{
    typedef Mf<double>::v4 v4;
    __shared__ double si[SNB][SNB + 1];
    int job;
    int64_t tile;
    if (!step_item(a, a.seg_poff, a.seg_ppj, job, tile)) return;
    const int n = (int)a.patch_n[a.job_patch[job]];
    const int k0 = a.k0;
    const int kb = min(SNB, n - k0);
    const double* li = a.linv + a.job_inv[job] + (int64_t)(k0 / SNB) * SNB * SNB;
    block64_to_lds<256>(si, li, (int)threadIdx.x);
    __syncthreads();
    TS* A = (TS*)a.work + a.job_work[job];
    constexpr int KS = SNB / 4, NT = SNB / 16;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lm = lane & 15, lk = lane >> 4;
    for (int half = 0; half < SPR / 128; ++half) {
        const int i0 = k0 + kb + (int)tile * SPR + 128 * half + 32 * wave;
        if (i0 >= n) break;
        // the wave's 32 rows, read completely before any of them is written
        double av[KS][2];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int kk = 4 * s + lk;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int ri = i0 + 16 * h + lm;
                av[s][h] = (double)load_or_zero((const TS*)A, ri + (int64_t)(k0 + kk) * n, kk < kb && ri < n);
            }
        }
#pragma unroll
        for (int b = 0; b < NT; ++b) {
            v4 acc[2] = {v4{0, 0, 0, 0}, v4{0, 0, 0, 0}};
            const int cc = 16 * b + lm;
            // L^-T[kk][cc] = L^-1[cc][kk] is zero for kk > cc: K-steps past 16 b + 15 are skipped
#pragma unroll
            for (int s = 0; s < 4 * (b + 1); ++s) {
                const int kk = 4 * s + lk;
                const double bv = (kk < kb && cc < kb) ? si[cc][kk] : 0.0;
#pragma unroll
                for (int h = 0; h < 2; ++h) acc[h] = Mf<double>::run(av[s][h], bv, acc[h]);
            }
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = i0 + 16 * h + lk + 4 * r;
                    if (row < n && cc < kb) A[row + (int64_t)(k0 + cc) * n] = (TS)acc[h][r];
                }
        }
    }
}

// Trailing updates with look-ahead, so that the trailing matrix is read and
// written once per SLA panels.  Step k (k0 = 64 k, r0 = k0 + kb the first
// row below the diagonal block), phase ph = k mod SLA:
//   ph < SLA - 1: the next block column only, from the panels since the last
//           trailing update, A[r0:, r0:r0+64] -= Q Q[r0:r0+64]^T with
//           Q = [P_{k-ph} .. P_k] (K = 64 ph + kb; tiles (ti, 0));
//   ph = SLA - 1: the whole trailing matrix from the last SLA panels,
//           A[r0:, r0:] -= Q Q^T, Q = [P_{k-SLA+1} .. P_k].
// The workgroup stages the two 64-row operand blocks in LDS one 64-deep K
// chunk at a time (k-major, rows contiguous; a diagonal tile stages one),
// with coalesced column loads, and fetches its stored tile at the start;
// wave w takes the 32 x 32 quadrant (w / 2, w % 2) on MFMA (the upper
// quadrant of a diagonal tile is skipped).  A k-row of 64 entries padded to
// 80 keeps the MFMA operand reads (16 rows x 2 k-rows per half-wave) free of
// LDS bank conflicts for 4- and 8-byte entries.
template <typename TS, typename TC>
__global__ __launch_bounds__(BT) void sw_update_kernel(StepArgs a)
{
    typedef typename Mf<TC>::v4 v4;
    constexpr bool kF64 = sizeof(TC) == 8;
    constexpr int LDT = 80;
    static_assert(LDT >= SNB, "a staged k-row holds the 64 rows of a tile");
    __shared__ TC sa[SNB][LDT];
    __shared__ TC sb[SNB][LDT];
    int job;
    int64_t tile;
    if (!step_item(a, a.seg_uoff, a.seg_upj, job, tile)) return;
    const int ph = (a.k0 / SNB) % SLA;
    const bool wide = ph == SLA - 1;
    int ti, tj;
    if (wide) tri_tile((int)tile, ti, tj);
    else ti = (int)tile, tj = 0;
    const int n = (int)a.patch_n[a.job_patch[job]];
    const int k0 = a.k0;
    const int kb = min(SNB, n - k0);
    const int r0 = k0 + kb;
    const int kbeg = k0 - ph * SNB, kend = k0 + kb;
    const int iA = r0 + ti * SNB, iB = r0 + tj * SNB;
    const bool diag = ti == tj;
    TS* A = (TS*)a.work + a.job_work[job];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wa = wave >> 1, wb = wave & 1;
    const int lm = lane & 15, lk = lane >> 4;
    const bool active = !(diag && wb > wa);
    // the stored quadrant (MFMA output layout)
    TS cv[2][2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = iA + 32 * wa + 16 * x + lm;
                const int col = iB + 32 * wb + 16 * y + (kF64 ? lk + 4 * r : 4 * lk + r);
                cv[x][y][r] = load_or_zero((const TS*)A, row + (int64_t)col * n, active && row < n && col < n && row >= col);
            }
    const TC (*pb)[LDT] = diag ? sa : sb;
    v4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = v4{0, 0, 0, 0};
    // operand blocks, thread (row = lane, k = wave + 4 p): all 32 loads of a
    // chunk issued together (a diagonal tile loads its block twice rather
    // than branch); loading the next chunk during this chunk's products
    // measured no faster
    TS va[SNB / 4], vb[SNB / 4];
    auto fetch = [&](int kc) {
        const int kw = min(SNB, kend - kc);
#pragma unroll
        for (int p = 0; p < SNB / 4; ++p) {
            const int k = wave + 4 * p;
            const TS* col = A + (int64_t)(kc + (k < kw ? k : 0)) * n;
            const int ra = iA + lane, rb = iB + lane;
            va[p] = load_or_zero(col, ra, k < kw && ra < n);
            vb[p] = load_or_zero(col, rb, k < kw && rb < n);
        }
    };
    for (int kc = kbeg; kc < kend; kc += SNB) {
        fetch(kc);
        if (kc != kbeg) __syncthreads();
#pragma unroll
        for (int p = 0; p < SNB / 4; ++p) {
            const int k = wave + 4 * p;
            sa[k][lane] = (TC)va[p];
            sb[k][lane] = (TC)vb[p];
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int s = 0; s < SNB / 4; ++s) {
                const int k = 4 * s + lk;
                TC av[2], bv[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    av[h] = sa[k][32 * wa + 16 * h + lm];
                    bv[h] = pb[k][32 * wb + 16 * h + lm];
                }
                // D = P_j P_i^T (the transposed tile), as update_tile32
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y) acc[x][y] = Mf<TC>::run(bv[y], av[x], acc[x][y]);
            }
        }
    }
    if (!active) return;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = iA + 32 * wa + 16 * x + lm;
                const int col = iB + 32 * wb + 16 * y + (kF64 ? lk + 4 * r : 4 * lk + r);
                if (row < n && col < n && row >= col) A[row + (int64_t)col * n] = (TS)(cv[x][y][r] - (TS)acc[x][y][r]);
            }
}

// Solves, step-synchronous and in place on each job's FP64 vector rv: the
// forward step k takes z_k = L_kk^-1 r_k (the stored FP64 block inverse) and
// subtracts L[below, k] z_k from the rows below; the backward step k (last
// block first) forms r_k - L[below, k]^T x_below, one wave per 16 columns
// with lane-strided rows and a fixed-order wave reduction, then x_k =
// L_kk^-T of it.  One workgroup per job and step, every active job at once.

// rv = D x (mode 0), or the refinement residual D (x - sum_k c_k (M_k Y)_job)
// from the patch products (mode 1).
__global__ __launch_bounds__(BT) void sw_rhs_kernel(BatchArgs a, double* rv, int mode)
{
    const int job = blockIdx.x;
    if (a.info[job] != 0) return;
    const int p = a.job_patch[job];
    const int n = (int)a.patch_n[p];
    const double* x = a.xs + a.patch_x[p];
    const double* d = a.dsc + a.job_vec[job];
    double* r = rv + a.job_vec[job];
    if (mode == 0) {
        for (int i = threadIdx.x; i < n; i += BT) r[i] = d[i] * x[i];
        return;
    }
    const double* t = a.T + a.job_tcol[job];
    double c[MAXC];
#pragma unroll
    for (int k = 0; k < MAXC; ++k) c[k] = k < a.ncomp ? a.coef[(int64_t)job * a.ncomp + k] : 0.0;
    for (int i = threadIdx.x; i < n; i += BT) {
        double cy = 0.0;
#pragma unroll
        for (int k = 0; k < MAXC; ++k)
            if (k < a.ncomp) cy += c[k] * t[(int64_t)k * n + i];
        r[i] = d[i] * (x[i] - cy);
    }
}

template <typename TS>
__global__ __launch_bounds__(BT) void sw_fwd_kernel(StepArgs a)
{
    __shared__ double rk[SNB];
    __shared__ double zk[SNB];
    const int job = a.order[blockIdx.x];
    if (a.info[job] != 0) return;
    const int n = (int)a.patch_n[a.job_patch[job]];
    const int k0 = a.k0;
    const int kb = min(SNB, n - k0);
    double* r = a.rv + a.job_vec[job];
    const double* li = a.linv + a.job_inv[job] + (int64_t)(k0 / SNB) * SNB * SNB;
    const int tid = threadIdx.x;
    if (tid < SNB) rk[tid] = tid < kb ? r[k0 + tid] : 0.0;
    __syncthreads();
    if (tid < SNB) {
        double z = 0.0;
#pragma unroll 16
        for (int c = 0; c < SNB; ++c) z += li[tid + c * SNB] * rk[c];
        zk[tid] = tid < kb ? z : 0.0;   // zero past kb: the clamped column loads below add nothing
        if (tid < kb) r[k0 + tid] = z;
    }
    __syncthreads();
    const TS* L = (const TS*)a.work + a.job_work[job] + (int64_t)k0 * n;
    for (int i = k0 + kb + tid; i < n; i += BT) {
        double acc = 0.0;
#pragma unroll
        for (int c0 = 0; c0 < SNB; c0 += 16) {
            TS lv[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) lv[c] = L[i + (int64_t)min(c0 + c, kb - 1) * n];
            // all 16 in flight before the first use (the chain of sums
            // would otherwise take them one at a time)
#pragma unroll
            for (int c = 0; c < 16; ++c) asm volatile("" ::"v"(lv[c]));
#pragma unroll
            for (int c = 0; c < 16; ++c) acc += (double)lv[c] * zk[c0 + c];
        }
        r[i] -= acc;
    }
}

template <typename TS>
__global__ __launch_bounds__(BT) void sw_bwd_kernel(StepArgs a)
{
    __shared__ double sl[SNB][SNB + 1];
    __shared__ double rk[SNB];
    const int job = a.order[blockIdx.x];
    if (a.info[job] != 0) return;
    const int n = (int)a.patch_n[a.job_patch[job]];
    const int k0 = a.k0;
    const int kb = min(SNB, n - k0);
    double* r = a.rv + a.job_vec[job];
    const double* li = a.linv + a.job_inv[job] + (int64_t)(k0 / SNB) * SNB * SNB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    block64_to_lds<BT>(sl, li, tid);
    // t_c = sum_{i >= k0 + kb} L[i][k0 + c] x_i: wave w takes columns
    // 16 w .. 16 w + 15 together (16 independent sums per lane over the rows
    // lane, lane + 64, ...), then one fixed-order wave reduction per column
    const TS* L = (const TS*)a.work + a.job_work[job] + (int64_t)k0 * n;
    constexpr int CW = SNB / (BT / 64);
    double part[CW];
#pragma unroll
    for (int u = 0; u < CW; ++u) part[u] = 0.0;
    for (int i = k0 + kb + lane; i < n; i += 64) {
        const double xi = r[i];
#pragma unroll
        for (int u = 0; u < CW; ++u) {
            const int c = CW * wave + u;
            part[u] += (double)L[i + (int64_t)min(c, kb - 1) * n] * xi;   // columns past kb are never read back
        }
    }
#pragma unroll
    for (int u = 0; u < CW; ++u) {
        double t = part[u];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        const int c = CW * wave + u;
        if (lane == 0 && c < kb) rk[c] = r[k0 + c] - t;
    }
    __syncthreads();
    if (tid < kb) {
        // x = L_kk^-T rk: x[t] = sum_c L^-1[c][t] rk[c]
        double x = 0.0;
#pragma unroll 16
        for (int c = 0; c < SNB; ++c)
            if (c < kb) x += sl[c][tid] * rk[c];
        r[k0 + tid] = x;
    }
}

// After the solves.  mode 0: ln|C| and chi2 = (D x)^T y; 1: ln|C|, y kept
// for refinement (ys, and the unscaled y as the job's column of Y); 2:
// ys += y (a refinement step that is not the last); 3: chi2 from ys + y.
__global__ __launch_bounds__(BT) void sw_finish_kernel(BatchArgs a, const double* rv, const double* ldacc, int mode)
{
    __shared__ double red[BT / 64];
    const int tid = threadIdx.x;
    const int job = blockIdx.x;
    const int p = a.job_patch[job];
    const int n = (int)a.patch_n[p];
    if (a.info[job] != 0) {
        if (mode != 0 && mode != 3) {
            double* yc = a.Y + a.job_ycol[job];
            for (int i = tid; i < n; i += BT) yc[i] = 0.0;
        }
        if (tid == 0 && mode <= 1) {
            a.logdet[job] = __builtin_nan("");
            a.chi2[job] = __builtin_nan("");
        }
        return;
    }
    const double* x = a.xs + a.patch_x[p];
    const double* d = a.dsc + a.job_vec[job];
    const double* y = rv + a.job_vec[job];
    if (mode <= 1) {
        double ldd = 0.0;
        for (int i = tid; i < n; i += BT) ldd += log(d[i]);
        const double sum_logd = block_sum(ldd, red);
        if (tid == 0) a.logdet[job] = ldacc[job] - 2.0 * sum_logd;
    }
    if (mode == 1 || mode == 2) {
        double* ys = a.ys + a.job_vec[job];
        double* yc = a.Y + a.job_ycol[job];
        for (int i = tid; i < n; i += BT) {
            const double v = mode == 1 ? y[i] : ys[i] + y[i];
            ys[i] = v;
            yc[i] = d[i] * v;
        }
        return;
    }
    const double* ys = a.ys + a.job_vec[job];
    double part = 0.0;
    for (int i = tid; i < n; i += BT) part += d[i] * x[i] * (mode == 0 ? y[i] : ys[i] + y[i]);
    const double chi = block_sum(part, red);
    if (tid == 0) a.chi2[job] = chi;
}

__global__ void reset_queue_kernel(unsigned* q)
{
    if (threadIdx.x < 2) q[threadIdx.x] = 0u;
}

}  // namespace

struct cfmi_batch {
    cfmi_ctx* ctx = nullptr;
    int mode = CFMI_BATCH_MIXED;
    int nb = 32;
    int wg_per_cu = 2;
    int iters = 1;
    int ncomp = 0;
    int njobs = 0;
    int ncu = 256;
    BatchArgs args = {};
    void* d_work = nullptr;
    double* d_dsc = nullptr;
    int* d_order = nullptr;
    int* d_job_patch = nullptr;
    int64_t* d_patch = nullptr;  // patch_n | patch_comp | patch_x
    int64_t* d_job = nullptr;    // job_work | job_vec
    unsigned* d_queue = nullptr;
    // refinement: per-job y, the patches' Y_p and T_p = [M_1 .. M_c]^T Y_p
    double* d_ys = nullptr;
    double* d_Y = nullptr;
    double* d_T = nullptr;
    int64_t* d_jobext = nullptr;  // job_ycol | job_tcol
    std::vector<int64_t> pn, pcomp, pY, pT, pJ;   // per patch (host)
    const double* comps = nullptr;
    // step-synchronous engine (nb = 64): per-job block inverses and ln|C|
    // accumulators, the per-step segment tables and the assembly tables
    double* d_linv = nullptr;
    double* d_ldacc = nullptr;
    double* d_rv = nullptr;
    int64_t* d_plan = nullptr;
    int* d_pjobs = nullptr;
    struct Step {
        int64_t jobs, panel, update;   // grid of each stage
        int nseg;
        int64_t off;                   // of the step's 5 nseg-long arrays in d_plan
    };
    std::vector<Step> steps;
    int64_t asm_wgs = 0, asm_off = 0, pjoff_off = 0;   // assembly grid; table offsets in d_plan
    int64_t inv_off = 0, inv_total = 0;                 // per-job block-inverse offsets in d_plan; their size
    StepArgs sargs = {};
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

template <typename TS, typename TC, int NB>
void launch_factor(cfmi_batch* b, hipStream_t s)
{
    dim3 grid((unsigned)(b->ncu * b->wg_per_cu));
    hipLaunchKernelGGL((batch_factor_kernel<TS, TC, NB>), grid, dim3(BT), 0, s, b->args);
}

template <typename TS, typename TC>
void launch_nb(cfmi_batch* b, hipStream_t s)
{
    if (b->nb == 16) launch_factor<TS, TC, 16>(b, s);
    else launch_factor<TS, TC, 32>(b, s);
}

template <typename TS>
void launch_refine(cfmi_batch* b, hipStream_t s, int last)
{
    BatchArgs a = b->args;
    a.last = last;
    dim3 grid((unsigned)(b->ncu * b->wg_per_cu));
    if (b->nb == 16) hipLaunchKernelGGL((batch_refine_kernel<TS, 16>), grid, dim3(BT), 0, s, a);
    else hipLaunchKernelGGL((batch_refine_kernel<TS, 32>), grid, dim3(BT), 0, s, a);
}

// T_p = [M_1 .. M_c]^T Y_p for every patch (FP64 GEMMs on stream s, no
// split-K workspace so that the sequence can be captured).
int patch_products(cfmi_batch* b, hipStream_t s)
{
    cfmi_ctx* c = b->ctx;
    hipStream_t keep = c->stream;
    c->stream = s;
    int rc = CFMI_OK;
    for (size_t p = 0; p < b->pn.size() && rc == CFMI_OK; ++p) {
        const int64_t n = b->pn[p], J = b->pJ[p];
        if (J == 0) continue;
        rc = cfmi_gemm_impl(c, CFMI_F64, 1, 0, b->ncomp * n, J, n, 1.0, b->comps + b->pcomp[p], n,
                            b->d_Y + b->pY[p], n, 0.0, b->d_T + b->pT[p], b->ncomp * n, CFMI_TRI_NOSPLIT);
    }
    c->stream = keep;
    return rc;
}

// The step-synchronous factorisation and the solves (nb = 64).
template <typename TS, typename TC>
int enqueue_steps(cfmi_batch* b, hipStream_t s)
{
    cfmi_ctx* c = b->ctx;
    StepArgs sa = b->sargs;
    sa.coef = b->args.coef;
    sa.info = b->args.info;
    hipLaunchKernelGGL(sw_scale_kernel, dim3((unsigned)b->njobs), dim3(BT), 0, s, sa);
    CFMI_LAUNCH_CHECK(c);
    hipLaunchKernelGGL((sw_assemble_kernel<TS>), dim3((unsigned)b->asm_wgs), dim3(BT), 0, s, sa);
    CFMI_LAUNCH_CHECK(c);
    for (size_t k = 0; k < b->steps.size(); ++k) {
        const cfmi_batch::Step& st = b->steps[k];
        sa.k0 = (int)(k * SNB);
        sa.nseg = st.nseg;
        const int64_t* t = b->d_plan + st.off;
        sa.seg_q0 = t;
        sa.seg_ppj = t + st.nseg;
        sa.seg_upj = t + 2 * st.nseg;
        sa.seg_poff = t + 3 * st.nseg;
        sa.seg_uoff = t + 4 * st.nseg;
        hipLaunchKernelGGL((sw_diag_kernel<TS>), dim3((unsigned)st.jobs), dim3(64), 0, s, sa);
        CFMI_LAUNCH_CHECK(c);
        if (st.panel > 0) {
            hipLaunchKernelGGL((sw_panel_kernel<TS>), dim3((unsigned)st.panel), dim3(256), 0, s, sa);
            CFMI_LAUNCH_CHECK(c);
        }
        if (st.update > 0) {
            hipLaunchKernelGGL((sw_update_kernel<TS, TC>), dim3((unsigned)st.update), dim3(BT), 0, s, sa);
            CFMI_LAUNCH_CHECK(c);
        }
    }
    for (int it = 0; it <= b->iters; ++it) {
        if (it > 0) {
            int rc = patch_products(b, s);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(sw_rhs_kernel, dim3((unsigned)b->njobs), dim3(BT), 0, s, b->args, b->d_rv, it > 0 ? 1 : 0);
        CFMI_LAUNCH_CHECK(c);
        for (size_t k = 0; k < b->steps.size(); ++k) {
            sa.k0 = (int)(k * SNB);
            hipLaunchKernelGGL((sw_fwd_kernel<TS>), dim3((unsigned)b->steps[k].jobs), dim3(BT), 0, s, sa);
            CFMI_LAUNCH_CHECK(c);
        }
        for (size_t k = b->steps.size(); k-- > 0;) {
            sa.k0 = (int)(k * SNB);
            hipLaunchKernelGGL((sw_bwd_kernel<TS>), dim3((unsigned)b->steps[k].jobs), dim3(BT), 0, s, sa);
            CFMI_LAUNCH_CHECK(c);
        }
        const int mode = it == 0 ? (b->iters > 0 ? 1 : 0) : (it < b->iters ? 2 : 3);
        hipLaunchKernelGGL(sw_finish_kernel, dim3((unsigned)b->njobs), dim3(BT), 0, s, b->args,
                           (const double*)b->d_rv, (const double*)b->d_ldacc, mode);
        CFMI_LAUNCH_CHECK(c);
    }
    return CFMI_OK;
}

int enqueue(cfmi_batch* b, hipStream_t s)
{
    cfmi_ctx* c = b->ctx;
    if (b->nb == SNB) {
        int rc = b->mode == CFMI_BATCH_MIXED ? enqueue_steps<float, float>(b, s)
               : b->mode == CFMI_BATCH_MIXED_F64_STORAGE ? enqueue_steps<double, float>(b, s)
                                                          : enqueue_steps<double, double>(b, s);
        return rc;
    }
    hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(64), 0, s, b->d_queue);
    CFMI_LAUNCH_CHECK(c);
    if (b->mode == CFMI_BATCH_MIXED) launch_nb<float, float>(b, s);
    else if (b->mode == CFMI_BATCH_MIXED_F64_STORAGE) launch_nb<double, float>(b, s);
    else launch_nb<double, double>(b, s);
    CFMI_LAUNCH_CHECK(c);
    for (int it = 0; it < b->iters; ++it) {
        int rc = patch_products(b, s);
        if (rc) return rc;
        hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(64), 0, s, b->d_queue);
        CFMI_LAUNCH_CHECK(c);
        const int last = it == b->iters - 1;
        if (b->mode == CFMI_BATCH_MIXED) launch_refine<float>(b, s, last);
        else launch_refine<double>(b, s, last);
        CFMI_LAUNCH_CHECK(c);
    }
    return CFMI_OK;
}

void release(cfmi_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    (void)hipFree(b->d_work);
    (void)hipFree(b->d_dsc);
    (void)hipFree(b->d_order);
    (void)hipFree(b->d_job_patch);
    (void)hipFree(b->d_patch);
    (void)hipFree(b->d_job);
    (void)hipFree(b->d_queue);
    (void)hipFree(b->d_ys);
    (void)hipFree(b->d_Y);
    (void)hipFree(b->d_T);
    (void)hipFree(b->d_jobext);
    (void)hipFree(b->d_linv);
    (void)hipFree(b->d_ldacc);
    (void)hipFree(b->d_rv);
    (void)hipFree(b->d_plan);
    (void)hipFree(b->d_pjobs);
}

// Plans of the step-synchronous engine: per step the active jobs (a prefix
// of the order, largest n first) grouped by equal n, with their panel and
// trailing tile counts; per patch the assembly tiles and job chunks.
void plan_steps(cfmi_batch* b, int npatch, const int64_t* n, const std::vector<int>& order, const int* job_patch,
                std::vector<int64_t>& tab, std::vector<int>& pjobs)
{
    const int njobs = (int)order.size();
    struct Seg { int64_t q0, cnt, n; };
    std::vector<Seg> segs;
    for (int q = 0; q < njobs; ++q) {
        const int64_t m = n[job_patch[order[q]]];
        if (segs.empty() || segs.back().n != m) segs.push_back({q, 0, m});
        segs.back().cnt += 1;
    }
    const int64_t nmax = segs.empty() ? 0 : segs.front().n;
    b->steps.clear();
    for (int64_t k0 = 0; k0 < nmax; k0 += SNB) {
        cfmi_batch::Step st{0, 0, 0, 0, (int64_t)tab.size()};
        std::vector<int64_t> q0, ppj, upj, poff, uoff;
        for (const Seg& g : segs) {
            if (g.n <= k0) break;
            const int64_t kb = std::min<int64_t>(SNB, g.n - k0);
            const int64_t t = (g.n - k0 - kb + SNB - 1) / SNB;      // trailing 64-tiles per side
            const int64_t tp = (g.n - k0 - kb + SPR - 1) / SPR;     // panel workgroups
            // look-ahead: the next block column, or every SLA-th step the trailing matrix
            const int64_t tu = (k0 / SNB) % SLA == SLA - 1 ? t * (t + 1) / 2 : t;
            q0.push_back(g.q0);
            ppj.push_back(tp);
            upj.push_back(tu);
            poff.push_back(st.panel);
            uoff.push_back(st.update);
            st.jobs += g.cnt;
            st.panel += g.cnt * tp;
            st.update += g.cnt * tu;
        }
        st.nseg = (int)q0.size();
        for (auto* v : {&q0, &ppj, &upj, &poff, &uoff}) tab.insert(tab.end(), v->begin(), v->end());
        b->steps.push_back(st);
    }
    // assembly: per patch its lower tiles x job chunks, jobs listed by patch
    std::vector<int64_t> tiles(npatch), chunks(npatch), off(npatch), pjoff(npatch + 1, 0);
    std::vector<int64_t> cnt(npatch, 0);
    for (int j = 0; j < njobs; ++j) cnt[job_patch[j]] += 1;
    for (int p = 0; p < npatch; ++p) pjoff[p + 1] = pjoff[p] + cnt[p];
    pjobs.assign(njobs, 0);
    {
        std::vector<int64_t> fill(pjoff.begin(), pjoff.end() - 1);
        for (int j = 0; j < njobs; ++j) pjobs[fill[job_patch[j]]++] = j;
    }
    b->asm_wgs = 0;
    for (int p = 0; p < npatch; ++p) {
        const int64_t t = (n[p] + SNB - 1) / SNB;
        tiles[p] = t * (t + 1) / 2;
        chunks[p] = (cnt[p] + SJC - 1) / SJC;
        off[p] = b->asm_wgs;
        b->asm_wgs += tiles[p] * chunks[p];
    }
    b->asm_off = (int64_t)tab.size();
    for (auto* v : {&tiles, &chunks, &off}) tab.insert(tab.end(), v->begin(), v->end());
    b->pjoff_off = (int64_t)tab.size();
    tab.insert(tab.end(), pjoff.begin(), pjoff.end());
    // every job keeps its ceil(n / 64) FP64 diagonal-block inverses (the solves use them)
    b->inv_off = (int64_t)tab.size();
    b->inv_total = 0;
    for (int j = 0; j < njobs; ++j) {
        tab.push_back(b->inv_total);
        b->inv_total += (n[job_patch[j]] + SNB - 1) / SNB * SNB * SNB;
    }
}

}  // namespace

extern "C" {

int cfmi_batch_create(cfmi_ctx* c, int mode, int npatch, const int64_t* n, int ncomp, const double* comps,
                      const int64_t* comp_off, const double* xs, const int64_t* x_off, int njobs,
                      const int* job_patch, int nb, int wg_per_cu, int refine_iters, cfmi_batch** out)
{
    if (!c || !out) return CFMI_EINVAL;
    *out = nullptr;
    if (mode != CFMI_BATCH_MIXED && mode != CFMI_BATCH_MIXED_F64_STORAGE && mode != CFMI_BATCH_FP64)
        return cfmi_fail(c, CFMI_EINVAL, "batch: unknown mode %d", mode);
    if (npatch <= 0 || njobs <= 0) return cfmi_fail(c, CFMI_EINVAL, "batch: empty batch");
    if (ncomp < 1 || ncomp > MAXC) return cfmi_fail(c, CFMI_EINVAL, "batch: 1..%d component matrices", MAXC);
    if (nb != 16 && nb != 32 && nb != SNB)
        return cfmi_fail(c, CFMI_EINVAL, "batch: panel width must be 16, 32 or 64");
    if (wg_per_cu < 1 || wg_per_cu > 8) return cfmi_fail(c, CFMI_EINVAL, "batch: 1..8 workgroups per CU");
    if (refine_iters < 0 || refine_iters > 8) return cfmi_fail(c, CFMI_EINVAL, "batch: 0..8 refinement iterations");
    for (int p = 0; p < npatch; ++p)
        if (n[p] < 1 || n[p] > VMAX)
            return cfmi_fail(c, CFMI_EINVAL, "batch: patch %d has %lld rows (1..%d supported)", p, (long long)n[p],
                             VMAX);
    for (int j = 0; j < njobs; ++j)
        if (job_patch[j] < 0 || job_patch[j] >= npatch)
            return cfmi_fail(c, CFMI_EINVAL, "batch: job %d names patch %d of %d", j, job_patch[j], npatch);
    cfmi_batch* b = new cfmi_batch();
    b->ctx = c;
    b->mode = mode;
    b->nb = nb;
    b->wg_per_cu = wg_per_cu;
    b->iters = refine_iters;
    b->ncomp = ncomp;
    b->njobs = njobs;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && ncu > 0)
        b->ncu = ncu;
    const size_t esz = mode == CFMI_BATCH_MIXED ? 4 : 8;
    std::vector<int64_t> jw(njobs), jv(njobs);
    int64_t w = 0, v = 0;
    for (int j = 0; j < njobs; ++j) {
        int64_t m = n[job_patch[j]];
        jw[j] = w;
        jv[j] = v;
        w += m * m;
        v += m;
    }
    std::vector<int> order(njobs);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return n[job_patch[x]] > n[job_patch[y]]; });
    std::vector<int64_t> pt(3 * (size_t)npatch);
    for (int p = 0; p < npatch; ++p) {
        pt[p] = n[p];
        pt[npatch + p] = comp_off[p];
        pt[2 * npatch + p] = x_off[p];
    }
    std::vector<int64_t> jb(2 * (size_t)njobs);
    for (int j = 0; j < njobs; ++j) {
        jb[j] = jw[j];
        jb[njobs + j] = jv[j];
    }
    // refinement layout: patch p's jobs are the columns of Y_p (n_p x J_p)
    // and T_p (ncomp n_p x J_p), in job order
    b->pn.assign(n, n + npatch);
    b->pcomp.assign(comp_off, comp_off + npatch);
    b->pY.assign(npatch, 0);
    b->pT.assign(npatch, 0);
    b->pJ.assign(npatch, 0);
    for (int j = 0; j < njobs; ++j) b->pJ[job_patch[j]] += 1;
    int64_t ytot = 0;
    for (int p = 0; p < npatch; ++p) {
        b->pY[p] = ytot;
        b->pT[p] = ytot * ncomp;
        ytot += n[p] * b->pJ[p];
    }
    std::vector<int64_t> je(2 * (size_t)njobs);
    {
        std::vector<int64_t> rank(npatch, 0);
        for (int j = 0; j < njobs; ++j) {
            const int p = job_patch[j];
            je[j] = b->pY[p] + rank[p] * n[p];
            je[njobs + j] = b->pT[p] + rank[p] * ncomp * n[p];
            rank[p] += 1;
        }
    }
    b->comps = comps;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t r) { if (e == hipSuccess) e = r; };
    chk(hipMalloc(&b->d_work, (size_t)w * esz));
    chk(hipMalloc((void**)&b->d_dsc, (size_t)v * sizeof(double)));
    chk(hipMalloc((void**)&b->d_order, (size_t)njobs * sizeof(int)));
    chk(hipMalloc((void**)&b->d_job_patch, (size_t)njobs * sizeof(int)));
    chk(hipMalloc((void**)&b->d_patch, pt.size() * sizeof(int64_t)));
    chk(hipMalloc((void**)&b->d_job, jb.size() * sizeof(int64_t)));
    chk(hipMalloc((void**)&b->d_queue, 2 * sizeof(unsigned)));
    if (refine_iters > 0) {
        chk(hipMalloc((void**)&b->d_ys, (size_t)v * sizeof(double)));
        chk(hipMalloc((void**)&b->d_Y, (size_t)ytot * sizeof(double)));
        chk(hipMalloc((void**)&b->d_T, (size_t)ytot * ncomp * sizeof(double)));
        chk(hipMalloc((void**)&b->d_jobext, je.size() * sizeof(int64_t)));
        if (e == hipSuccess)
            chk(hipMemcpy(b->d_jobext, je.data(), je.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    }
    std::vector<int64_t> tab;
    std::vector<int> pjobs;
    if (nb == SNB) {
        plan_steps(b, npatch, n, order, job_patch, tab, pjobs);
        chk(hipMalloc((void**)&b->d_linv, (size_t)b->inv_total * sizeof(double)));
        chk(hipMalloc((void**)&b->d_ldacc, (size_t)njobs * sizeof(double)));
        chk(hipMalloc((void**)&b->d_rv, (size_t)v * sizeof(double)));
        chk(hipMalloc((void**)&b->d_plan, tab.size() * sizeof(int64_t)));
        chk(hipMalloc((void**)&b->d_pjobs, pjobs.size() * sizeof(int)));
        if (e == hipSuccess) {
            chk(hipMemcpy(b->d_plan, tab.data(), tab.size() * sizeof(int64_t), hipMemcpyHostToDevice));
            chk(hipMemcpy(b->d_pjobs, pjobs.data(), pjobs.size() * sizeof(int), hipMemcpyHostToDevice));
        }
    }
    if (e == hipSuccess) {
        chk(hipMemcpy(b->d_order, order.data(), (size_t)njobs * sizeof(int), hipMemcpyHostToDevice));
        chk(hipMemcpy(b->d_job_patch, job_patch, (size_t)njobs * sizeof(int), hipMemcpyHostToDevice));
        chk(hipMemcpy(b->d_patch, pt.data(), pt.size() * sizeof(int64_t), hipMemcpyHostToDevice));
        chk(hipMemcpy(b->d_job, jb.data(), jb.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    }
    if (e != hipSuccess) {
        release(b);
        delete b;
        return cfmi_fail(c, CFMI_ENOMEM, "batch: device allocation failed (%s)", hipGetErrorString(e));
    }
    BatchArgs& a = b->args;
    a.njobs = njobs;
    a.ncomp = ncomp;
    a.iters = refine_iters;
    a.order = b->d_order;
    a.job_patch = b->d_job_patch;
    a.patch_n = b->d_patch;
    a.patch_comp = b->d_patch + npatch;
    a.patch_x = b->d_patch + 2 * npatch;
    a.comps = comps;
    a.xs = xs;
    a.work = b->d_work;
    a.job_work = b->d_job;
    a.job_vec = b->d_job + njobs;
    a.dsc = b->d_dsc;
    a.queue = b->d_queue;
    a.ys = b->d_ys;
    a.Y = b->d_Y;
    a.T = b->d_T;
    a.job_ycol = b->d_jobext;
    a.job_tcol = b->d_jobext ? b->d_jobext + njobs : nullptr;
    if (nb == SNB) {
        StepArgs& t = b->sargs;
        t.order = a.order;
        t.job_patch = a.job_patch;
        t.patch_n = a.patch_n;
        t.patch_comp = a.patch_comp;
        t.comps = comps;
        t.ncomp = ncomp;
        t.work = a.work;
        t.job_work = a.job_work;
        t.job_vec = a.job_vec;
        t.dsc = a.dsc;
        t.linv = b->d_linv;
        t.job_inv = b->d_plan + b->inv_off;
        t.ldacc = b->d_ldacc;
        t.rv = b->d_rv;
        t.njobs = njobs;
        t.npatch = npatch;
        t.asm_tiles = b->d_plan + b->asm_off;
        t.asm_chunks = b->d_plan + b->asm_off + npatch;
        t.asm_off = b->d_plan + b->asm_off + 2 * npatch;
        t.pjobs = b->d_pjobs;
        t.pjoff = b->d_plan + b->pjoff_off;
    }
    *out = b;
    return CFMI_OK;
}

int cfmi_batch_run(cfmi_batch* b, const double* coef, double* chi2, double* logdet, int* info)
{
    if (!b) return CFMI_EINVAL;
    b->args.coef = coef;
    b->args.chi2 = chi2;
    b->args.logdet = logdet;
    b->args.info = info;
    return enqueue(b, b->ctx->stream);
}

int cfmi_batch_capture(cfmi_batch* b, const double* coef, double* chi2, double* logdet, int* info)
{
    if (!b) return CFMI_EINVAL;
    cfmi_ctx* c = b->ctx;
    if (b->exec) {
        hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    if (!b->cap_stream) CFMI_HIP(c, hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
    b->args.coef = coef;
    b->args.chi2 = chi2;
    b->args.logdet = logdet;
    b->args.info = info;
    CFMI_HIP(c, hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue(b, b->cap_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return cfmi_fail(c, CFMI_EHIP, "batch: graph capture failed (%s)", hipGetErrorString(e));
    b->graph = g;
    CFMI_HIP(c, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    return CFMI_OK;
}

int cfmi_batch_replay(cfmi_batch* b)
{
    if (!b) return CFMI_EINVAL;
    if (!b->exec) return cfmi_fail(b->ctx, CFMI_EINVAL, "batch: no captured graph (call cfmi_batch_capture)");
    CFMI_HIP(b->ctx, hipGraphLaunch(b->exec, b->ctx->stream));
    return CFMI_OK;
}

int cfmi_batch_destroy(cfmi_batch* b)
{
    if (!b) return CFMI_OK;
    release(b);
    delete b;
    return CFMI_OK;
}

}  // extern "C"
