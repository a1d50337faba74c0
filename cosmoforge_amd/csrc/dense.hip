// Dense linear algebra for gfx950: MFMA GEMM (FP64 / FP32), blocked
// Cholesky, triangular inverse, Jacobi scaling, log-determinant, Gram.
//
// GEMM: 128 x 128 x 16 tiles, 256 threads (4 waves in a 2 x 2 grid, each
// wave owns 64 x 64 = 4 x 4 MFMA 16x16x4 accumulators).  Global loads for
// the next K-slab are issued into registers while the current slab is
// consumed from LDS.  Split-K writes per-slab partials that a second
// kernel reduces in a fixed order, so results do not depend on scheduling.
#include "cfmi_internal.h"
#include "chol64.h"

#include <algorithm>
#include <type_traits>
#include <utility>
#include <vector>

namespace {

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 run(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 run(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

#ifndef CFMI_GEMM_BK
#define CFMI_GEMM_BK 16
#endif
#ifndef CFMI_GEMM_WAVES
#define CFMI_GEMM_WAVES 2
#endif
constexpr int BM = 128, BN = 128, BK = CFMI_GEMM_BK, NT = 256, PAD = 2;
constexpr int BSMALL = 96;  // tile edge for outputs of at most ~300 x 300 (Gram matrices)

// Grid of an elementwise grid-stride kernel over tot elements (256 threads).
inline dim3 elem_grid(int64_t tot)
{
    int64_t b = (tot + 255) / 256;
    return dim3((unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536));
}

// Where the output block sits and how it is stored.
//   mode 0: dense column-major (ldc)
//   mode 1: packed lower triangle, column-major, of a pn x pn matrix
//   mode 2: packed lower triangle, row-major: element (i, j <= i) at
//           i (i + 1) / 2 + j - pbase (a contiguous band of rows)
struct OutMap {
    int mode = 0;
    int64_t row0 = 0, col0 = 0;  // absolute coordinates of C(0, 0)
    int64_t pn = 0;
    int64_t pbase = 0;
    double pdiag = 1.0;          // packed: scale of diagonal entries
    double poff = 1.0;           // packed: scale of off-diagonal entries
};

// T: input type; TO: output type; TE: epilogue / partial-sum type (double
// when FP32 products are accumulated in FP64 across K-chunks).
template <typename T, typename TO, typename TE>
struct GemmArgs {
    int64_t M, N, K;
    double alpha, beta;
    const T* A;
    int64_t lda;
    const T* B;
    int64_t ldb;
    TO* C;
    int64_t ldc;
    int tri;
    int64_t offA = 0, offB = 0;  // diagonal offsets of the triangular K-trims
    int64_t kchunk;
    TE* part;
    int64_t tm, tn;
    OutMap o;
    int64_t sA = 0, sB = 0, sC = 0;  // per-batch element strides (blockIdx.z)
    // Batched launches without split-K: tiles in groups of gm x gn per XCD
    // over a one-dimensional grid (group_tile); 0: tile_of / tri_tile_of.
    int grouped = 0;
    int64_t gm = 0, gn = 0, gtm = 0, gtmp = 0, gtn = 0, ngroups = 0;
};

__device__ __forceinline__ int64_t packed_index(int64_t i, int64_t j, int64_t n)
{
    return j * (2 * n - j + 1) / 2 + (i - j);
}

// Offset of output element (gm, gn) in g.C and the scale its value takes.
template <typename T, typename TO, typename TE>
__device__ __forceinline__ int64_t out_offset(const GemmArgs<T, TO, TE>& g, int64_t gm, int64_t gn, TE& scale)
{
    const int64_t i = g.o.row0 + gm, j = g.o.col0 + gn;
    scale = TE(1);
    if (g.o.mode == 1) {
        scale = (TE)(i == j ? g.o.pdiag : g.o.poff);
        return packed_index(i, j, g.o.pn);
    }
    if (g.o.mode == 2) {
        scale = (TE)(i == j ? g.o.pdiag : g.o.poff);
        return i * (i + 1) / 2 + j - g.o.pbase;
    }
    return gm + gn * g.ldc;
}

template <typename T, typename TO, typename TE>
__device__ __forceinline__ void store_out(const GemmArgs<T, TO, TE>& g, int64_t gm, int64_t gn, TE v)
{
    TE scale;
    TO* p = g.C + out_offset(g, gm, gn, scale);
    v *= scale;
    *p = g.beta == 0.0 ? (TO)v : (TO)(v + (TE)g.beta * (TE)*p);
}

// Tile order: with a triangular K-trim the heaviest tiles are issued first
// (interleaved over the XCDs); otherwise each XCD gets a contiguous range of
// tiles so that tiles sharing a B panel meet in the same L2.
__device__ __forceinline__ void tile_of(int64_t b, int64_t tm, int64_t tn, int tri, int64_t& ti, int64_t& tj)
{
    const int64_t G = tm * tn;
    int64_t t = b;
    if ((tri & 31) == 0 && G >= 64) {
        const int64_t q = G / 8, r = G % 8, x = b % 8, loc = b / 8;
        t = x < r ? x * (q + 1) + loc : r * (q + 1) + (x - r) * q + loc;
    }
    ti = t % tm;
    tj = t / tm;
    if (tri & 1) ti = tm - 1 - ti;
    if (tri & 4) tj = tn - 1 - tj;
}

// Symmetric result (bit 6, tm == tn): the grid enumerates the lower tiles
// only, row-major (t -> (ti, tj <= ti)), each XCD taking a contiguous,
// equal-count range of them (round-robin dispatch sends block b to XCD b % 8).
__device__ __forceinline__ void tri_tile_of(int64_t b, int64_t tm, int64_t& ti, int64_t& tj)
{
    const int64_t G = tm * (tm + 1) / 2;
    int64_t t = b;
    if (G >= 64) {
        const int64_t q = G / 8, r = G % 8, x = b % 8, loc = b / 8;
        t = x < r ? x * (q + 1) + loc : r * (q + 1) + (x - r) * q + loc;
    }
    int64_t i = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    while (i * (i + 1) / 2 > t) --i;
    ti = i;
    tj = t - i * (i + 1) / 2;
}

template <typename T> struct Vec2;
template <> struct Vec2<double> { typedef double2 type; };
template <> struct Vec2<float> { typedef float2 type; };

// Per thread BK/2 A and BK/2 B elements of a BM x BK / BK x BN slab, loaded as
// pairs contiguous in memory: (m, m+1) for column-major A, (k, k+1) for a
// transposed A; likewise for B.  Interior slabs with 2-element alignment use
// unguarded vector loads; edge slabs use guarded scalar loads.
template <typename T, int TA, int TB, int TBM, int TBN>
struct Slab {
    static constexpr int PA = TBM * BK / 2 / NT;  // element pairs per thread
    static constexpr int PB = TBN * BK / 2 / NT;
    static_assert(PA * 2 * NT == TBM * BK && PB * 2 * NT == TBN * BK, "tile does not split evenly");
    T ra[2 * PA], rb[2 * PB];

    __device__ __forceinline__ static void amap(int idx, int& m, int& k)
    {
        if (TA == 0) { m = (idx % (TBM / 2)) * 2; k = idx / (TBM / 2); }
        else { k = (idx % (BK / 2)) * 2; m = idx / (BK / 2); }
    }
    __device__ __forceinline__ static void bmap(int idx, int& k, int& n)
    {
        if (TB == 0) { k = (idx % (BK / 2)) * 2; n = idx / (BK / 2); }
        else { n = (idx % (TBN / 2)) * 2; k = idx / (TBN / 2); }
    }

    // Pairs along k (TA = 1 for A, TB = 0 for B) take rows beyond M / N
    // clamped to the last row: they feed only outputs that are never stored,
    // so edge tiles keep the vector loads; pairs along m / n need interior
    // rows.
    __device__ __forceinline__ void load_fast(const T* __restrict__ A, int64_t lda, const T* __restrict__ B,
                                              int64_t ldb, int64_t m0, int64_t n0, int64_t k0, int64_t M,
                                              int64_t N, int tid)
    {
        typedef typename Vec2<T>::type V2;
#pragma unroll
        for (int r = 0; r < PA; ++r) {
            int idx = tid + NT * r, m, k, n;
            amap(idx, m, k);
            const int64_t mr = TA == 0 ? m0 + m : (m0 + m < M ? m0 + m : M - 1);
            const T* pa = TA == 0 ? A + mr + (k0 + k) * lda : A + (k0 + k) + mr * lda;
            V2 va = *reinterpret_cast<const V2*>(pa);
            ra[2 * r] = va.x;
            ra[2 * r + 1] = va.y;
            bmap(idx, k, n);
            const int64_t nr = TB == 0 ? (n0 + n < N ? n0 + n : N - 1) : n0 + n;
            const T* pb = TB == 0 ? B + (k0 + k) + nr * ldb : B + nr + (k0 + k) * ldb;
            V2 vb = *reinterpret_cast<const V2*>(pb);
            rb[2 * r] = vb.x;
            rb[2 * r + 1] = vb.y;
        }
    }

    // Edge slabs: every load from a clamped (valid) index, issued together,
    // then the out-of-range entries zeroed (a load under its own condition
    // is issued and waited for one at a time).  kend > k0, M, N >= 1.
    __device__ __forceinline__ void load_guarded(const T* __restrict__ A, int64_t lda, const T* __restrict__ B,
                                                 int64_t ldb, int64_t m0, int64_t n0, int64_t k0, int64_t M,
                                                 int64_t N, int64_t kend, int tid)
    {
#pragma unroll
        for (int r = 0; r < PA; ++r) {
            int idx = tid + NT * r, m, k, n;
            amap(idx, m, k);
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int64_t gm = min(m0 + m + (TA == 0 ? e : 0), M - 1);
                const int64_t gk = min(k0 + k + (TA == 0 ? 0 : e), kend - 1);
                ra[2 * r + e] = TA == 0 ? A[gm + gk * lda] : A[gk + gm * lda];
            }
            bmap(idx, k, n);
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int64_t gk = min(k0 + k + (TB == 0 ? e : 0), kend - 1);
                const int64_t gn = min(n0 + n + (TB == 0 ? 0 : e), N - 1);
                rb[2 * r + e] = TB == 0 ? B[gk + gn * ldb] : B[gn + gk * ldb];
            }
        }
#pragma unroll
        for (int r = 0; r < 2 * PA; ++r) asm volatile("" ::"v"(ra[r]));
#pragma unroll
        for (int r = 0; r < 2 * PB; ++r) asm volatile("" ::"v"(rb[r]));
#pragma unroll
        for (int r = 0; r < PA; ++r) {
            int idx = tid + NT * r, m, k, n;
            amap(idx, m, k);
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int64_t gm = m0 + m + (TA == 0 ? e : 0), gk = k0 + k + (TA == 0 ? 0 : e);
                ra[2 * r + e] = (gm < M && gk < kend) ? ra[2 * r + e] : T(0);
            }
            bmap(idx, k, n);
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int64_t gk = k0 + k + (TB == 0 ? e : 0), gn = n0 + n + (TB == 0 ? 0 : e);
                rb[2 * r + e] = (gn < N && gk < kend) ? rb[2 * r + e] : T(0);
            }
        }
    }

    __device__ __forceinline__ void store(T (*As)[TBM + PAD], T (*Bs)[TBN + PAD], int tid) const
    {
#pragma unroll
        for (int r = 0; r < PA; ++r) {
            int idx = tid + NT * r, m, k, n;
            amap(idx, m, k);
            if (TA == 0) { As[k][m] = ra[2 * r]; As[k][m + 1] = ra[2 * r + 1]; }
            else { As[k][m] = ra[2 * r]; As[k + 1][m] = ra[2 * r + 1]; }
            bmap(idx, k, n);
            if (TB == 0) { Bs[k][n] = rb[2 * r]; Bs[k + 1][n] = rb[2 * r + 1]; }
            else { Bs[k][n] = rb[2 * r]; Bs[k][n + 1] = rb[2 * r + 1]; }
        }
    }
};

// FP32 products accumulated in FP64: the MFMA accumulators are flushed into
// FP64 registers every FLUSH_SLABS K-slabs (FLUSH_SLABS * BK terms).
constexpr int FLUSH_SLABS = 16;

template <typename T, typename TO, typename TE, int TA, int TB, int TBM, int TBN>
__global__ __launch_bounds__(NT, CFMI_GEMM_WAVES) void gemm_kernel(GemmArgs<T, TO, TE> g)
{
    typedef typename Mfma<T>::v4 v4;
    constexpr bool ACC64 = sizeof(TE) > sizeof(T);
    constexpr int BM = TBM, BN = TBN;
    constexpr int FM = BM / 32, FN = BN / 32;  // 16 x 16 MFMA tiles per wave (2 x 2 waves)
    __shared__ T As[BK][BM + PAD];
    __shared__ T Bs[BK][BN + PAD];
    const int tid = threadIdx.x;
    int64_t ti, tj, zb = blockIdx.z;
    if (g.grouped) {
        if (!group_tile(g, blockIdx.x, zb, ti, tj)) return;
    } else if ((g.tri & 64) && g.tm == g.tn) {
        tri_tile_of(blockIdx.x, g.tm, ti, tj);
    } else {
        tile_of(blockIdx.x, g.tm, g.tn, g.tri, ti, tj);
    }
    const int64_t m0 = ti * BM;
    const int64_t n0 = tj * BN;
    const int z = blockIdx.y;
    const int64_t M = g.M, N = g.N, K = g.K;
    if (zb) {
        g.A += zb * g.sA;
        g.B += zb * g.sB;
        g.C += zb * g.sC;
    }
    const bool lower_only = (g.tri & 16) != 0;
    if (lower_only && g.o.col0 + n0 > g.o.row0 + m0 + BM - 1) return;
    if ((g.tri & 64) && tj > ti) return;  // symmetric result: lower tiles only

    int64_t kbeg = (int64_t)z * g.kchunk;
    int64_t kend = kbeg + g.kchunk < K ? kbeg + g.kchunk : K;
    if (g.tri & 1) kend = kend < m0 + BM + g.offA ? kend : m0 + BM + g.offA;
    if (g.tri & 2) kbeg = kbeg > m0 + g.offA ? kbeg : m0 + g.offA;
    if (g.tri & 4) kend = kend < n0 + BN + g.offB ? kend : n0 + BN + g.offB;
    if (g.tri & 8) kbeg = kbeg > n0 + g.offB ? kbeg : n0 + g.offB;

    const int wave = tid >> 6, lane = tid & 63;
    const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * (BN / 2);
    const int lm = lane & 15, lk = lane >> 4;
    const T* __restrict__ A = g.A;
    const T* __restrict__ B = g.B;
    const int64_t lda = g.lda, ldb = g.ldb;
    // Vector loads need 2-element alignment of every column start.
    const bool aligned = ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) % (2 * sizeof(T)) == 0)
                         && (lda % 2 == 0) && (ldb % 2 == 0) && (kbeg % 2 == 0);
    const bool rows_ok = (TA == 1 || m0 + BM <= M) && (TB == 0 || n0 + BN <= N);

    v4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = v4{0, 0, 0, 0};
    TE acc64[ACC64 ? FM : 1][ACC64 ? FN : 1][4];
    if (ACC64) {
#pragma unroll
        for (int i = 0; i < (ACC64 ? FM : 1); ++i)
#pragma unroll
            for (int j = 0; j < (ACC64 ? FN : 1); ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc64[i][j][r] = 0;
    }
    auto flush = [&]() {
        if (ACC64) {
#pragma unroll
            for (int i = 0; i < (ACC64 ? FM : 1); ++i)
#pragma unroll
                for (int j = 0; j < (ACC64 ? FN : 1); ++j) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc64[i][j][r] += (TE)acc[i][j][r];
                    acc[i][j] = v4{0, 0, 0, 0};
                }
        }
    };

    Slab<T, TA, TB, BM, BN> sl;
    auto fetch = [&](int64_t k0) {
        if (aligned && rows_ok && k0 + BK <= kend) sl.load_fast(A, lda, B, ldb, m0, n0, k0, M, N, tid);
        else sl.load_guarded(A, lda, B, ldb, m0, n0, k0, M, N, kend, tid);
    };

    if (kbeg < kend) {
        fetch(kbeg);
        int slab = 0;
        for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
            __syncthreads();
            sl.store(As, Bs, tid);
            __syncthreads();
            if (k0 + BK < kend) fetch(k0 + BK);
#pragma unroll
            for (int kk = 0; kk < BK / 4; ++kk) {
                T a[FM], b[FN];
#pragma unroll
                for (int i = 0; i < FM; ++i) a[i] = As[kk * 4 + lk][wm + i * 16 + lm];
#pragma unroll
                for (int j = 0; j < FN; ++j) b[j] = Bs[kk * 4 + lk][wn + j * 16 + lm];
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<T>::run(a[i], b[j], acc[i][j]);
            }
            if (ACC64 && ++slab == FLUSH_SLABS) {
                flush();
                slab = 0;
            }
        }
    }
    flush();

    // Epilogue: lane holds column lm of each 16 x 16 tile; its four values
    // are rows lk + 4r (f64 16x16x4) or 4 lk + r (f32 16x16x4), as measured
    // on gfx950 with an explicit layout probe.  One 16-row group of tiles at
    // a time: the C values that beta needs are fetched together (from a
    // valid address for elements that are not stored), then the stores.
    constexpr bool kF64 = sizeof(T) == 8;
    const bool read_c = g.part == nullptr && g.beta != 0.0;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        TE old[FN][4];
        if (read_c) {
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t gm = m0 + wm + i * 16 + (kF64 ? lk + 4 * r : lk * 4 + r);
                    const int64_t gn = n0 + wn + j * 16 + lm;
                    const bool ok = gm < M && gn < N && !(lower_only && g.o.col0 + gn > g.o.row0 + gm);
                    TE sc;
                    old[j][r] = (TE)g.C[ok ? out_offset(g, gm, gn, sc) : 0];
                }
            // Consume the values here, so that the compiler cannot sink each
            // load into its element's store branch (one wait per element).
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) asm volatile("" ::"v"(old[j][r]));
        }
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gm = m0 + wm + i * 16 + (kF64 ? lk + 4 * r : lk * 4 + r);
                const int64_t gn = n0 + wn + j * 16 + lm;
                if (gm >= M || gn >= N) continue;
                if (lower_only && g.o.col0 + gn > g.o.row0 + gm) continue;
                TE v;
                if (ACC64) v = acc64[ACC64 ? i : 0][ACC64 ? j : 0][r];
                else v = (TE)acc[i][j][r];
                v *= (TE)g.alpha;
                if (g.part) {
                    g.part[(int64_t)z * M * N + gm + gn * M] = v;
                } else {
                    TE sc;
                    const int64_t off = out_offset(g, gm, gn, sc);
                    v *= sc;
                    g.C[off] = read_c ? (TO)(v + (TE)g.beta * old[j][r]) : (TO)v;
                }
            }
    }
}

template <typename T, typename TO, typename TE>
__global__ void splitk_reduce_kernel(GemmArgs<T, TO, TE> g, int nz)
{
    const int64_t M = g.M, N = g.N;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < M * N;
         idx += (int64_t)gridDim.x * blockDim.x) {
        int64_t m = idx % M, n = idx / M;
        if ((g.tri & 16) && g.o.col0 + n > g.o.row0 + m) continue;
        if ((g.tri & 64) && n > m) continue;
        TE s = 0;
        for (int z = 0; z < nz; ++z) s += g.part[(int64_t)z * M * N + idx];
        store_out(g, m, n, s);
    }
}

// Upper triangle of each batch's n x n C (blockIdx.z, sC apart) from its lower.
template <typename T>
__global__ void mirror_lower_kernel(int64_t n, T* C, int64_t ldc, int64_t sC)
{
    C += (int64_t)blockIdx.z * sC;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n * n;
         idx += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = idx % n, j = idx / n;
        if (j > i) C[i + j * ldc] = C[j + i * ldc];
    }
}

struct GemmShape {
    int64_t offA = 0, offB = 0;
    OutMap om;
    int64_t batch = 1, sA = 0, sB = 0, sC = 0;
    int gm = 0, gn = 0;              // tile group of batched products (group_tile); 0: default
};

// Matrix-vector products in FP64 (M = 1 or N = 1 of a plain product),
// bound by reading the matrix once: a GEMM tile would leave 127 of its 128
// columns idle.
//   gemv_n: y[i] = alpha sum_k A[i + k lda] x[k incx] + beta y[i incy]
//           (rows contiguous): a thread per row over one chunk of GV_KC
//           columns (x chunk in LDS), the chunks' partial sums reduced in a
//           fixed order;
//   gemv_t: y[j] = alpha sum_k A[k + j lda] x[k incx] + beta y[j incy]
//           (columns contiguous): a wave per column, a fixed-order wave
//           reduction.
constexpr int GV_KC = 512;

__global__ __launch_bounds__(256) void gemv_n_kernel(int64_t m, int64_t k, const double* __restrict__ A, int64_t lda,
                                                     const double* __restrict__ x, int64_t incx,
                                                     double* __restrict__ part)
{
    __shared__ double xs[GV_KC];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t k0 = (int64_t)blockIdx.y * GV_KC;
    const int64_t kc = min((int64_t)GV_KC, k - k0);
    for (int j = threadIdx.x; j < kc; j += 256) xs[j] = x[(k0 + j) * incx];
    __syncthreads();
    if (i >= m) return;
    const double* a = A + i + k0 * lda;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int j = 0;
    for (; j + 4 <= kc; j += 4) {
        s0 += a[(j + 0) * lda] * xs[j + 0];
        s1 += a[(j + 1) * lda] * xs[j + 1];
        s2 += a[(j + 2) * lda] * xs[j + 2];
        s3 += a[(j + 3) * lda] * xs[j + 3];
    }
    for (; j < kc; ++j) s0 += a[j * lda] * xs[j];
    part[(int64_t)blockIdx.y * m + i] = (s0 + s1) + (s2 + s3);
}

__global__ void gemv_reduce_kernel(int64_t m, int64_t nz, const double* __restrict__ part, double alpha, double beta,
                                   double* __restrict__ y, int64_t incy)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int64_t z = 0; z < nz; ++z) s += part[z * m + i];
        double* p = y + i * incy;
        *p = alpha * s + (beta == 0.0 ? 0.0 : beta * *p);
    }
}

__global__ __launch_bounds__(256) void gemv_t_kernel(int64_t n, int64_t k, const double* __restrict__ A, int64_t lda,
                                                     const double* __restrict__ x, int64_t incx, double alpha,
                                                     double beta, double* __restrict__ y, int64_t incy)
{
    const int lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= n) return;
    const double* a = A + j * lda;
    double s0 = 0.0, s1 = 0.0;
    int64_t q = lane;
    for (; q + 64 < k; q += 128) {
        s0 += a[q] * x[q * incx];
        s1 += a[q + 64] * x[(q + 64) * incx];
    }
    if (q < k) s0 += a[q] * x[q * incx];
    double s = s0 + s1;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) {
        double* p = y + j * incy;
        *p = alpha * s + (beta == 0.0 ? 0.0 : beta * *p);
    }
}

// The GEMV cases of gemm_x (FP64, plain output); returns false when the
// product is not one.
bool gemv_dispatch(cfmi_ctx* c, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
                   int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int tri, int& rc)
{
    rc = CFMI_OK;
    if (tri != 0 || K <= 0) return false;
    if (N == 1 && !ta) {
        // y (M) = A (M x K, rows contiguous) x, x = B's column (tb = 0) or row (tb = 1)
        const int64_t nz = (K + GV_KC - 1) / GV_KC;
        if (nz > 65535) return false;
        double* part = (double*)cfmi_workspace(c, (size_t)nz * M * sizeof(double));
        if (!part) {
            rc = cfmi_fail(c, CFMI_ENOMEM, "gemv: workspace allocation failed");
            return true;
        }
        hipLaunchKernelGGL(gemv_n_kernel, dim3((unsigned)((M + 255) / 256), (unsigned)nz), dim3(256), 0, c->stream, M,
                           K, A, lda, B, tb ? ldb : (int64_t)1, part);
        if (hipGetLastError() != hipSuccess) { rc = cfmi_fail(c, CFMI_EHIP, "gemv launch failed"); return true; }
        hipLaunchKernelGGL(gemv_reduce_kernel, elem_grid(M), dim3(256), 0, c->stream, M, nz, (const double*)part,
                           alpha, beta, C, (int64_t)1);
        if (hipGetLastError() != hipSuccess) rc = cfmi_fail(c, CFMI_EHIP, "gemv launch failed");
        return true;
    }
    if (M == 1 && !tb) {
        // y (N, stride ldc) = B^T x, B K x N (columns contiguous), x = A's row (ta = 0) or column (ta = 1)
        hipLaunchKernelGGL(gemv_t_kernel, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, c->stream, N, K, B, ldb, A,
                           ta ? (int64_t)1 : lda, alpha, beta, C, ldc);
        if (hipGetLastError() != hipSuccess) rc = cfmi_fail(c, CFMI_EHIP, "gemv launch failed");
        return true;
    }
    return false;
}

template <typename T, typename TO, typename TE>
int gemm_x(cfmi_ctx* c, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
           int64_t lda, const void* B, int64_t ldb, double beta, void* C, int64_t ldc, int tri,
           const GemmShape& sh)
{
    const int64_t batch = sh.batch;
    if (M <= 0 || N <= 0 || batch <= 0) return CFMI_OK;
    if constexpr (std::is_same<T, double>::value && std::is_same<TO, double>::value &&
                  std::is_same<TE, double>::value) {
        int rc = CFMI_OK;
        if (batch == 1 && sh.om.mode == 0 &&
            gemv_dispatch(c, ta, tb, M, N, K, alpha, (const double*)A, lda, (const double*)B, ldb, beta, (double*)C,
                          ldc, tri, rc))
            return rc;
    }
    // FP64 accumulators of FP32 products fit in registers only with 96 x 96 tiles.
    constexpr bool kAcc64 = sizeof(TE) > sizeof(T);
    const bool small = kAcc64 || (M <= 3 * BSMALL && N <= 3 * BSMALL && K > 4096);
    const int tbm = small ? BSMALL : BM, tbn = small ? BSMALL : BN;
    int64_t tm = (M + tbm - 1) / tbm, tn = (N + tbn - 1) / tbn;
    // Split K when the output grid alone cannot fill the 256 CUs: aim at
    // 2048 working blocks (four rounds of two per CU, so the last round's
    // imbalance stays small); a symmetric result only works its lower tiles.
    int64_t nz = 1;
    int64_t tiles = tm * tn * batch;
    const int64_t active = (tri & 64) && tm == tn ? tm * (tm + 1) / 2 * batch : tiles;
    if (!(tri & CFMI_TRI_NOSPLIT) && K > 4 * BK && tiles < 512 && batch == 1) {
        nz = std::min<int64_t>((2048 + active - 1) / active, (K + 255) / 256);
        nz = std::max<int64_t>(nz, 1);
        // Bound the partial buffer to 1 GiB.
        while (nz > 1 && (size_t)nz * M * N * sizeof(TE) > ((size_t)1 << 30)) nz = (nz + 1) / 2;
    }
    int64_t kchunk = BK;
    if (K > 0) {
        kchunk = (K + nz - 1) / nz;
        kchunk = ((kchunk + BK - 1) / BK) * BK;
        nz = (K + kchunk - 1) / kchunk;
    }
    if (nz < 1) nz = 1;
    TE* part = nullptr;
    if (nz > 1) {
        part = (TE*)cfmi_workspace(c, (size_t)nz * M * N * sizeof(TE));
        if (!part) return cfmi_fail(c, CFMI_ENOMEM, "gemm: workspace allocation failed");
    }
    GemmArgs<T, TO, TE> g;
    g.M = M; g.N = N; g.K = K > 0 ? K : 0;
    g.alpha = alpha; g.beta = beta;
    g.A = (const T*)A; g.lda = lda; g.B = (const T*)B; g.ldb = ldb; g.C = (TO*)C; g.ldc = ldc;
    tri &= ~CFMI_TRI_NOSPLIT;
    g.tri = tri; g.offA = sh.offA; g.offB = sh.offB;
    g.kchunk = kchunk; g.part = part; g.tm = tm; g.tn = tn; g.o = sh.om;
    g.sA = sh.sA; g.sB = sh.sB; g.sC = sh.sC;
    const bool sym_tiles = (tri & 64) && tm == tn;
    // tri bit 256: batched tiles in the plain order (tile_of over blockIdx.z)
    g.grouped = batch > 1 && nz == 1 && !sym_tiles && !(tri & 256);
    dim3 grid((unsigned)(sym_tiles ? tm * (tm + 1) / 2 : tm * tn), (unsigned)nz, (unsigned)batch);
    if (g.grouped) {
        // default groups of 1 x 16 tiles (measured on the nside-16
        // whitening: FP64 444 ms at 1 x 16, 446 at 1 x 8, 458 at 2 x 8, 487
        // at 4 x 8, 535 at 8 x 8; FP32 243 at 1 x 16, 244 at 1 x 8, 251 at 2 x 8)
        g.gm = sh.gm > 0 ? sh.gm : 1;
        g.gn = sh.gn > 0 ? sh.gn : 16;
        g.gtm = (tm + g.gm - 1) / g.gm;
        g.gtn = (tn + g.gn - 1) / g.gn;
        g.gtmp = g.gtm | 1;
        g.ngroups = g.gtmp * g.gtn * batch;
        const int64_t nwg = (g.ngroups + 7) / 8 * 8 * g.gm * g.gn;
        if (nwg > 0x7fffffffLL) return cfmi_fail(c, CFMI_EINVAL, "gemm: too many tiles");
        grid = dim3((unsigned)nwg, 1, 1);
    }
#define CFMI_GEMM_LAUNCH(TA_, TB_)                                                                        \
    do {                                                                                                  \
        if constexpr (kAcc64) {                                                                           \
            hipLaunchKernelGGL((gemm_kernel<T, TO, TE, TA_, TB_, BSMALL, BSMALL>), grid, dim3(NT), 0,     \
                               c->stream, g);                                                             \
        } else {                                                                                          \
            if (small) hipLaunchKernelGGL((gemm_kernel<T, TO, TE, TA_, TB_, BSMALL, BSMALL>), grid,       \
                                          dim3(NT), 0, c->stream, g);                                     \
            else hipLaunchKernelGGL((gemm_kernel<T, TO, TE, TA_, TB_, BM, BN>), grid, dim3(NT), 0,        \
                                    c->stream, g);                                                        \
        }                                                                                                 \
    } while (0)
    if (!ta && !tb) CFMI_GEMM_LAUNCH(0, 0);
    else if (!ta && tb) CFMI_GEMM_LAUNCH(0, 1);
    else if (ta && !tb) CFMI_GEMM_LAUNCH(1, 0);
    else CFMI_GEMM_LAUNCH(1, 1);
#undef CFMI_GEMM_LAUNCH
    CFMI_LAUNCH_CHECK(c);
    if (nz > 1) {
        int64_t tot = M * N;
        hipLaunchKernelGGL((splitk_reduce_kernel<T, TO, TE>), elem_grid(tot), dim3(256), 0,
                           c->stream, g, (int)nz);
        CFMI_LAUNCH_CHECK(c);
    }
    if ((tri & 64) && M == N && sh.om.mode == 0) {
        dim3 mg = elem_grid(M * M);
        mg.z = (unsigned)batch;
        hipLaunchKernelGGL((mirror_lower_kernel<TO>), mg, dim3(256), 0, c->stream, M, (TO*)C, ldc, sh.sC);
        CFMI_LAUNCH_CHECK(c);
    }
    return CFMI_OK;
}

// Same-type GEMM (FP64 or FP32 accumulation), the form used inside the library.
template <typename T>
int gemm_t(cfmi_ctx* c, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
           int64_t lda, const void* B, int64_t ldb, double beta, void* C, int64_t ldc, int tri,
           const OutMap& om = OutMap(), int64_t batch = 1, int64_t sA = 0, int64_t sB = 0, int64_t sC = 0)
{
    GemmShape sh;
    sh.om = om; sh.batch = batch; sh.sA = sA; sh.sB = sB; sh.sC = sC;
    return gemm_x<T, T, T>(c, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tri, sh);
}

// ---------------------------------------------------------------------------
// Cholesky: 64-column panels.  The diagonal block is factored in LDS by one
// workgroup, which also forms its inverse; the panel below is multiplied by
// that inverse (TRSM), and the trailing lower triangle is updated by GEMM.
// ---------------------------------------------------------------------------
constexpr int NB = 64;
constexpr int64_t OB2 = 2048; // outer panel width (a multiple of NB)
constexpr int64_t S8_PANEL_MIN_K = 256;   // sliced in-panel updates from this K on

// One wave: factor and inverse of the kb x kb diagonal block (chol64.h);
// the inverse (FP64: blocked, on the matrix cores) leaves through LDS so
// that the global stores are by rows.
template <typename T>
__global__ __launch_bounds__(64) void potrf_diag_kernel(T* A, int64_t lda, int kb, int* info, int64_t koff,
                                                        T* linv, T* U, int64_t ldu)
{
    __shared__ T lt[NB][NB + 1];
    __shared__ T rdiag[NB];
    __shared__ T ut[std::is_same<T, double>::value ? NB : 1][NB + 1];
    __shared__ T tb[std::is_same<T, double>::value ? 32 : 1][33];
    if (*info != 0) return;
    const int r = threadIdx.x;
    T row[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) row[c] = A[min(r, kb - 1) + (int64_t)min(c, kb - 1) * lda];
    // all 64 loads in flight before the first use, then the mask
#pragma unroll
    for (int c = 0; c < NB; ++c) asm volatile("" ::"v"(row[c]));
#pragma unroll
    for (int c = 0; c < NB; ++c) row[c] = (r < kb && c <= r) ? row[c] : T(0);
    const int bad = wave_chol64<T>(row, kb, lt, rdiag);
    if (bad) {
        if (r == 0) atomicCAS(info, 0, (int)(koff + bad));
        return;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (r < kb && c < kb) A[r + c * lda] = lt[c][r];   // L[r][c], zero above the diagonal
    T (*uo)[NB + 1] = lt;
    if constexpr (std::is_same<T, double>::value) {
        wave_inv64_mfma(lt, rdiag, ut, tb);
        uo = ut;
    } else {
        wave_inv64_to_lds<T>(lt, rdiag);
    }
    // lane i stores row i of U: U[i][c] = uo[c][i], consecutive lanes at consecutive addresses
    const int i = r;
#pragma unroll
    for (int cc = 0; cc < NB; ++cc) {
        const T v = uo[cc][i];
        linv[i + cc * NB] = v;
        if (U && i < kb && cc < kb) U[i + cc * ldu] = v;
    }
}

// P (m x kb) <- P Linv^T, TRSM_ROWS rows per workgroup on MFMA (16 x 16 x 4):
// wave w takes rows 32 w .. of each 128-row group, reads its 32 rows
// completely before storing any of them, and skips the K-steps past each
// 16-column block's diagonal (Linv^T is upper triangular).
template <typename T> struct PanelMf;
template <> struct PanelMf<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 run(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <> struct PanelMf<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 run(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

constexpr int TRSM_ROWS = 128;   // panel rows per workgroup (more workgroups: the panel is latency-bound)

template <typename T>
__global__ __launch_bounds__(256) void trsm_panel_kernel(T* P, int64_t ldp, int64_t m, int kb,
                                                         const T* linv, const int* info)
{
    typedef typename PanelMf<T>::v4 v4;
    __shared__ T sl[NB][NB + 1];
    if (*info != 0) return;
    const int tid = threadIdx.x;
    block64_to_lds<256>(sl, linv, tid);   // sl[i][j] = Linv[i][j]
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63;
    const int lm = lane & 15, lk = lane >> 4;
    constexpr int KS = NB / 4, NT = NB / 16;
    for (int half = 0; half < TRSM_ROWS / 128; ++half) {
        const int64_t i0 = (int64_t)blockIdx.x * TRSM_ROWS + 128 * half + 32 * wave;
        if (i0 >= m) break;
        T av[KS][2];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int kk = 4 * s + lk;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int64_t ri = i0 + 16 * h + lm;
                av[s][h] = load_or_zero((const T*)P, ri + (int64_t)kk * ldp, kk < kb && ri < m);
            }
        }
#pragma unroll
        for (int b = 0; b < NT; ++b) {
            v4 acc[2] = {v4{0, 0, 0, 0}, v4{0, 0, 0, 0}};
            const int cc = 16 * b + lm;
#pragma unroll
            for (int s = 0; s < 4 * (b + 1); ++s) {
                const int kk = 4 * s + lk;
                const T bv = (kk < kb && cc < kb) ? sl[cc][kk] : T(0);   // Linv^T[kk][cc]
#pragma unroll
                for (int h = 0; h < 2; ++h) acc[h] = PanelMf<T>::run(av[s][h], bv, acc[h]);
            }
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t row = i0 + 16 * h + PanelMf<T>::row(lk, r);
                    if (row < m && cc < kb) P[row + (int64_t)cc * ldp] = acc[h][r];
                }
        }
    }
}


// The Cholesky's narrow in-panel updates, C (m x w, lower trapezoid: only
// entries with column <= row) -= P P[0:w]^T with P m x K, K <= 128 and
// w <= min(m, 128): one workgroup per 64 x 64 tile of C, its stored values,
// and both operand blocks (one 64-deep K chunk at a time, staged in LDS,
// k-major), fetched with every load of a round in flight; wave w takes the
// 32 x 32 quadrant (w / 2, w % 2) on MFMA.  (A general GEMM tile for these
// shapes spends its time waiting on one K-slab after another.)
constexpr int CU_LDT = 72;   // LDS row pitch of the staged operands (64 rows + pad)

template <typename T>
__global__ __launch_bounds__(256) void chol_update_kernel(int64_t m, int w, int K, const T* __restrict__ P,
                                                          int64_t ldp, T* __restrict__ C, int64_t ldc)
{
    typedef typename PanelMf<T>::v4 v4;
    __shared__ T sa[NB][CU_LDT];
    __shared__ T sb[NB][CU_LDT];
    // tile (ti, tj >= 0) of the trapezoid, column-major over the tiles
    const int64_t tm = (m + NB - 1) / NB;
    int64_t b = blockIdx.x, tj = 0;
    while (b >= tm - tj) {
        b -= tm - tj;
        ++tj;
    }
    const int64_t ti = tj + b;
    const int64_t r0 = ti * NB, c0 = tj * NB;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wa = wave >> 1, wb = wave & 1, lm = lane & 15, lk = lane >> 4;
    const bool diag = ti == tj;
    const bool active = !(diag && wb > wa);
    // the stored quadrant (MFMA output layout)
    T cv[2][2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = r0 + 32 * wa + 16 * x + PanelMf<T>::row(lk, r);
                const int64_t col = c0 + 32 * wb + 16 * y + lm;
                cv[x][y][r] = load_or_zero((const T*)C, row + col * ldc, active && row < m && col < w && col <= row);
            }
    v4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = v4{0, 0, 0, 0};
    for (int kc = 0; kc < K; kc += NB) {
        const int kw = min(NB, K - kc);
        T va[NB / 4], vb[NB / 4];
#pragma unroll
        for (int p = 0; p < NB / 4; ++p) {
            const int k = wave + 4 * p;
            const T* col = P + (int64_t)(kc + (k < kw ? k : 0)) * ldp;
            const int64_t ra = r0 + lane, rb = c0 + lane;
            va[p] = load_or_zero(col, ra, k < kw && ra < m);
            vb[p] = load_or_zero(col, rb, k < kw && rb < w);
        }
        if (kc) __syncthreads();
#pragma unroll
        for (int p = 0; p < NB / 4; ++p) {
            const int k = wave + 4 * p;
            sa[k][lane] = va[p];
            sb[k][lane] = vb[p];
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int s4 = 0; s4 < NB / 4; ++s4) {
                const int k = 4 * s4 + lk;
                T av[2], bv[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    av[h] = sa[k][32 * wa + 16 * h + lm];
                    bv[h] = sb[k][32 * wb + 16 * h + lm];
                }
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y) acc[x][y] = PanelMf<T>::run(av[x], bv[y], acc[x][y]);
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
                const int64_t row = r0 + 32 * wa + 16 * x + PanelMf<T>::row(lk, r);
                const int64_t col = c0 + 32 * wb + 16 * y + lm;
                if (row < m && col < w && col <= row) C[row + col * ldc] = cv[x][y][r] - acc[x][y][r];
            }
}

template <typename T>
__global__ void fill_kernel(int64_t m, int64_t n, T* A, int64_t lda, T v, int mode)
{
    // mode 0: whole block; mode 1: strict upper triangle only.
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < m * n;
         idx += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = idx % m, j = idx / m;
        if (mode == 1 && j <= i) continue;
        A[i + j * lda] = v;
    }
}

// Contiguous fill: 16-byte vector stores (the 16-byte-aligned body), scalar
// head and tail.
template <typename T>
__global__ void fill_flat_kernel(int64_t count, T* A, T v)
{
    constexpr int PER = 16 / sizeof(T);
    typedef T vec __attribute__((ext_vector_type(PER)));
    const int64_t head = (int64_t)(((16 - (reinterpret_cast<uintptr_t>(A) & 15)) & 15) / sizeof(T));
    const int64_t h = head < count ? head : count;
    const int64_t nvec = (count - h) / PER;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, step = (int64_t)gridDim.x * blockDim.x;
    vec* body = reinterpret_cast<vec*>(A + h);
    vec w;
#pragma unroll
    for (int e = 0; e < PER; ++e) w[e] = v;
    for (int64_t i = tid; i < nvec; i += step) body[i] = w;
    const int64_t rest = h + nvec * PER;
    if (tid < h) A[tid] = v;
    if (tid < count - rest) A[rest + tid] = v;
}

template <typename T>
int fill_t(cfmi_ctx* c, int64_t m, int64_t n, void* A, int64_t lda, double v, int mode)
{
    if (m <= 0 || n <= 0) return CFMI_OK;
    int64_t tot = m * n;
    if (mode == 0 && (lda == m || n == 1)) {
        hipLaunchKernelGGL((fill_flat_kernel<T>), elem_grid(tot / (16 / sizeof(T)) + 1), dim3(256), 0, c->stream,
                           tot, (T*)A, (T)v);
        CFMI_LAUNCH_CHECK(c);
        return CFMI_OK;
    }
    hipLaunchKernelGGL((fill_kernel<T>), elem_grid(tot), dim3(256), 0, c->stream,
                       m, n, (T*)A, lda, (T)v, mode);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

// Columns [k0, k0 + w) of the n x n lower triangle, all rows from k0 down,
// already updated by every column left of k0: factored recursively.  The
// left half (a multiple of 64 columns, at least the right half) is factored,
// the right half's columns are updated by one product with K = the left
// half's width, then the right half is factored; 64-column leaves are the
// one-wave diagonal factor and the triangular solve of all rows below.  A
// 2048-column panel thus takes updates of K = 1024, 512, ..., 64 whose
// outputs are only as wide as K: far less traffic than a rank-64 update of
// all the panel's remaining columns after every 64-column step.
template <typename T>
int potrf_panel(cfmi_ctx* c, int64_t n, T* A, int64_t lda, T* U, int64_t ldu, T* linv, int64_t k0, int64_t w,
                int slices)
{
    if (w <= NB) {
        const int kb = (int)w;
        hipLaunchKernelGGL((potrf_diag_kernel<T>), dim3(1), dim3(64), 0, c->stream, A + k0 + k0 * lda, lda, kb,
                           c->d_info, k0, linv, U ? U + k0 + k0 * ldu : (T*)nullptr, ldu);
        CFMI_LAUNCH_CHECK(c);
        const int64_t m = n - k0 - kb;
        if (m > 0) {
            hipLaunchKernelGGL((trsm_panel_kernel<T>), dim3((unsigned)((m + TRSM_ROWS - 1) / TRSM_ROWS)), dim3(256),
                               0, c->stream, A + (k0 + kb) + k0 * lda, lda, m, kb, (const T*)linv,
                               (const int*)c->d_info);
            CFMI_LAUNCH_CHECK(c);
        }
        return CFMI_OK;
    }
    const int64_t blocks = (w + NB - 1) / NB;
    const int64_t h = (blocks + 1) / 2 * NB;
    int rc = potrf_panel<T>(c, n, A, lda, U, ldu, linv, k0, h, slices);
    if (rc) return rc;
    const int64_t m1 = n - (k0 + h);
    T* P = A + (k0 + h) + k0 * lda;
    T* C = A + (k0 + h) + (k0 + h) * lda;
    if (h <= 2 * NB) {
        // the narrow updates (K = 64, 128): 64 x 64 tiles with all loads of a round in flight
        const int64_t tm = (m1 + NB - 1) / NB, tw = (w - h + NB - 1) / NB;
        const int64_t tiles = tw * tm - tw * (tw - 1) / 2;
        hipLaunchKernelGGL((chol_update_kernel<T>), dim3((unsigned)tiles), dim3(256), 0, c->stream, m1, (int)(w - h),
                           (int)h, (const T*)P, lda, C, lda);
        CFMI_LAUNCH_CHECK(c);
    } else if constexpr (std::is_same<T, double>::value) {
        // sliced (FP64-accurate on the int8 matrix cores) once K is deep
        // enough for the products to outweigh the slicing
        if (slices && h >= S8_PANEL_MIN_K) rc = cfmi_s8_update_lower(c, slices, m1, w - h, h, -1.0, P, lda, C, lda);
        else rc = gemm_t<T>(c, 0, 1, m1, w - h, h, -1.0, P, lda, P, lda, 1.0, C, lda, 16);
    } else {
        rc = gemm_t<T>(c, 0, 1, m1, w - h, h, -1.0, P, lda, P, lda, 1.0, C, lda, 16);
    }
    if (rc) return rc;
    return potrf_panel<T>(c, n, A, lda, U, ldu, linv, k0 + h, w - h, slices);
}

// The look-ahead stream and its two events, created on first use.
int lookahead_init(cfmi_ctx* c)
{
    if (c->aux_stream && c->ev_fork && c->ev_join) return CFMI_OK;
    if (!c->aux_stream) CFMI_HIP(c, hipStreamCreateWithFlags(&c->aux_stream, hipStreamNonBlocking));
    if (!c->ev_fork) CFMI_HIP(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    if (!c->ev_join) CFMI_HIP(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    return CFMI_OK;
}

// f() with the context's stream and workspace exchanged for the look-ahead
// stream's: every launch and workspace request inside f goes to that stream.
template <typename F>
int on_aux_stream(cfmi_ctx* c, F&& f)
{
    std::swap(c->stream, c->aux_stream);
    std::swap(c->d_work, c->aux_work);
    std::swap(c->work_bytes, c->aux_work_bytes);
    const int rc = f();
    std::swap(c->stream, c->aux_stream);
    std::swap(c->d_work, c->aux_work);
    std::swap(c->work_bytes, c->aux_work_bytes);
    return rc;
}

// The panel-slices buffer, stream-ordered on the context's stream (every
// factorisation ends with that stream waiting for the look-ahead stream, so
// no look-ahead work still reads a buffer freed here).
void* panel_buffer(cfmi_ctx* c, size_t bytes)
{
    if (bytes <= c->panel_bytes) return c->d_panel;
    if (c->d_panel) hipFreeAsync(c->d_panel, c->stream);
    c->d_panel = nullptr;
    c->panel_bytes = 0;
    if (hipMallocAsync(&c->d_panel, bytes, c->stream) != hipSuccess) {
        c->d_panel = nullptr;
        return nullptr;
    }
    c->panel_bytes = bytes;
    return c->d_panel;
}

// OB2-column panels, each factored recursively (potrf_panel); after each
// panel one rank-OB2 update of the trailing matrix, which carries almost all
// of the n^3/3 flops at a K that keeps the matrix cores busy (a narrower
// update of the whole trailing matrix is bound by re-reading C).
//
// Look-ahead: the update after panel k is split into the next panel's
// columns (on the context's stream, at once) and the rest of the trailing
// matrix (on the look-ahead stream).  Panel k + 1, mostly latency-bound
// one-wave and narrow launches, is then factored while the rest of panel
// k's update fills the matrix cores; panel k + 1's own update waits for it.
// The split changes no entry's arithmetic: each is the same product, with
// the same slices on the sliced route.  (The overlap is partial: a sliced
// update holds every CU with one workgroup each, so the panel's launches
// wait for CUs to free up.  Measured at n = 24192 on one device: FP64
// 108 -> 103 ms, sliced 62.7 -> 59.3 ms with the same bits.)
template <typename T>
int potrf_t(cfmi_ctx* c, int64_t n, void* Av, int64_t lda, void* Uv, int64_t ldu, int* info, int slices = 0)
{
    T* A = (T*)Av;
    T* U = (T*)Uv;
    T* linv = (T*)c->d_red;  // 64 x 64 staging inside the reduction buffer
    CFMI_HIP(c, hipMemsetAsync(c->d_info, 0, sizeof(int), c->stream));
    if (U) {
        int rc = fill_t<T>(c, n, n, U, ldu, 0.0, 0);
        if (rc) return rc;
    }
    constexpr bool dbl = std::is_same<T, double>::value;
    const bool sliced = dbl && slices > 0;
    if (n > 2 * OB2) {
        int rc = lookahead_init(c);
        if (rc) return rc;
    }
    char* ws = nullptr;
    if (sliced && n > OB2) {
        // slices of the largest panel below its diagonal block: n - OB2 rows of OB2 columns
        const size_t bytes = (size_t)slices * (size_t)(n - OB2) * OB2 + 8 * (size_t)(n - OB2);
        ws = (char*)panel_buffer(c, bytes);
        if (!ws) return cfmi_fail(c, CFMI_ENOMEM, "potrf: panel slices of %lld bytes", (long long)bytes);
    }
    bool pending = false;   // look-ahead work the context's stream has not waited for
    auto panels = [&]() -> int {
        for (int64_t K0 = 0; K0 < n; K0 += OB2) {
            const int64_t ob2 = std::min<int64_t>(OB2, n - K0);
            int rc = potrf_panel<T>(c, n, A, lda, U, ldu, linv, K0, ob2, slices);
            if (rc) return rc;
            const int64_t m2 = n - K0 - ob2;
            if (m2 <= 0) continue;
            T* P2 = A + (K0 + ob2) + K0 * lda;
            T* A22 = A + (K0 + ob2) + (K0 + ob2) * lda;
            const int64_t wa = std::min<int64_t>(OB2, m2);   // the next panel's columns
            // the previous panel's look-ahead update covers these columns and reads the panel slices
            if (pending) CFMI_HIP(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
            pending = false;
            const int64_t ldk = (ob2 + 63) / 64 * 64, ps = m2 * ldk;
            double* mu = sliced ? reinterpret_cast<double*>(ws + slices * ps) : nullptr;   // ps: a multiple of 64
            if (sliced) {
                // FP64-accurate on the int8 matrix cores: P2's rows as int8 slices
                rc = cfmi_slice8(c, slices, 1, m2, ob2, (const double*)P2, lda, ws, ldk, ps, mu);
                if (!rc)
                    rc = cfmi_gemm_s8(c, slices, m2, wa, ob2, -1.0, ws, ldk, ps, mu, 0, ws, ldk, ps, mu, 0, 1.0,
                                      (double*)A22, lda, 32, 0, 0, 0, 1.0);
            } else {
                rc = gemm_t<T>(c, 0, 1, m2, wa, ob2, -1.0, P2, lda, P2, lda, 1.0, A22, lda, 16);
            }
            if (rc) return rc;
            if (m2 == wa) continue;
            const int64_t mb = m2 - wa;
            CFMI_HIP(c, hipEventRecord(c->ev_fork, c->stream));
            CFMI_HIP(c, hipStreamWaitEvent(c->aux_stream, c->ev_fork, 0));
            pending = true;
            rc = on_aux_stream(c, [&]() -> int {
                if (sliced)
                    return cfmi_gemm_s8(c, slices, mb, mb, ob2, -1.0, ws + wa * ldk, ldk, ps, mu + wa, 0,
                                        ws + wa * ldk, ldk, ps, mu + wa, 0, 1.0, (double*)(A22 + wa + wa * lda), lda,
                                        32, 0, 0, 0, 1.0);
                return gemm_t<T>(c, 0, 1, mb, mb, ob2, -1.0, P2 + wa, lda, P2 + wa, lda, 1.0, A22 + wa + wa * lda,
                                 lda, 16);
            });
            if (rc) return rc;
            CFMI_HIP(c, hipEventRecord(c->ev_join, c->aux_stream));
        }
        return CFMI_OK;
    };
    int rc = panels();
    if (pending) {
        const hipError_t e = hipStreamWaitEvent(c->stream, c->ev_join, 0);
        if (!rc && e != hipSuccess) rc = cfmi_fail(c, CFMI_EHIP, "potrf: join: %s", hipGetErrorString(e));
    }
    if (rc) return rc;
    rc = fill_t<T>(c, n, n, A, lda, 0.0, 1);
    if (rc) return rc;
    int h = 0;
    CFMI_HIP(c, hipMemcpyAsync(&h, c->d_info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CFMI_HIP(c, hipStreamSynchronize(c->stream));
    if (info) *info = h;
    return h > 0 ? h : CFMI_OK;
}

// Inverse of L's diagonal blocks [j nb, (j + 1) nb) (nb = 64 * 2^k; nb >= n
// inverts the whole of L), bottom up.  The 64 x 64 diagonal blocks of U come
// from potrf.  Level b joins each pair of adjacent inverted b-blocks at
// lo = t 2b, [[U11, 0], [U21, U22]] with U21 = -U22 (L21 U11): the pairs
// of a level are 2b (1 + ld) apart on the diagonal, so every full pair of
// the level is one batched product (two launches per level instead of two
// per pair), and a trailing pair whose second block is short takes one
// more product.  A pair never straddles an nb boundary (2b divides nb).
template <typename T>
int trtri_levels(cfmi_ctx* c, int64_t n, int64_t nb, const T* L, int64_t ldl, T* U, int64_t ldu, T* tmp)
{
    for (int64_t b = NB; b < nb && b < n; b *= 2) {
        const int64_t full = n / (2 * b);
        const int64_t sL = 2 * b * (1 + ldl), sU = 2 * b * (1 + ldu);
        for (int64_t t0 = 0; t0 < full; t0 += 65535) {
            const int64_t cnt = std::min<int64_t>(65535, full - t0);
            const T* L0 = L + t0 * sL;
            T* U0 = U + t0 * sU;
            // tmp_t (b x b) = L21 U11   (U11 lower: op(B)[k,n] = 0 for k < n)
            int rc = gemm_t<T>(c, 0, 0, b, b, b, 1.0, L0 + b, ldl, U0, ldu, 0.0, tmp, b, 8, OutMap(), cnt, sL, sU,
                               b * b);
            if (rc) return rc;
            // U21 = -U22 tmp   (U22 lower: op(A)[m,k] = 0 for k > m)
            rc = gemm_t<T>(c, 0, 0, b, b, b, -1.0, U0 + b + b * ldu, ldu, tmp, b, 0.0, U0 + b, ldu, 1, OutMap(),
                           cnt, sU, b * b, sU);
            if (rc) return rc;
        }
        const int64_t lo = full * 2 * b, m2 = n - lo - b;
        if (m2 > 0) {
            const int64_t mid = lo + b;
            int rc = gemm_t<T>(c, 0, 0, m2, b, b, 1.0, L + mid + lo * ldl, ldl, U + lo + lo * ldu, ldu, 0.0, tmp,
                               m2, 8);
            if (rc) return rc;
            rc = gemm_t<T>(c, 0, 0, m2, b, m2, -1.0, U + mid + mid * ldu, ldu, tmp, m2, 0.0, U + mid + lo * ldu, ldu,
                           1);
            if (rc) return rc;
        }
    }
    return CFMI_OK;
}

__global__ void diag_scale_kernel(int64_t n, const double* A, int64_t lda, double* d)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double a = fabs(A[i + i * lda]);
        if (a == 0.0) a = 1.0;
        d[i] = 1.0 / sqrt(a);
    }
}

// Column-blocked grid for the elementwise m x n passes: blockIdx.y walks the
// columns, the x-blocks cover EW_ROWS rows of a column each (no 64-bit
// divisions per element, and each thread's loads of a column in flight).
constexpr int EW_PER = 8;                    // elements per thread and column
constexpr int64_t EW_ROWS = 256 * EW_PER;

inline dim3 ew_grid(int64_t m, int64_t n)
{
    const int64_t bx = (m + EW_ROWS - 1) / EW_ROWS;
    return dim3((unsigned)(bx > 0 ? bx : 1), (unsigned)(n < 65535 ? (n > 0 ? n : 1) : 65535));
}

__global__ void scale_flat_kernel(int mode, int64_t m, int64_t n, double* A, int64_t lda, const double* d)
{
    // the flat form for short columns; mode as scale_kernel
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < m * n;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx % m, j = idx / m;
        A[i + j * lda] *= mode == 0 ? d[j] : mode == 1 ? d[i] : d[i] * d[j];
    }
}

__global__ void scale_kernel(int mode, int64_t m, int64_t n, double* A, int64_t lda, const double* d)
{
    // mode 0: columns, 1: rows, 2: both (D A D)
    for (int64_t j = blockIdx.y; j < n; j += gridDim.y) {
        const double dj = mode == 1 ? 1.0 : d[j];
        double* col = A + j * lda;
        for (int64_t i0 = (int64_t)blockIdx.x * EW_ROWS + threadIdx.x; i0 < m; i0 += (int64_t)gridDim.x * EW_ROWS) {
            double v[EW_PER], di[EW_PER];
#pragma unroll
            for (int e = 0; e < EW_PER; ++e) {
                const int64_t i = i0 + 256 * e;
                v[e] = load_or_zero((const double*)col, i, i < m);
                di[e] = load_or_zero(d, mode == 0 ? 0 : i, mode != 0 && i < m);   // d has n entries in mode 0
            }
#pragma unroll
            for (int e = 0; e < EW_PER; ++e) {
                const int64_t i = i0 + 256 * e;
                if (i < m) col[i] = v[e] * (mode == 0 ? dj : (mode == 1 ? di[e] : di[e] * dj));
            }
        }
    }
}

// The flat form for short columns (one element per column, a strided diagonal):
__global__ void axpby_flat_kernel(int64_t m, int64_t n, double alpha, const double* A, int64_t lda, double beta,
                                  double* B, int64_t ldb)
{
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < m * n;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx % m, j = idx / m;
        double* p = B + i + j * ldb;
        *p = alpha * A[i + j * lda] + (beta == 0.0 ? 0.0 : beta * *p);
    }
}

__global__ void axpby_kernel(int64_t m, int64_t n, double alpha, const double* A, int64_t lda, double beta,
                             double* B, int64_t ldb)
{
    for (int64_t j = blockIdx.y; j < n; j += gridDim.y) {
        const double* a = A + j * lda;
        double* b = B + j * ldb;
        for (int64_t i0 = (int64_t)blockIdx.x * EW_ROWS + threadIdx.x; i0 < m; i0 += (int64_t)gridDim.x * EW_ROWS) {
            double va[EW_PER], vb[EW_PER];
#pragma unroll
            for (int e = 0; e < EW_PER; ++e) {
                const int64_t i = i0 + 256 * e;
                va[e] = load_or_zero(a, i, i < m);
                vb[e] = load_or_zero((const double*)b, i, i < m);   // not used when beta = 0
            }
#pragma unroll
            for (int e = 0; e < EW_PER; ++e) {
                const int64_t i = i0 + 256 * e;
                if (i < m) b[i] = alpha * va[e] + (beta == 0.0 ? 0.0 : beta * vb[e]);
            }
        }
    }
}

template <typename S, typename D>
__global__ void convert_kernel(int64_t m, int64_t n, const S* src, int64_t lds, D* dst, int64_t ldd)
{
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < m * n;
         idx += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = idx % m, j = idx / m;
        dst[i + j * ldd] = (D)src[i + j * lds];
    }
}

// Deterministic single-workgroup reduction: 2 sum log L_ii - 2 sum log d_i.
__global__ __launch_bounds__(256) void logdet_kernel(int64_t n, const double* L, int64_t ldl,
                                                     const double* d, double* out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        s += 2.0 * log(L[i + i * ldl]);
        if (d) s -= 2.0 * log(d[i]);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

}  // namespace

int cfmi_gemm_impl(cfmi_ctx* c, int dtype, int ta, int tb, int64_t M, int64_t N, int64_t K,
                   double alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                   double beta, void* C, int64_t ldc, int tri)
{
    if (dtype == CFMI_F64) return gemm_t<double>(c, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tri);
    if (dtype == CFMI_F32) return gemm_t<float>(c, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tri);
    return cfmi_fail(c, CFMI_EINVAL, "gemm: unsupported dtype %d", dtype);
}

extern "C" {

int cfmi_gemm(cfmi_ctx* c, int dtype, int transa, int transb, int64_t m, int64_t n, int64_t k,
              double alpha, const void* A, int64_t lda, const void* B, int64_t ldb, double beta,
              void* C, int64_t ldc, int flags)
{
    if (!c) return CFMI_EINVAL;
    if (m < 0 || n < 0 || k < 0) return cfmi_fail(c, CFMI_EINVAL, "gemm: negative size");
    if (ldc < (m > 1 ? m : 1)) return cfmi_fail(c, CFMI_EINVAL, "gemm: ldc < m");
    if (lda < 1 || ldb < 1) return cfmi_fail(c, CFMI_EINVAL, "gemm: bad leading dimension");
    return cfmi_gemm_impl(c, dtype, transa != 0, transb != 0, m, n, k, alpha, A, lda, B, ldb, beta, C,
                          ldc, flags);
}

int cfmi_gemm_packed(cfmi_ctx* c, int dtype, int transa, int transb, int64_t m, int64_t n, int64_t k,
                     double alpha, const void* A, int64_t lda, const void* B, int64_t ldb, double beta, void* P,
                     int64_t pn, int64_t row0, int64_t col0, double pdiag, double poff, int flags)
{
    return cfmi_gemm_packed_batched(c, dtype, transa, transb, m, n, k, alpha, A, lda, 0, B, ldb, 0, beta, P, 0,
                                    pn, row0, col0, pdiag, poff, 1, flags);
}

int cfmi_gemm_packed_batched(cfmi_ctx* c, int dtype, int transa, int transb, int64_t m, int64_t n, int64_t k,
                             double alpha, const void* A, int64_t lda, int64_t sA, const void* B, int64_t ldb,
                             int64_t sB, double beta, void* P, int64_t sP, int64_t pn, int64_t row0, int64_t col0,
                             double pdiag, double poff, int64_t batch, int flags)
{
    if (!c) return CFMI_EINVAL;
    if (m < 0 || n < 0 || k < 0 || pn <= 0 || row0 < 0 || col0 < 0 || row0 + m > pn || col0 + n > pn)
        return cfmi_fail(c, CFMI_EINVAL, "gemm_packed: block outside the %lld x %lld matrix", (long long)pn,
                         (long long)pn);
    if (batch < 1 || batch > 65535) return cfmi_fail(c, CFMI_EINVAL, "gemm_packed: batch out of range");
    OutMap om;
    om.mode = 1; om.row0 = row0; om.col0 = col0; om.pn = pn; om.pdiag = pdiag; om.poff = poff;
    int tri = (flags & 15) | 16;
    if (dtype == CFMI_F64)
        return gemm_t<double>(c, transa != 0, transb != 0, m, n, k, alpha, A, lda, B, ldb, beta, P, 1, tri, om,
                              batch, sA, sB, sP);
    if (dtype == CFMI_F32)
        return gemm_t<float>(c, transa != 0, transb != 0, m, n, k, alpha, A, lda, B, ldb, beta, P, 1, tri, om,
                             batch, sA, sB, sP);
    return cfmi_fail(c, CFMI_EINVAL, "gemm_packed: unsupported dtype %d", dtype);
}

int cfmi_gemm_batched(cfmi_ctx* c, int dtype, int transa, int transb, int64_t m, int64_t n, int64_t k,
                      double alpha, const void* A, int64_t lda, int64_t sA, const void* B, int64_t ldb, int64_t sB,
                      double beta, void* C, int64_t ldc, int64_t sC, int64_t batch, int flags)
{
    if (!c) return CFMI_EINVAL;
    if (m < 0 || n < 0 || k < 0) return cfmi_fail(c, CFMI_EINVAL, "gemm_batched: negative size");
    if (batch < 1 || batch > 65535) return cfmi_fail(c, CFMI_EINVAL, "gemm_batched: batch out of range");
    if (ldc < (m > 1 ? m : 1)) return cfmi_fail(c, CFMI_EINVAL, "gemm_batched: ldc < m");
    if (dtype == CFMI_F64)
        return gemm_t<double>(c, transa != 0, transb != 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, flags & 31,
                              OutMap(), batch, sA, sB, sC);
    if (dtype == CFMI_F32)
        return gemm_t<float>(c, transa != 0, transb != 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, flags & 31,
                             OutMap(), batch, sA, sB, sC);
    return cfmi_fail(c, CFMI_EINVAL, "gemm_batched: unsupported dtype %d", dtype);
}

int cfmi_potrf(cfmi_ctx* c, int dtype, int64_t n, void* A, int64_t lda, void* U, int64_t ldu, int* info)
{
    if (!c) return CFMI_EINVAL;
    if (n < 0 || lda < n) return cfmi_fail(c, CFMI_EINVAL, "potrf: bad size n=%lld lda=%lld",
                                           (long long)n, (long long)lda);
    if (U && ldu < n) return cfmi_fail(c, CFMI_EINVAL, "potrf: ldu < n");
    if (info) *info = 0;
    if (n == 0) return CFMI_OK;
    if (dtype == CFMI_F64) return potrf_t<double>(c, n, A, lda, U, ldu, info);
    if (dtype == CFMI_F32) return potrf_t<float>(c, n, A, lda, U, ldu, info);
    return cfmi_fail(c, CFMI_EINVAL, "potrf: unsupported dtype %d", dtype);
}

int cfmi_potrf_s8(cfmi_ctx* c, int s, int64_t n, double* A, int64_t lda, double* U, int64_t ldu, int* info)
{
    if (!c) return CFMI_EINVAL;
    if (s < 3 || s > 5) return cfmi_fail(c, CFMI_EINVAL, "potrf_s8: %d slices (3 to 5 are built)", s);
    if (n < 0 || lda < n) return cfmi_fail(c, CFMI_EINVAL, "potrf_s8: bad size n=%lld lda=%lld",
                                           (long long)n, (long long)lda);
    if (U && ldu < n) return cfmi_fail(c, CFMI_EINVAL, "potrf_s8: ldu < n");
    if (info) *info = 0;
    if (n == 0) return CFMI_OK;
    return potrf_t<double>(c, n, A, lda, U, ldu, info, s);
}

// Device scratch of trtri_levels: the largest level's b x b products.
static int64_t trtri_tmp_elems(int64_t n, int64_t nb)
{
    int64_t big = 0;
    for (int64_t b = NB; b < nb && b < n; b *= 2) {
        const int64_t full = std::min<int64_t>(n / (2 * b), 65535);
        big = std::max<int64_t>(big, std::max<int64_t>(full * b * b, (n - n / (2 * b) * 2 * b - b) * b));
    }
    return big;
}

static int trtri_run(cfmi_ctx* c, int dtype, int64_t n, int64_t nb, const void* L, int64_t ldl, void* U, int64_t ldu)
{
    const int64_t elems = trtri_tmp_elems(n, nb);
    if (elems <= 0) return CFMI_OK;
    const size_t esz = dtype == CFMI_F64 ? 8 : 4;
    // Kept apart from the shared workspace, which gemm's split-K may grow.
    void* tmp2 = nullptr;
    CFMI_HIP(c, hipMallocAsync(&tmp2, (size_t)elems * esz, c->stream));
    int rc;
    if (dtype == CFMI_F64)
        rc = trtri_levels<double>(c, n, nb, (const double*)L, ldl, (double*)U, ldu, (double*)tmp2);
    else
        rc = trtri_levels<float>(c, n, nb, (const float*)L, ldl, (float*)U, ldu, (float*)tmp2);
    hipFreeAsync(tmp2, c->stream);
    return rc;
}

int cfmi_trtri(cfmi_ctx* c, int dtype, int64_t n, const void* L, int64_t ldl, void* U, int64_t ldu)
{
    if (!c) return CFMI_EINVAL;
    if (dtype != CFMI_F64 && dtype != CFMI_F32) return cfmi_fail(c, CFMI_EINVAL, "trtri: unsupported dtype %d", dtype);
    if (n < 0 || ldl < n || ldu < n) return cfmi_fail(c, CFMI_EINVAL, "trtri: bad size or leading dimension");
    return trtri_run(c, dtype, n, n, L, ldl, U, ldu);
}

int cfmi_trtri_blocks(cfmi_ctx* c, int dtype, int64_t n, int64_t nb, const void* L, int64_t ldl, void* U,
                      int64_t ldu)
{
    if (!c) return CFMI_EINVAL;
    if (dtype != CFMI_F64 && dtype != CFMI_F32)
        return cfmi_fail(c, CFMI_EINVAL, "trtri_blocks: unsupported dtype %d", dtype);
    if (n < 0 || ldl < n || ldu < n) return cfmi_fail(c, CFMI_EINVAL, "trtri_blocks: bad size or leading dimension");
    if (nb < NB || (nb / NB) & (nb / NB - 1) || nb % NB)
        return cfmi_fail(c, CFMI_EINVAL, "trtri_blocks: block size %lld is not 64 times a power of two",
                         (long long)nb);
    return trtri_run(c, dtype, n, nb, L, ldl, U, ldu);
}

int cfmi_jacobi_scale(cfmi_ctx* c, int64_t n, double* A, int64_t lda, double* d)
{
    if (!c) return CFMI_EINVAL;
    if (n <= 0) return CFMI_OK;
    hipLaunchKernelGGL(diag_scale_kernel, elem_grid(n), dim3(256), 0, c->stream, n,
                       (const double*)A, lda, d);
    CFMI_LAUNCH_CHECK(c);
    hipLaunchKernelGGL(scale_kernel, ew_grid(n, n), dim3(256), 0, c->stream, 2, n, n, A, lda, (const double*)d);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_scale(cfmi_ctx* c, int side, int64_t m, int64_t n, double* A, int64_t lda, const double* d)
{
    if (!c) return CFMI_EINVAL;
    if (m <= 0 || n <= 0) return CFMI_OK;
    if (m >= 256)
        hipLaunchKernelGGL(scale_kernel, ew_grid(m, n), dim3(256), 0, c->stream, side ? 1 : 0, m, n, A, lda, d);
    else
        hipLaunchKernelGGL(scale_flat_kernel, elem_grid(m * n), dim3(256), 0, c->stream, side ? 1 : 0, m, n, A, lda,
                           d);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_axpby(cfmi_ctx* c, int64_t m, int64_t n, double alpha, const double* A, int64_t lda, double beta,
               double* B, int64_t ldb)
{
    if (!c) return CFMI_EINVAL;
    if (m <= 0 || n <= 0) return CFMI_OK;
    if (m >= 256)
        hipLaunchKernelGGL(axpby_kernel, ew_grid(m, n), dim3(256), 0, c->stream, m, n, alpha, A, lda, beta, B, ldb);
    else
        hipLaunchKernelGGL(axpby_flat_kernel, elem_grid(m * n), dim3(256), 0, c->stream, m, n, alpha, A, lda, beta,
                           B, ldb);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_convert(cfmi_ctx* c, int dir, int64_t m, int64_t n, const void* src, int64_t lds, void* dst,
                 int64_t ldd)
{
    if (!c) return CFMI_EINVAL;
    if (m <= 0 || n <= 0) return CFMI_OK;
    int64_t tot = m * n;
    dim3 g = elem_grid(tot);
    if (dir == 0)
        hipLaunchKernelGGL((convert_kernel<double, float>), g, dim3(256), 0, c->stream, m, n,
                           (const double*)src, lds, (float*)dst, ldd);
    else
        hipLaunchKernelGGL((convert_kernel<float, double>), g, dim3(256), 0, c->stream, m, n,
                           (const float*)src, lds, (double*)dst, ldd);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_copy(cfmi_ctx* c, int dtype, int64_t m, int64_t n, const void* src, int64_t lds, void* dst, int64_t ldd)
{
    if (!c) return CFMI_EINVAL;
    if (m <= 0 || n <= 0) return CFMI_OK;
    int64_t tot = m * n;
    dim3 g = elem_grid(tot);
    if (dtype == CFMI_F64)
        hipLaunchKernelGGL((convert_kernel<double, double>), g, dim3(256), 0, c->stream, m, n, (const double*)src,
                           lds, (double*)dst, ldd);
    else if (dtype == CFMI_F32)
        hipLaunchKernelGGL((convert_kernel<float, float>), g, dim3(256), 0, c->stream, m, n, (const float*)src,
                           lds, (float*)dst, ldd);
    else
        return cfmi_fail(c, CFMI_EINVAL, "copy: unsupported dtype %d", dtype);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_fill(cfmi_ctx* c, int dtype, int64_t m, int64_t n, void* A, int64_t lda, double v)
{
    if (!c) return CFMI_EINVAL;
    if (dtype == CFMI_F64) return fill_t<double>(c, m, n, A, lda, v, 0);
    if (dtype == CFMI_F32) return fill_t<float>(c, m, n, A, lda, v, 0);
    return cfmi_fail(c, CFMI_EINVAL, "fill: unsupported dtype %d", dtype);
}

int cfmi_logdet_chol(cfmi_ctx* c, int64_t n, const double* L, int64_t ldl, const double* d, double* out)
{
    if (!c) return CFMI_EINVAL;
    double* dev = c->d_red + 8192;
    hipLaunchKernelGGL(logdet_kernel, dim3(1), dim3(256), 0, c->stream, n, L, ldl, d, dev);
    CFMI_LAUNCH_CHECK(c);
    CFMI_HIP(c, hipMemcpyAsync(out, dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CFMI_HIP(c, hipStreamSynchronize(c->stream));
    return CFMI_OK;
}

int cfmi_gram(cfmi_ctx* c, int dtype, int64_t m, int64_t n, int64_t K, const void* X, int64_t ldx,
              const void* Y, int64_t ldy, double* G, int64_t ldg)
{
    if (!c) return CFMI_EINVAL;
    const int sym = (X == Y && m == n && ldx == ldy) ? 64 : 0;
    GemmShape sh;
    if (dtype == CFMI_F64)
        return gemm_x<double, double, double>(c, 1, 0, m, n, K, 1.0, X, ldx, Y, ldy, 0.0, G, ldg, sym, sh);
    // FP32 products, FP32 MFMA accumulation within K-chunks, FP64 across them.
    if (dtype == CFMI_F32)
        return gemm_x<float, double, double>(c, 1, 0, m, n, K, 1.0, X, ldx, Y, ldy, 0.0, G, ldg, sym, sh);
    return cfmi_fail(c, CFMI_EINVAL, "gram: unsupported dtype %d", dtype);
}

int cfmi_gemm_ex(cfmi_ctx* c, int dtype, int out_dtype, int transa, int transb, int64_t m, int64_t n, int64_t k,
                 double alpha, const void* A, int64_t lda, int64_t sA, int64_t off_a, const void* B, int64_t ldb,
                 int64_t sB, int64_t off_b, double beta, void* C, int64_t ldc, int64_t sC, int64_t batch, int flags,
                 int out_mode, int64_t row0, int64_t col0, int64_t pbase, double pdiag, double poff)
{
    if (!c) return CFMI_EINVAL;
    if (m < 0 || n < 0 || k < 0) return cfmi_fail(c, CFMI_EINVAL, "gemm_ex: negative size");
    if (batch < 1 || batch > 65535) return cfmi_fail(c, CFMI_EINVAL, "gemm_ex: batch out of range");
    if (lda < 1 || ldb < 1) return cfmi_fail(c, CFMI_EINVAL, "gemm_ex: bad leading dimension");
    if (out_mode < 0 || out_mode > 2) return cfmi_fail(c, CFMI_EINVAL, "gemm_ex: bad out_mode %d", out_mode);
    if (out_mode == 0 && ldc < (m > 1 ? m : 1)) return cfmi_fail(c, CFMI_EINVAL, "gemm_ex: ldc < m");
    if (out_mode == 1 && (row0 < 0 || col0 < 0 || row0 + m > ldc || col0 + n > ldc))
        return cfmi_fail(c, CFMI_EINVAL, "gemm_ex: block outside the packed %lld x %lld matrix", (long long)ldc,
                         (long long)ldc);
    if (out_mode == 2 && (row0 < 0 || col0 < 0 || pbase < 0 || pbase > row0 * (row0 + 1) / 2))
        return cfmi_fail(c, CFMI_EINVAL, "gemm_ex: packed row band starts after row %lld", (long long)row0);
    GemmShape sh;
    sh.offA = off_a; sh.offB = off_b;
    sh.om.mode = out_mode; sh.om.row0 = row0; sh.om.col0 = col0; sh.om.pn = out_mode == 1 ? ldc : 0;
    sh.om.pbase = pbase; sh.om.pdiag = pdiag; sh.om.poff = poff;
    sh.batch = batch; sh.sA = sA; sh.sB = sB; sh.sC = sC;
    // bits 12-15 / 16-19: v > 0 gives 2^(v - 1) group rows / columns (1 .. 64)
    if ((flags >> 12) & 15) sh.gm = 1 << std::min(((flags >> 12) & 15) - 1, 6);
    if ((flags >> 16) & 15) sh.gn = 1 << std::min(((flags >> 16) & 15) - 1, 6);
    int tri = (flags & (31 | 64)) | ((flags & 1024) ? 256 : 0);
    if (out_mode != 0) tri = (tri & (31 | 256)) | 16;
    if (out_mode != 0) ldc = 1;
    const bool acc64 = (flags & CFMI_GEMM_ACC64) != 0;
    if (dtype == CFMI_F64 && out_dtype == CFMI_F64)
        return gemm_x<double, double, double>(c, transa != 0, transb != 0, m, n, k, alpha, A, lda, B, ldb, beta, C,
                                              ldc, tri, sh);
    if (dtype == CFMI_F32 && out_dtype == CFMI_F32 && !acc64)
        return gemm_x<float, float, float>(c, transa != 0, transb != 0, m, n, k, alpha, A, lda, B, ldb, beta, C,
                                           ldc, tri, sh);
    if (dtype == CFMI_F32 && out_dtype == CFMI_F32)
        return gemm_x<float, float, double>(c, transa != 0, transb != 0, m, n, k, alpha, A, lda, B, ldb, beta, C,
                                            ldc, tri, sh);
    if (dtype == CFMI_F32 && out_dtype == CFMI_F64)
        return gemm_x<float, double, double>(c, transa != 0, transb != 0, m, n, k, alpha, A, lda, B, ldb, beta, C,
                                             ldc, tri, sh);
    return cfmi_fail(c, CFMI_EINVAL, "gemm_ex: unsupported dtype pair %d -> %d", dtype, out_dtype);
}

}  // extern "C"
