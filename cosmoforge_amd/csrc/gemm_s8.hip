// FP64 products on the int8 matrix cores (gfx950 v_mfma_i32_16x16x64_i8),
// exact integer accumulation, operands held as int8 slices.
//
// Slicing.  Row r of an operand (the k-contiguous rows of A, or of B^T) is
// scaled by mu_r = max_k |x_rk| and split into S signed int8 slices:
//     t = (x / mu) * 127,   s_1 = rint(t),   t <- (t - s_1) * 254,
//     s_2 = rint(t),        t <- (t - s_2) * 254,  ...
// so x = mu (s_1 / 127 + s_2 / (127 254) + ... ) up to |x - sum| <=
// mu / (2 127 254^(S-1)); every |s_i| <= 127 (the remainders stay in
// [-1/2, 1/2]).  S = 4 keeps 2^-32 of the row maximum, S = 5 2^-40.
//
// Product.  C_mn = mu_m nu_n / 127^2 sum_d 254^-d sum_{p+q=d} <A_p[m], B_q[n]>
// over the levels d = 0 .. S-1 (S (S+1) / 2 int8 products; the dropped
// levels d >= S are below the slicing error).  The inner sums are exact
// int32 MFMA sums: level d has d + 1 products of |.| <= 127^2 per k, so
// S K 127^2 < 2^31 bounds K per launch (the host splits longer K and adds
// the chunks in FP64).  The epilogue combines the levels in FP64 (Horner,
// no multiply-add contraction in this file), scales by the row factors
// and applies alpha / beta, so the result is a deterministic function of
// the slices: tests reproduce it bit for bit on the host.
//
// Tiles 128 x 128, 8 waves of 32 x 64 (2 x 4 MFMA tiles of 16 x 16, S
// int32 accumulators each), one 64-deep K-slab of every slice in LDS per
// step (64-byte rows, 16-byte chunks XOR-swizzled by row), the next slab
// copied into the other LDS stage during the MFMAs.  Operand lane layout
// (measured on gfx950): lane l reads A[m = l % 16][k = 16 (l / 16) + j] and
// B[k = 16 (l / 16) + j][n = l % 16], j = 0..15 (16 bytes); output row
// 4 (l / 16) + r, column l % 16.
#include "cfmi_internal.h"

#include <algorithm>
#include <cstring>

// Every floating-point operation below is rounded on its own (no fused
// multiply-add), so the host restatement can reproduce the results exactly.
// (The header intrinsics __dmul_rn / __dadd_rn are plain operators compiled
// with contraction on, so they are not used here.)
#pragma clang fp contract(off)

namespace {

typedef int i4 __attribute__((ext_vector_type(4)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));

constexpr int SK = 64;     // k per slab (one MFMA deep)
constexpr int BT = 128;    // tile edge
constexpr int NTH = 512;   // 8 waves: 4 along M x 2 along N, 32 x 64 each
constexpr double kInv254 = 1.0 / 254.0;
constexpr double kInv127sq = 1.0 / 16129.0;

__device__ __forceinline__ int swz(int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

// Slices of 16 consecutive values of one row into S 16-byte vectors.
template <int S>
__device__ __forceinline__ void slice16(const double (&x)[16], double inv_mu, u4 (&out)[S])
{
    unsigned w[S][4];
#pragma unroll
    for (int p = 0; p < S; ++p)
#pragma unroll
        for (int e = 0; e < 4; ++e) w[p][e] = 0u;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        double t = (x[j] * inv_mu) * 127.0;
        if (!(fabs(t) <= 127.0)) t = 0.0;              // NaN / Inf rows: the scale carries the NaN
#pragma unroll
        for (int p = 0; p < S; ++p) {
            const double s = rint(t);
            const unsigned b = (unsigned)(int)s & 0xffu;
            w[p][j >> 2] |= b << (8 * (j & 3));
            t = (t - s) * 254.0;
        }
    }
#pragma unroll
    for (int p = 0; p < S; ++p) out[p] = u4{w[p][0], w[p][1], w[p][2], w[p][3]};
}

// mu = max |x| over the row (NaN if any entry is NaN); 1 for an all-zero row.
__device__ __forceinline__ double finish_mu(double m) { return m == 0.0 ? 1.0 : m; }
__device__ __forceinline__ double absmax(double m, double v)
{
    const double a = fabs(v);
    return (a > m || a != a) ? a : m;
}

// Layout 0: row r at src + r lds (k contiguous).  One wave per row: the
// row maximum by a wave reduction, then 1024-value chunks, read coalesced
// (64 consecutive values per load) into the wave's LDS buffer and sliced
// as 16 consecutive values per lane (a pad entry every 16 keeps both LDS
// sides free of bank conflicts); zero beyond k.  NCH > 0: the row (at most
// NCH chunks) is read once and held in registers for both passes; NCH = 0:
// read twice.
template <int S>
__device__ __forceinline__ void slice_chunk(double* b, const double (&v)[16], int lane, int64_t base, int64_t r,
                                            int64_t ldk, int64_t ps, double inv, char* __restrict__ dst)
{
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const int e = t * 64 + lane;                  // chunk element
        b[e + (e >> 4)] = v[t];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int64_t c = base / 16 + lane;              // this lane's 16-value group
    if (c * 16 < ldk) {
        double x[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) x[j] = b[lane * 17 + j];
        u4 o[S];
        slice16<S>(x, inv, o);
#pragma unroll
        for (int p = 0; p < S; ++p) *reinterpret_cast<u4*>(dst + p * ps + r * ldk + c * 16) = o[p];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int S, int NCH>
__global__ __launch_bounds__(256) void slice_rows_kernel(int64_t rows, int64_t k, const double* __restrict__ src,
                                                         int64_t lds, char* __restrict__ dst, int64_t ldk,
                                                         int64_t ps, double* __restrict__ mu)
{
    __shared__ double buf[4][1024 + 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t r = (int64_t)blockIdx.x * 4 + wave;
    if (r >= rows) return;
    const double* row = src + r * lds;
    double* b = buf[wave];
    double m = 0.0;
    if constexpr (NCH > 0) {
        double v[NCH][16];
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch)
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int64_t e = ch * 1024 + t * 64 + lane;
                v[ch][t] = load_or_zero(row, e, e < k);
            }
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch)
#pragma unroll
            for (int t = 0; t < 16; ++t) m = absmax(m, v[ch][t]);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = absmax(m, __shfl_xor(m, off, 64));
        m = finish_mu(m);
        if (lane == 0) mu[r] = m;
        const double inv = 1.0 / m;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch)
            if (ch * 1024 < ldk) slice_chunk<S>(b, v[ch], lane, ch * 1024, r, ldk, ps, inv, dst);
    } else {
        for (int64_t j0 = 0; j0 < k; j0 += 8 * 64) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = load_or_zero(row, j0 + 64 * u + lane, j0 + 64 * u + lane < k);
#pragma unroll
            for (int u = 0; u < 8; ++u) m = absmax(m, v[u]);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = absmax(m, __shfl_xor(m, off, 64));
        m = finish_mu(m);
        if (lane == 0) mu[r] = m;
        const double inv = 1.0 / m;
        for (int64_t base = 0; base < ldk; base += 1024) {
            double v[16];
#pragma unroll
            for (int t = 0; t < 16; ++t) v[t] = load_or_zero(row, base + t * 64 + lane, base + t * 64 + lane < k);
            slice_chunk<S>(b, v, lane, base, r, ldk, ps, inv, dst);
        }
    }
}

// Layout 1: element (r, k) at src[r + k lds] (the rows of the operand are
// the rows of a column-major matrix).  Pass 1: per-row maxima over k chunks
// (thread per row, coalesced along r), merged with an integer atomicMax on
// the bit patterns (non-negative doubles, NaN above +Inf, order as
// integers).  Pass 2: 64 x 64 tiles transposed through LDS.
__global__ __launch_bounds__(256) void rowmax_cm_kernel(int64_t rows, int64_t k, const double* __restrict__ src,
                                                        int64_t lds, int64_t kchunk,
                                                        unsigned long long* __restrict__ bits)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= rows) return;
    const int64_t k0 = (int64_t)blockIdx.y * kchunk;
    const int64_t k1 = k0 + kchunk < k ? k0 + kchunk : k;
    double m = 0.0;
    for (int64_t j0 = k0; j0 < k1; j0 += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = load_or_zero(src, r + (j0 + u) * lds, j0 + u < k1);
#pragma unroll
        for (int u = 0; u < 8; ++u) m = absmax(m, v[u]);
    }
    unsigned long long b;
    memcpy(&b, &m, 8);
    if (b) atomicMax(bits + r, b);
}

__global__ void zero_kernel(int64_t count, double* p)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void finish_mu_kernel(int64_t rows, double* mu)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x)
        mu[r] = finish_mu(mu[r]);
}

template <int S>
__global__ __launch_bounds__(256) void slice_cm_kernel(int64_t rows, int64_t k, const double* __restrict__ src,
                                                       int64_t lds, char* __restrict__ dst, int64_t ldk, int64_t ps,
                                                       const double* __restrict__ mu)
{
    __shared__ double tile[64][65];
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int idx = tid + 256 * e, rr = idx & 63, kk = idx >> 6;
        const int64_t r = r0 + rr, kg = k0 + kk;
        tile[kk][rr] = load_or_zero(src, r + kg * lds, r < rows && kg < k);
    }
    __syncthreads();
    const int rr = tid >> 2, c = tid & 3;
    const int64_t r = r0 + rr;
    if (r >= rows || k0 + 16 * c >= ldk) return;
    double x[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) x[j] = tile[16 * c + j][rr];
    u4 o[S];
    slice16<S>(x, 1.0 / mu[r], o);
#pragma unroll
    for (int p = 0; p < S; ++p) *reinterpret_cast<u4*>(dst + p * ps + r * ldk + k0 + 16 * c) = o[p];
}

// Slices -> FP64 (the value the slices represent), for tests.
__global__ void unslice_kernel(int s, int64_t rows, int64_t k, const signed char* __restrict__ src, int64_t ldk,
                               int64_t ps, const double* __restrict__ mu, double* __restrict__ dst, int64_t ldd)
{
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < rows * k;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = idx / k, kk = idx % k;
        double v = (double)src[(s - 1) * ps + r * ldk + kk];
        for (int p = s - 2; p >= 0; --p) v = (double)src[p * ps + r * ldk + kk] + v * kInv254;
        dst[r * ldd + kk] = (v * (1.0 / 127.0)) * mu[r];
    }
}

struct S8Args {
    int64_t M, N, K;           // K a multiple of SK, within both operands' row pitch
    double alpha, beta;
    const char* A;
    int64_t lda, psA;
    const double* muA;
    const char* B;
    int64_t ldb, psB;
    const double* muB;
    double* C;
    int64_t ldc;
    int lower;                 // compute the tiles on or below the diagonal, store m >= n only
    int tri;                   // bit 0: A[m][k] = 0 for k > m + offA; bit 2: B[n][k] = 0 for k > n + offB
    int64_t offA, offB;
    int packed;                // 1: C is the packed lower row band (row0, col0, pbase, pdiag)
    int64_t row0, col0, pbase;
    double pdiag;
    int64_t tm, tn;
    int64_t gm, gn, gtm, gtmp, gtn, ngroups;
};

// Staging: thread t holds 16-byte chunk t % 4 of row t / 4 of every slice
// of the A and B tiles (one slab).  Named registers (no arrays), as arrays
// of vectors indexed in unrolled loops have ended up in scratch memory.
#define CFMI_S8_EACH(X) X(0) X(1) X(2) X(3) X(4)
template <int S>
struct Stage8 {
    u4 a0, a1, a2, a3, a4, b0, b1, b2, b3, b4;

    __device__ __forceinline__ void load(const S8Args& g, int64_t m0, int64_t n0, int64_t k0, int tid)
    {
        const int sr = tid >> 2, ck = (tid & 3) * 16;
        const int64_t ra = m0 + sr < g.M ? sr : g.M - 1 - m0;
        const int64_t rb = n0 + sr < g.N ? sr : g.N - 1 - n0;
        const char* pa = g.A + m0 * g.lda + k0;
        const char* pb = g.B + n0 * g.ldb + k0;
        const unsigned oa = (unsigned)(ra * g.lda + ck), ob = (unsigned)(rb * g.ldb + ck);
#define CFMI_S8_LD(p)                                                               \
        if constexpr (S > p) {                                                      \
            a##p = *reinterpret_cast<const u4*>(pa + p * g.psA + oa);               \
            b##p = *reinterpret_cast<const u4*>(pb + p * g.psB + ob);               \
        }
        CFMI_S8_EACH(CFMI_S8_LD)
#undef CFMI_S8_LD
    }
    __device__ __forceinline__ void store(char (*As)[BT][SK], char (*Bs)[BT][SK], int tid) const
    {
        const int row = tid >> 2, kc = ((tid & 3) ^ swz(row)) * 16;
#define CFMI_S8_ST(p)                                                               \
        if constexpr (S > p) {                                                      \
            *reinterpret_cast<u4*>(&As[p][row][kc]) = a##p;                         \
            *reinterpret_cast<u4*>(&Bs[p][row][kc]) = b##p;                         \
        }
        CFMI_S8_EACH(CFMI_S8_ST)
#undef CFMI_S8_ST
    }
};
#undef CFMI_S8_EACH

// Per-lane source of the asynchronous slab copies: wave w moves rows
// 16 w .. 16 w + 15 of every slice plane of both operands, one 1 KB
// global_load_lds_dwordx4 per plane and operand (lane l lands at byte 16 l
// of the wave's LDS destination: row 16 w + l / 4, chunk l % 4, so it reads
// the global chunk that the XOR swizzle stores there).  Rows past the
// matrix edge repeat the last row; the epilogue drops them.
struct DmaSrc {
    const char* a;             // tile bases (wave-uniform)
    const char* b;
    unsigned oa, ob;           // this lane's byte offsets from them

    __device__ __forceinline__ DmaSrc(const S8Args& g, int64_t m0, int64_t n0, int wave, int lane)
    {
        const int r = 16 * wave + (lane >> 2);
        const int cg = ((lane & 3) ^ swz(r)) * 16;
        const int64_t ra = m0 + r < g.M ? r : g.M - 1 - m0;
        const int64_t rb = n0 + r < g.N ? r : g.N - 1 - n0;
        a = g.A + m0 * g.lda;
        b = g.B + n0 * g.ldb;
        oa = (unsigned)(ra * g.lda + cg);
        ob = (unsigned)(rb * g.ldb + cg);
    }
};

typedef __attribute__((address_space(3))) void lds_void;

template <int S>
__device__ __forceinline__ void dma_slab(const S8Args& g, const DmaSrc& src, int64_t k0, char (*As)[BT][SK],
                                         char (*Bs)[BT][SK], int wave)
{
#pragma unroll
    for (int p = 0; p < S; ++p) {
        const char* pa = src.a + (p * g.psA + k0);
        const char* pb = src.b + (p * g.psB + k0);
        __builtin_amdgcn_global_load_lds(pa + src.oa, (lds_void*)&As[p][16 * wave][0], 16, 0, 0);
        __builtin_amdgcn_global_load_lds(pb + src.ob, (lds_void*)&Bs[p][16 * wave][0], 16, 0, 0);
    }
}

// The copies this wave issued have landed in LDS.
__device__ __forceinline__ void dma_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// TN = 1: the products are issued with the operands exchanged, so each
// accumulator holds the transposed 16 x 16 tile (the same exact integers):
// consecutive lanes then own consecutive rows of C, and a column-major C is
// read and written in 128-byte runs.  TN = 0 keeps consecutive columns per
// lane (contiguous in the packed row band).
template <int S, int TN>
__device__ __forceinline__ void slab_mfma(const char (*As)[BT][SK], const char (*Bs)[BT][SK], i4 (&acc)[S][2][4],
                                          int wm, int wn, int lm, int lk)
{
    const int kk = 16 * (lk ^ swz(lm));   // wm, wn, 16 i, 16 j are multiples of 16
    i4 a[2][S];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int p = 0; p < S; ++p) a[i][p] = *reinterpret_cast<const i4*>(&As[p][wm + 16 * i + lm][kk]);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        i4 b[S];
#pragma unroll
        for (int p = 0; p < S; ++p) b[p] = *reinterpret_cast<const i4*>(&Bs[p][wn + 16 * j + lm][kk]);
        // p outer: consecutive products go to different accumulators, so a
        // product waits on the one before it into the same accumulator only
        // after 2 (S - p) others have issued
#pragma unroll
        for (int p = 0; p < S; ++p)
#pragma unroll
            for (int d = p; d < S; ++d)
#pragma unroll
                for (int i = 0; i < 2; ++i)
                    acc[d][i][j] = TN ? __builtin_amdgcn_mfma_i32_16x16x64_i8(b[d - p], a[i][p], acc[d][i][j], 0, 0, 0)
                                      : __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i][p], b[d - p], acc[d][i][j], 0, 0, 0);
    }
}

// K range of tile (m0, n0): the triangular trims shorten it to the last
// k that can be nonzero, rounded up to a whole slab (the operands hold zeros
// there, and up to their padded row pitch).
__device__ __forceinline__ int64_t tile_kend(const S8Args& g, int64_t m0, int64_t n0)
{
    int64_t k = g.K;
    if (g.tri & 1) k = min(k, (m0 + BT + g.offA + SK - 1) / SK * SK);
    if (g.tri & 4) k = min(k, (n0 + BT + g.offB + SK - 1) / SK * SK);
    return k;
}

// DB = 1: two LDS stages filled by asynchronous global -> LDS copies
// (global_load_lds_dwordx4, no registers in between), one barrier per
// slab: the copy of slab k + 1 into one stage is in flight while slab k is
// multiplied out of the other.  DB = 0: one stage, staged through registers
// and stored between two barriers.
template <int S, int DB, int TN>
__global__ __launch_bounds__(NTH, 1) void gemm_s8_kernel(S8Args g)
{
    __shared__ __attribute__((aligned(16))) char As[S][BT][SK];
    __shared__ __attribute__((aligned(16))) char Bs[S][BT][SK];
    __shared__ __attribute__((aligned(16))) char As1[DB ? S : 1][DB ? BT : 1][SK];
    __shared__ __attribute__((aligned(16))) char Bs1[DB ? S : 1][DB ? BT : 1][SK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int64_t z, ti, tj;
    if (!group_tile(g, blockIdx.x, z, ti, tj)) return;
    const int64_t m0 = ti * BT, n0 = tj * BT;
    if (g.lower && n0 > m0 + BT - 1) return;
    if (g.packed && g.col0 + n0 > g.row0 + m0 + BT - 1) return;   // band tile above the diagonal
    const int64_t K = tile_kend(g, m0, n0);
    const int wm = (wave & 3) * 32, wn = (wave >> 2) * 64;
    const int lm = lane & 15, lk = lane >> 4;

    i4 acc[S][2][4];
#pragma unroll
    for (int d = 0; d < S; ++d)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[d][i][j] = i4{0, 0, 0, 0};

    if (K <= 0) {
        // the trims leave no k: the product is zero
    } else if constexpr (DB) {
        const int uw = __builtin_amdgcn_readfirstlane(wave);   // wave-uniform: LDS targets in scalars
        const DmaSrc src(g, m0, n0, uw, lane);
        dma_slab<S>(g, src, 0, As, Bs, uw);
        dma_wait();
        __syncthreads();
        for (int64_t k0 = 0;; k0 += SK) {
            const bool odd = (k0 / SK) & 1;
            const bool more = k0 + SK < K;
            if (more) {
                if (odd) dma_slab<S>(g, src, k0 + SK, As, Bs, uw);
                else dma_slab<S>(g, src, k0 + SK, As1, Bs1, uw);
            }
            slab_mfma<S, TN>(odd ? As1 : As, odd ? Bs1 : Bs, acc, wm, wn, lm, lk);
            if (!more) break;
            dma_wait();
            __syncthreads();
        }
    } else {
        Stage8<S> st;
        st.load(g, m0, n0, 0, tid);
        int64_t k0 = 0;
        for (; k0 + SK < K; k0 += SK) {
            __syncthreads();
            st.store(As, Bs, tid);
            __syncthreads();
            st.load(g, m0, n0, k0 + SK, tid);
            slab_mfma<S, TN>(As, Bs, acc, wm, wn, lm, lk);
        }
        __syncthreads();
        st.store(As, Bs, tid);
        __syncthreads();
        slab_mfma<S, TN>(As, Bs, acc, wm, wn, lm, lk);
    }

    if constexpr (TN) {
        // lane (lm, lk), register r: row m0 + wm + 16 i + lm, column n0 + wn + 16 j + 4 lk + r.
        // Per column block j the eight stored values of C are fetched together
        // (unconditional loads at clamped addresses), then combined and stored.
        double sa[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int64_t gm = m0 + wm + 16 * i + lm;
            const double mu = g.muA[gm < g.M ? gm : g.M - 1];
            sa[i] = gm < g.M ? g.alpha * mu : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double sb[4], cold[2][4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gn = n0 + wn + 16 * j + 4 * lk + r;
                const double mu = g.muB[gn < g.N ? gn : g.N - 1];
                sb[r] = gn < g.N ? mu : 0.0;
            }
            if (g.beta != 0.0) {
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int64_t gm = m0 + wm + 16 * i + lm, gn = n0 + wn + 16 * j + 4 * lk + r;
                        const bool ok = gm < g.M && gn < g.N && !(g.lower && gn > gm);
                        cold[i][r] = g.C[ok ? gm + gn * g.ldc : 0];
                    }
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t gm = m0 + wm + 16 * i + lm, gn = n0 + wn + 16 * j + 4 * lk + r;
                    if (gm >= g.M || gn >= g.N || (g.lower && gn > gm)) continue;
                    double v = (double)acc[S - 1][i][j][r];
#pragma unroll
                    for (int d = S - 2; d >= 0; --d) v = (double)acc[d][i][j][r] + v * kInv254;
                    v = (v * kInv127sq) * (sa[i] * sb[r]);
                    g.C[gm + gn * g.ldc] = g.beta == 0.0 ? v : v + g.beta * cold[i][r];
                }
        }
        return;
    }
    double sa[2][4], sb[4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t gm = m0 + wm + 16 * i + 4 * lk + r;
            sa[i][r] = gm < g.M ? g.alpha * g.muA[gm] : 0.0;
        }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int64_t gn = n0 + wn + 16 * j + lm;
        sb[j] = gn < g.N ? g.muB[gn] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gm = m0 + wm + 16 * i + 4 * lk + r;
                const int64_t gn = n0 + wn + 16 * j + lm;
                if (gm >= g.M || gn >= g.N || (g.lower && gn > gm)) continue;
                double v = (double)acc[S - 1][i][j][r];
#pragma unroll
                for (int d = S - 2; d >= 0; --d) v = (double)acc[d][i][j][r] + v * kInv254;
                v = (v * kInv127sq) * (sa[i][r] * sb[j]);
                double* p;
                if (g.packed) {
                    const int64_t ii = g.row0 + gm, cc = g.col0 + gn;
                    if (cc > ii) continue;
                    if (cc == ii) v = v * g.pdiag;
                    p = g.C + (ii * (ii + 1) / 2 + cc - g.pbase);
                } else {
                    p = g.C + gm + gn * g.ldc;
                }
                *p = g.beta == 0.0 ? v : v + g.beta * *p;
            }
}

// C(n, m) = C(m, n) for m > n, 64 x 64 tiles through LDS (both sides coalesced).
__global__ __launch_bounds__(256) void mirror_lower_s8_kernel(int64_t n, double* C, int64_t ldc)
{
    __shared__ double t[64][65];
    const int64_t bi = blockIdx.x, bj = blockIdx.y;   // source tile (rows bi, columns bj), bi >= bj
    if (bj > bi) return;
    const int tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int idx = tid + 256 * e, rr = idx & 63, cc = idx >> 6;
        const int64_t r = bi * 64 + rr, c = bj * 64 + cc;
        t[cc][rr] = load_or_zero((const double*)C, r + c * ldc, r < n && c < n);
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int idx = tid + 256 * e, rr = idx & 63, cc = idx >> 6;
        // destination (row bj*64 + rr, column bi*64 + cc) = source (bi*64 + cc, bj*64 + rr)
        const int64_t r = bj * 64 + rr, c = bi * 64 + cc;
        if (r < n && c < n && c > r) C[r + c * ldc] = t[rr][cc];
    }
}

// dst (n x m, column-major) = src^T for src (m x n, column-major), 64 x 64
// tiles through LDS.
__global__ __launch_bounds__(256) void transpose_kernel(int64_t m, int64_t n, const double* __restrict__ src,
                                                        int64_t lds, double* __restrict__ dst, int64_t ldd)
{
    __shared__ double t[64][65];
    const int64_t bi = blockIdx.x, bj = blockIdx.y;   // source rows bi*64.., columns bj*64..
    const int tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int idx = tid + 256 * e, rr = idx & 63, cc = idx >> 6;
        const int64_t r = bi * 64 + rr, c = bj * 64 + cc;
        t[cc][rr] = load_or_zero(src, r + c * lds, r < m && c < n);
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int idx = tid + 256 * e, rr = idx & 63, cc = idx >> 6;
        const int64_t r = bj * 64 + rr, c = bi * 64 + cc;   // dst (r, c) = src (c, r)
        if (r < n && c < m) dst[r + c * ldd] = t[rr][cc];
    }
}

inline dim3 grid_of(int64_t tot)
{
    int64_t b = (tot + 255) / 256;
    return dim3((unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536));
}

inline int64_t round64(int64_t k) { return (k + SK - 1) / SK * SK; }

template <int S>
int slice_t(cfmi_ctx* c, int layout, int64_t rows, int64_t k, const double* src, int64_t lds, char* dst,
            int64_t ldk, int64_t ps, double* mu)
{
    if (layout == 0) {
        // rows of up to 4096 values are read once (held in registers)
        const dim3 grid((unsigned)((rows + 3) / 4));
        if (ldk <= 1024)
            hipLaunchKernelGGL((slice_rows_kernel<S, 1>), grid, dim3(256), 0, c->stream, rows, k, src, lds, dst, ldk,
                               ps, mu);
        else if (ldk <= 2048)
            hipLaunchKernelGGL((slice_rows_kernel<S, 2>), grid, dim3(256), 0, c->stream, rows, k, src, lds, dst, ldk,
                               ps, mu);
        else if (ldk <= 4096)
            hipLaunchKernelGGL((slice_rows_kernel<S, 4>), grid, dim3(256), 0, c->stream, rows, k, src, lds, dst, ldk,
                               ps, mu);
        else
            hipLaunchKernelGGL((slice_rows_kernel<S, 0>), grid, dim3(256), 0, c->stream, rows, k, src, lds, dst, ldk,
                               ps, mu);
        CFMI_LAUNCH_CHECK(c);
        return CFMI_OK;
    }
    hipLaunchKernelGGL(zero_kernel, grid_of(rows), dim3(256), 0, c->stream, rows, mu);
    CFMI_LAUNCH_CHECK(c);
    const int64_t kchunk = 512;
    const int64_t nkc = (k + kchunk - 1) / kchunk;
    if (nkc > 0) {
        hipLaunchKernelGGL(rowmax_cm_kernel, dim3((unsigned)((rows + 255) / 256), (unsigned)nkc), dim3(256), 0,
                           c->stream, rows, k, src, lds, kchunk, reinterpret_cast<unsigned long long*>(mu));
        CFMI_LAUNCH_CHECK(c);
    }
    hipLaunchKernelGGL(finish_mu_kernel, grid_of(rows), dim3(256), 0, c->stream, rows, mu);
    CFMI_LAUNCH_CHECK(c);
    hipLaunchKernelGGL((slice_cm_kernel<S>), dim3((unsigned)((rows + 63) / 64), (unsigned)(ldk / 64)), dim3(256), 0,
                       c->stream, rows, k, src, lds, dst, ldk, ps, mu);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

hipEvent_t s8_event(cfmi_ctx* c)
{
    if (!c->s8_event_pool.empty()) {
        hipEvent_t e = c->s8_event_pool.back();
        c->s8_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

template <int S>
int gemm_s8_launch(cfmi_ctx* c, S8Args g, bool db)
{
    g.tm = (g.M + BT - 1) / BT;
    g.tn = (g.N + BT - 1) / BT;
    g.gm = 2;
    g.gn = 8;
    g.gtm = (g.tm + g.gm - 1) / g.gm;
    g.gtn = (g.tn + g.gn - 1) / g.gn;
    g.gtmp = g.gtm | 1;
    g.ngroups = g.gtmp * g.gtn;
    const int64_t nwg = (g.ngroups + 7) / 8 * 8 * g.gm * g.gn;
    if (nwg > 0x7fffffffLL) return cfmi_fail(c, CFMI_EINVAL, "gemm_s8: too many tiles");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->s8_stats) {
        e0 = s8_event(c);
        e1 = s8_event(c);
        if (!e0 || !e1) return cfmi_fail(c, CFMI_EHIP, "gemm_s8: hipEventCreate failed");
        CFMI_HIP(c, hipEventRecord(e0, c->stream));
    }
    // the transposed accumulator layout for a column-major C, the plain one for the packed band
    if (db && !g.packed)
        hipLaunchKernelGGL((gemm_s8_kernel<S, 1, 1>), dim3((unsigned)nwg), dim3(NTH), 0, c->stream, g);
    else if (db)
        hipLaunchKernelGGL((gemm_s8_kernel<S, 1, 0>), dim3((unsigned)nwg), dim3(NTH), 0, c->stream, g);
    else if (!g.packed)
        hipLaunchKernelGGL((gemm_s8_kernel<S, 0, 1>), dim3((unsigned)nwg), dim3(NTH), 0, c->stream, g);
    else
        hipLaunchKernelGGL((gemm_s8_kernel<S, 0, 0>), dim3((unsigned)nwg), dim3(NTH), 0, c->stream, g);
    CFMI_LAUNCH_CHECK(c);
    if (c->s8_stats) {
        CFMI_HIP(c, hipEventRecord(e1, c->stream));
        const double entries = (double)g.M * (double)g.N - (g.lower ? 0.5 * (double)g.N * (double)(g.N - 1) : 0.0);
        const double f = 2.0 * (double)g.K * entries;
        c->s8_launches.push_back({e0, e1, f, S});
    }
    return CFMI_OK;
}

bool aligned16(const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; }

}  // namespace

extern "C" {

int64_t cfmi_s8_kmax(int s)
{
    if (s < 1 || s > 8) return 0;
    return ((int64_t)0x7fffffff / (16129LL * s)) / SK * SK;
}

int cfmi_slice8(cfmi_ctx* c, int s, int layout, int64_t rows, int64_t k, const double* src, int64_t lds, void* dst,
                int64_t ldk, int64_t ps, double* mu)
{
    if (!c) return CFMI_EINVAL;
    if (rows < 0 || k < 0) return cfmi_fail(c, CFMI_EINVAL, "slice8: negative size");
    if (rows == 0) return CFMI_OK;
    if (s < 3 || s > 5) return cfmi_fail(c, CFMI_EINVAL, "slice8: %d slices (3 to 5 are built)", s);
    if (layout != 0 && layout != 1) return cfmi_fail(c, CFMI_EINVAL, "slice8: layout must be 0 or 1");
    if (ldk < round64(k) || ldk % SK != 0)
        return cfmi_fail(c, CFMI_EINVAL, "slice8: row pitch %lld must be a multiple of 64 and >= %lld",
                         (long long)ldk, (long long)round64(k));
    if (ps < rows * ldk || ps % 16 != 0) return cfmi_fail(c, CFMI_EINVAL, "slice8: slice planes overlap");
    if (!aligned16(dst)) return cfmi_fail(c, CFMI_EINVAL, "slice8: destination not 16-byte aligned");
    if ((layout == 0 && lds < k) || (layout == 1 && lds < rows))
        return cfmi_fail(c, CFMI_EINVAL, "slice8: source leading dimension too small");
    char* d = (char*)dst;
    switch (s) {
        case 3: return slice_t<3>(c, layout, rows, k, src, lds, d, ldk, ps, mu);
        case 4: return slice_t<4>(c, layout, rows, k, src, lds, d, ldk, ps, mu);
        default: return slice_t<5>(c, layout, rows, k, src, lds, d, ldk, ps, mu);
    }
}

int cfmi_transpose(cfmi_ctx* c, int64_t m, int64_t n, const double* src, int64_t lds, double* dst, int64_t ldd)
{
    if (!c) return CFMI_EINVAL;
    if (m <= 0 || n <= 0) return CFMI_OK;
    if (lds < m || ldd < n) return cfmi_fail(c, CFMI_EINVAL, "transpose: leading dimension too small");
    if ((m + 63) / 64 > 0x7fffffffLL || (n + 63) / 64 > 65535)
        return cfmi_fail(c, CFMI_EINVAL, "transpose: matrix too large");
    hipLaunchKernelGGL(transpose_kernel, dim3((unsigned)((m + 63) / 64), (unsigned)((n + 63) / 64)), dim3(256), 0,
                       c->stream, m, n, src, lds, dst, ldd);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_unslice8(cfmi_ctx* c, int s, int64_t rows, int64_t k, const void* src, int64_t ldk, int64_t ps,
                  const double* mu, double* dst, int64_t ldd)
{
    if (!c) return CFMI_EINVAL;
    if (rows <= 0 || k <= 0) return CFMI_OK;
    if (s < 1 || s > 8 || ldk < k || ldd < k) return cfmi_fail(c, CFMI_EINVAL, "unslice8: bad arguments");
    hipLaunchKernelGGL(unslice_kernel, grid_of(rows * k), dim3(256), 0, c->stream, s, rows, k,
                       (const signed char*)src, ldk, ps, mu, dst, ldd);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_s8_stats_enable(cfmi_ctx* c, int on)
{
    if (!c) return CFMI_EINVAL;
    CFMI_HIP(c, hipStreamSynchronize(c->stream));
    for (auto& l : c->s8_launches) {
        c->s8_event_pool.push_back(l.start);
        c->s8_event_pool.push_back(l.stop);
    }
    c->s8_launches.clear();
    c->s8_stats = on != 0;
    return CFMI_OK;
}

int cfmi_s8_stats_read(cfmi_ctx* c, int64_t* launches, double* ms, double* flops, double* products)
{
    if (!c) return CFMI_EINVAL;
    CFMI_HIP(c, hipStreamSynchronize(c->stream));
    double t = 0.0, f = 0.0, p = 0.0;
    for (auto& l : c->s8_launches) {
        float dt = 0.0f;
        CFMI_HIP(c, hipEventElapsedTime(&dt, l.start, l.stop));
        t += dt;
        f += l.flops;
        p += l.flops * l.s * (l.s + 1) / 2;
    }
    if (launches) *launches = (int64_t)c->s8_launches.size();
    if (ms) *ms = t;
    if (flops) *flops = f;
    if (products) *products = p;
    return CFMI_OK;
}

int cfmi_gemm_s8(cfmi_ctx* c, int s, int64_t m, int64_t n, int64_t k, double alpha, const void* A, int64_t lda,
                 int64_t psA, const double* muA, int64_t off_a, const void* B, int64_t ldb, int64_t psB,
                 const double* muB, int64_t off_b, double beta, double* C, int64_t ldc, int flags, int64_t row0,
                 int64_t col0, int64_t pbase, double pdiag)
{
    if (!c) return CFMI_EINVAL;
    if (m < 0 || n < 0 || k < 0) return cfmi_fail(c, CFMI_EINVAL, "gemm_s8: negative size");
    if (m == 0 || n == 0) return CFMI_OK;
    if (s < 3 || s > 5) return cfmi_fail(c, CFMI_EINVAL, "gemm_s8: %d slices (3 to 5 are built)", s);
    const int64_t kp = round64(k);
    if (lda < kp || ldb < kp || lda % 16 || ldb % 16)
        return cfmi_fail(c, CFMI_EINVAL, "gemm_s8: slice row pitch must be a multiple of 16 and >= %lld",
                         (long long)kp);
    if (psA < m * lda || psB < n * ldb || psA % 16 || psB % 16)
        return cfmi_fail(c, CFMI_EINVAL, "gemm_s8: slice planes overlap or are misaligned");
    if (!aligned16(A) || !aligned16(B)) return cfmi_fail(c, CFMI_EINVAL, "gemm_s8: operands not 16-byte aligned");
    const bool lower = flags & (1 | 32), mirror = flags & 1, packed = flags & 16;
    if (!packed && ldc < m) return cfmi_fail(c, CFMI_EINVAL, "gemm_s8: ldc < m");
    if (mirror && m != n) return cfmi_fail(c, CFMI_EINVAL, "gemm_s8: a symmetric result needs m == n");
    if (lower && (n > m || packed))
        return cfmi_fail(c, CFMI_EINVAL, "gemm_s8: a lower result needs n <= m and a dense C");
    if (packed && (row0 < 0 || col0 < 0 || pbase < 0 || pbase > row0 * (row0 + 1) / 2))
        return cfmi_fail(c, CFMI_EINVAL, "gemm_s8: packed row band starts after row %lld", (long long)row0);
    if (k == 0) return cfmi_fail(c, CFMI_EINVAL, "gemm_s8: k must be positive");
    // K chunks within the exact int32 range, added in FP64 (beta = 1 after the first)
    const int64_t kmax = cfmi_s8_kmax(s);
    double b = beta;
    for (int64_t k0 = 0; k0 < kp; k0 += kmax) {
        const int64_t kk = std::min<int64_t>(kmax, kp - k0);
        S8Args g = {};
        g.M = m; g.N = n; g.K = kk;
        g.alpha = alpha; g.beta = b;
        g.A = (const char*)A + k0; g.lda = lda; g.psA = psA; g.muA = muA;
        g.B = (const char*)B + k0; g.ldb = ldb; g.psB = psB; g.muB = muB;
        g.C = C; g.ldc = ldc; g.lower = lower ? 1 : 0;
        // triangular K-trims, relative to this chunk's first k
        g.tri = ((flags & 4) ? 1 : 0) | ((flags & 8) ? 4 : 0);
        g.offA = off_a - k0; g.offB = off_b - k0;
        g.packed = packed ? 1 : 0; g.row0 = row0; g.col0 = col0; g.pbase = pbase; g.pdiag = pdiag;
        // two LDS stages filled asynchronously; flag 2: one stage staged through registers
        const bool db = !(flags & 2);
        int rc = s == 3 ? gemm_s8_launch<3>(c, g, db) : s == 4 ? gemm_s8_launch<4>(c, g, db)
                                                              : gemm_s8_launch<5>(c, g, db);
        if (rc) return rc;
        b = 1.0;
    }
    if (mirror) {
        const unsigned t = (unsigned)((n + 63) / 64);
        hipLaunchKernelGGL(mirror_lower_s8_kernel, dim3(t, t), dim3(256), 0, c->stream, n, C, ldc);
        CFMI_LAUNCH_CHECK(c);
    }
    return CFMI_OK;
}

}  // extern "C"

int cfmi_s8_update_lower(cfmi_ctx* c, int s, int64_t m, int64_t w, int64_t k, double alpha, const double* P,
                         int64_t ldp, double* C, int64_t ldc)
{
    if (m <= 0 || w <= 0 || k <= 0) return CFMI_OK;
    if (w > m) return cfmi_fail(c, CFMI_EINVAL, "s8 update: %lld columns > %lld rows", (long long)w, (long long)m);
    const int64_t ldk = round64(k), ps = m * ldk;
    char* ws = (char*)cfmi_workspace(c, (size_t)(s * ps + 8 * m));
    if (!ws) return cfmi_fail(c, CFMI_ENOMEM, "s8 update: workspace of %lld bytes", (long long)(s * ps + 8 * m));
    double* mu = reinterpret_cast<double*>(ws + s * ps);    // s ps is a multiple of 64
    int rc = cfmi_slice8(c, s, 1, m, k, P, ldp, ws, ldk, ps, mu);
    if (rc) return rc;
    // B is the first w rows of the same slices (the same row scales)
    return cfmi_gemm_s8(c, s, m, w, k, alpha, ws, ldk, ps, mu, 0, ws, ldk, ps, mu, 0, 1.0, C, ldc, 32, 0, 0, 0, 1.0);
}
