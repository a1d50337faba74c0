// FP32-accurate GEMM on BF16 MFMA (gfx950 v_mfma_f32_16x16x32_bf16).
//
// Every operand value x is held as three BF16 pieces x = x0 + x1 + x2
// (x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1): 24 significant
// bits, the precision of FP32), stored as three planes.  A product sums the
// six piece products that reach FP32 precision,
//     a0 b0 + a0 b1 + a1 b0 + a0 b2 + a1 b1 + a2 b0,
// each exact in FP32, accumulated by the MFMA in FP32; the dropped terms are
// below 2^-24 of |a| |b|.  At the MI355X's 16:1 ratio of dense BF16 to FP32
// MFMA throughput this is up to 2.7x the FP32 MFMA rate at FP32 accuracy.
//
// Operands are "k-contiguous": A[m][k] at A + m lda + k, B[n][k] at
// B + n ldb + k (one plane; planes pA / pB elements apart), so a 32-deep
// K-slab of a tile is a 16-byte-vector copy into LDS ([row][k]) and each
// MFMA operand (8 consecutive k) one 16-byte LDS read.  Tiles 128 x 128
// (4 waves of 64 x 64) or 256 x 256 (8 waves of 64 x 128); the next slab is
// prefetched into registers during the MFMAs.
//
// Outputs: FP32 column-major (mode 0), three BF16 planes row-major
// C[m ldc + n] (mode 1, the next product's operand), or the FP32 packed
// lower row band of the whitening (mode 2: element (i, c <= i) at
// i (i + 1) / 2 + c - pbase, diagonal scaled by pdiag).
#include "cfmi_internal.h"

#include <algorithm>
#include <cstring>

namespace {

typedef unsigned short bf16;
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));

// LDS operand rows are 32 bf16 (64 bytes, unpadded); the 16-byte chunk c of
// row r sits at chunk c ^ swz(r).  With swz(r) = ((r >> 1) ^ (r >> 2)) & 3
// the staging stores (8 lanes: 2 rows x 4 chunks, 128 contiguous bytes) and
// the MFMA operand reads (16 lanes: 16 rows at one chunk) both hit distinct
// banks, for 32 and for 64 banks of 4 bytes.  swz(r) depends on r mod 16
// only.
constexpr int TK = 32;
constexpr int LROW = TK;
__device__ __forceinline__ int swz(int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

// Block configurations: BT x BT output tiles, 2 BT threads, waves of 64 x BT/2
// (BT / 64 along M, 2 along N).  BT = 256 halves the operand bytes per flop
// and gives every wave 4 x 8 MFMA tiles (36 LDS reads per 192 MFMAs); its
// 96 KB of LDS allow one block per CU (2 waves per SIMD).  BT = 128 (two
// blocks per CU) serves grids too small to fill the CUs with 256-tiles.
template <int BT> struct Cfg {
    static constexpr int NT = 2 * BT, HALF = BT / 2, WMC = BT / 64, FN = BT / 32, MINB = BT == 128 ? 2 : 1;
};

__device__ __forceinline__ bf16 f2bf(float f)
{
    unsigned u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 v)
{
    unsigned u = (unsigned)v << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
__device__ __forceinline__ void split3(double x, bf16& h, bf16& m, bf16& l)
{
    h = f2bf((float)x);
    double r = x - (double)bf2f(h);
    m = f2bf((float)r);
    r -= (double)bf2f(m);
    l = f2bf((float)r);
}

struct B3Args {
    int64_t M, N, K;
    float alpha, beta;
    const bf16* A;
    int64_t lda, sA, pA;
    const bf16* B;
    int64_t ldb, sB, pB;
    void* C;
    int64_t ldc, sC, pC;
    int tri;                  // bit 0: A[m][k] = 0 for k > m + offA; bit 2: B[n][k] = 0 for k > n + offB
    int64_t offA, offB;
    int64_t row0, col0, pbase;
    float pdiag;
    int64_t tm, tn, tmp;      // row / column tiles; tmp: enumerated rows (tm, or tm + 1)
    int vecA, vecB;           // 16-byte vector loads allowed
    int grouped;              // 1: XCD-grouped tile order (group_tile), 0: tile_order over blockIdx.z
    int64_t gm, gn;           // group of gm x gn tiles
    int64_t gtm, gtmp, gtn;   // group rows (gtmp: enumerated, odd), group columns
    int64_t ngroups;          // gtmp gtn batch
};

// Tile order: column tiles of B in turn, the row tiles of A in each; with a
// triangular K-trim the heaviest tiles (longest K) are issued first.
// Workgroups go to the 8 XCDs round-robin in launch order, so with an even
// row count every XCD would own a fixed subset of row tiles; under an A-trim
// or a lower-triangle output the work of a row tile grows with its index and
// that subset decides the XCD's share (measured: 2.4x more work on one XCD
// than on another for 16 rows of 256-tiles, and the whole launch as slow as
// the untrimmed product).  The enumeration then counts tmp = tm + 1 rows (the
// extra one exits at once), so the row-to-XCD assignment rotates from column
// to column and evens out.
__device__ __forceinline__ bool tile_order(int64_t b, int64_t tm, int64_t tmp, int64_t tn, int tri, int64_t& ti,
                                           int64_t& tj)
{
    ti = b % tmp;
    tj = b / tmp;
    if (ti >= tm) return false;
    if (tri & 1) ti = tm - 1 - ti;
    if (tri & 4) tj = tn - 1 - tj;
    return true;
}

// The XCD-grouped order (group_tile) is shared with the other GEMMs: cfmi_internal.h.

// 16 bytes (8 bf16 at k .. k + 7 of one row), zero outside rows / [.., kend).
__device__ __forceinline__ u4 load16_guarded(const bf16* src, bool row_ok, int64_t k, int64_t kend)
{
    unsigned w[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const unsigned lo = (row_ok && k + 2 * e < kend) ? src[2 * e] : 0u;
        const unsigned hi = (row_ok && k + 2 * e + 1 < kend) ? src[2 * e + 1] : 0u;
        w[e] = lo | (hi << 16);
    }
    return u4{w[0], w[1], w[2], w[3]};
}

// One 32-deep slab of a BT-row operand tile (3 planes x BT rows x 4 vectors
// of 8): thread t holds k chunk skc(t) of rows srow(t) and srow(t) + BT / 2 in
// each plane.  Six named registers rather than an array, so they never live
// in scratch memory.
// Thread t stages row srow(t) = t / 4 and k chunk t % 4 (elements
// skc(t) .. + 7): four consecutive lanes read one row's 64 contiguous bytes.
__device__ __forceinline__ int srow(int t) { return t >> 2; }
__device__ __forceinline__ int skc(int t) { return (t & 3) * 8; }

#define CFMI_B3_EACH(X) X(0, v0) X(1, v1) X(2, v2) X(3, v3) X(4, v4) X(5, v5)
template <int BT>
struct Stage {
    static constexpr int H = Cfg<BT>::HALF;
    u4 v0, v1, v2, v3, v4, v5;   // vector u: plane u / 2, row srow(t) + H (u % 2)

    // Whole vectors.  The tile base (block row, k0, plane) is uniform; each
    // lane adds a 32-bit byte offset for its two rows, which lets every load
    // use a scalar base address.  Rows beyond ``rows`` are clamped to the last
    // row: they feed only output rows / columns beyond M / N, which are never
    // stored, so edge tiles need no predication.
    __device__ __forceinline__ void load_vec(const bf16* __restrict__ P, int64_t ld, int64_t ps, int64_t r0,
                                             int64_t rows, int64_t k0, int tid)
    {
        const char* base = reinterpret_cast<const char*>(P + r0 * ld + k0);
        const int64_t last = rows - 1 - r0;
        const int sr = srow(tid);
        const int64_t ra = sr < last ? sr : last;
        const int64_t rb = sr + H < last ? sr + H : last;
        const unsigned o0 = (unsigned)((ra * ld + skc(tid)) * 2);
        const unsigned o1 = (unsigned)((rb * ld + skc(tid)) * 2);
#define CFMI_B3_LV(u, vv) vv = *reinterpret_cast<const u4*>(base + (u / 2) * ps * 2 + ((u % 2) ? o1 : o0));
        CFMI_B3_EACH(CFMI_B3_LV)
#undef CFMI_B3_LV
    }
    __device__ __forceinline__ void load_guarded(const bf16* __restrict__ P, int64_t ld, int64_t ps, int64_t r0,
                                                 int64_t rows, int64_t k0, int64_t kend, int tid)
    {
        const int64_t gr = r0 + srow(tid), gk = k0 + skc(tid);
        const bf16* base = P + gr * ld + gk;
#define CFMI_B3_LG(u, vv) vv = load16_guarded(base + (u / 2) * ps + (u % 2) * H * ld, gr + (u % 2) * H < rows, gk, kend);
        CFMI_B3_EACH(CFMI_B3_LG)
#undef CFMI_B3_LG
    }
    __device__ __forceinline__ void store(bf16 (*S)[BT][LROW], int tid) const
    {
        const int row = srow(tid), kc = ((tid & 3) ^ swz(row)) * 8;   // swz(row + H) = swz(row)
#define CFMI_B3_ST(u, vv) *reinterpret_cast<u4*>(&S[u / 2][row + (u % 2) * H][kc]) = vv;
        CFMI_B3_EACH(CFMI_B3_ST)
#undef CFMI_B3_ST
    }
};
#undef CFMI_B3_EACH

// 6 MFMAs per 16 x 16 tile and 32-deep slab, smallest terms first; the four
// row tiles of a column are interleaved so consecutive MFMAs are independent.
template <int BT>
__device__ __forceinline__ void slab_mfma(const bf16 (*As)[BT][LROW], const bf16 (*Bs)[BT][LROW],
                                          f4 (&acc)[4][Cfg<BT>::FN], int wm, int wn, int lm, int lk)
{
    const int kk = 8 * (lk ^ swz(lm));   // wm, wn, 16 i, 16 j are multiples of 16
    b8 a[4][3];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int p = 0; p < 3; ++p) a[i][p] = *reinterpret_cast<const b8*>(&As[p][wm + 16 * i + lm][kk]);
#pragma unroll
    for (int j = 0; j < Cfg<BT>::FN; ++j) {
        b8 b[3];
#pragma unroll
        for (int p = 0; p < 3; ++p) b[p] = *reinterpret_cast<const b8*>(&Bs[p][wn + 16 * j + lm][kk]);
#define CFMI_B3_TERM(pa, pb)                                                                            \
    _Pragma("unroll") for (int i = 0; i < 4; ++i)                                                       \
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][pa], b[pb], acc[i][j], 0, 0, 0);
        CFMI_B3_TERM(2, 0) CFMI_B3_TERM(1, 1) CFMI_B3_TERM(0, 2) CFMI_B3_TERM(1, 0) CFMI_B3_TERM(0, 1)
        CFMI_B3_TERM(0, 0)
#undef CFMI_B3_TERM
    }
}

// Staging variants: VEC (whole 16-byte vectors, slabs inside K) and ELEM
// (element-wise: partial last slab or unaligned operands).
enum { VEC = 0, ELEM = 1 };

template <int BT, int LD>
__device__ __forceinline__ void fetch(const B3Args& g, const bf16* A, const bf16* B, int64_t m0, int64_t n0,
                                      int64_t k0, int64_t kend, Stage<BT>& sa, Stage<BT>& sb, int tid)
{
    if constexpr (LD == VEC) {
        sa.load_vec(A, g.lda, g.pA, m0, g.M, k0, tid);
        sb.load_vec(B, g.ldb, g.pB, n0, g.N, k0, tid);
    } else {
        sa.load_guarded(A, g.lda, g.pA, m0, g.M, k0, kend, tid);
        sb.load_guarded(B, g.ldb, g.pB, n0, g.N, k0, kend, tid);
    }
}

// K range [kbeg, kend) in 32-deep slabs: the next slab's global loads are in
// flight (registers) while the current one is multiplied from LDS.
template <int BT, int LD>
__device__ __forceinline__ void k_loop(const B3Args& g, const bf16* A, const bf16* B, int64_t m0, int64_t n0,
                                       int64_t kbeg, int64_t kend, bf16 (*As)[BT][LROW], bf16 (*Bs)[BT][LROW],
                                       f4 (&acc)[4][Cfg<BT>::FN], int tid, int wm, int wn, int lm, int lk)
{
    Stage<BT> sa, sb;
    fetch<BT, LD>(g, A, B, m0, n0, kbeg, kend, sa, sb, tid);
    int64_t k0 = kbeg;
    for (; k0 + TK < kend; k0 += TK) {
        __syncthreads();
        sa.store(As, tid);
        sb.store(Bs, tid);
        __syncthreads();
        fetch<BT, LD>(g, A, B, m0, n0, k0 + TK, kend, sa, sb, tid);
        slab_mfma<BT>(As, Bs, acc, wm, wn, lm, lk);
    }
    __syncthreads();
    sa.store(As, tid);
    sb.store(Bs, tid);
    __syncthreads();
    slab_mfma<BT>(As, Bs, acc, wm, wn, lm, lk);
}

template <int MODE, int BT>
__global__ __launch_bounds__(Cfg<BT>::NT, Cfg<BT>::MINB) void gemm_b3_kernel(B3Args g)
{
    constexpr int FN = Cfg<BT>::FN;
    __shared__ __attribute__((aligned(16))) bf16 As[3][BT][LROW];
    __shared__ __attribute__((aligned(16))) bf16 Bs[3][BT][LROW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int64_t z = blockIdx.z, ti, tj;
    if (g.grouped) {
        if (!group_tile(g, blockIdx.x, z, ti, tj)) return;
    } else if (!tile_order(blockIdx.x, g.tm, g.tmp, g.tn, g.tri, ti, tj)) {
        return;
    }
    const int64_t m0 = ti * BT, n0 = tj * BT;
    if (MODE == 2 && g.col0 + n0 > g.row0 + m0 + BT - 1) return;  // tile above the diagonal
    const bf16* A = g.A + z * g.sA;
    const bf16* B = g.B + z * g.sB;
    int64_t kend = g.K;
    if (g.tri & 1) kend = kend < m0 + BT + g.offA ? kend : m0 + BT + g.offA;
    if (g.tri & 4) kend = kend < n0 + BT + g.offB ? kend : n0 + BT + g.offB;
    const int wm = (wave % Cfg<BT>::WMC) * 64, wn = (wave / Cfg<BT>::WMC) * Cfg<BT>::HALF;
    const int lm = lane & 15, lk = lane >> 4;

    f4 acc[4][FN];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = f4{0, 0, 0, 0};

    const bool vec = g.vecA && g.vecB;
    const int64_t kf = vec ? kend / TK * TK : 0;
    if (kf > 0) k_loop<BT, VEC>(g, A, B, m0, n0, 0, kf, As, Bs, acc, tid, wm, wn, lm, lk);
    if (kf < kend) k_loop<BT, ELEM>(g, A, B, m0, n0, kf, kend, As, Bs, acc, tid, wm, wn, lm, lk);

    // Output lane layout (measured on gfx950): row 4 (lane / 16) + r, column lane % 16.
    float* Cf = (float*)g.C;
    bf16* Cb = (bf16*)g.C;
    const int64_t zoff = z * g.sC;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gm = m0 + wm + 16 * i + 4 * lk + r;
                const int64_t gn = n0 + wn + 16 * j + lm;
                if (gm >= g.M || gn >= g.N) continue;
                float v = g.alpha * acc[i][j][r];
                if (MODE == 0) {
                    float* p = Cf + zoff + gm + gn * g.ldc;
                    *p = g.beta == 0.0f ? v : v + g.beta * *p;
                } else if (MODE == 1) {
                    bf16 h, m, l;
                    split3((double)v, h, m, l);
                    bf16* p = Cb + zoff + gm * g.ldc + gn;
                    p[0] = h;
                    p[g.pC] = m;
                    p[2 * g.pC] = l;
                } else {
                    const int64_t ii = g.row0 + gm, cc = g.col0 + gn;
                    if (cc > ii) continue;
                    if (ii == cc) v *= g.pdiag;
                    float* p = Cf + zoff + (ii * (ii + 1) / 2 + cc - g.pbase);
                    *p = g.beta == 0.0f ? v : v + g.beta * *p;
                }
            }
}

// (rows x cols) column-major FP64 / FP32 source -> three BF16 planes,
// row-major (element (i, j) at dst + i ldd + j), planes ps apart.
template <typename S>
__global__ void split3_kernel(int64_t rows, int64_t cols, const S* src, int64_t lds, bf16* dst, int64_t ldd,
                              int64_t ps)
{
    const int64_t tot = rows * cols;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < tot;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx % rows, j = idx / rows;
        bf16 h, m, l;
        split3((double)src[i + j * lds], h, m, l);
        bf16* p = dst + i * ldd + j;
        p[0] = h;
        p[ps] = m;
        p[2 * ps] = l;
    }
}

// Three planes -> FP64 (sum of the pieces), for tests.
__global__ void join3_kernel(int64_t count, const bf16* src, int64_t ps, double* dst)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (double)bf2f(src[i]) + (double)bf2f(src[i + ps]) + (double)bf2f(src[i + 2 * ps]);
}

inline dim3 grid_of(int64_t tot)
{
    int64_t b = (tot + 255) / 256;
    return dim3((unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536));
}

bool aligned16(const void* p, int64_t ld, int64_t ps)
{
    return (reinterpret_cast<uintptr_t>(p) % 16 == 0) && (ld % 8 == 0) && (ps % 8 == 0);
}

}  // namespace

extern "C" {

int cfmi_split3(cfmi_ctx* c, int dtype, int64_t rows, int64_t cols, const void* src, int64_t lds, void* dst,
                int64_t ldd, int64_t ps)
{
    if (!c) return CFMI_EINVAL;
    if (rows <= 0 || cols <= 0) return CFMI_OK;
    if (ldd < cols || ps < rows * ldd) return cfmi_fail(c, CFMI_EINVAL, "split3: planes overlap");
    if (dtype == CFMI_F64)
        hipLaunchKernelGGL((split3_kernel<double>), grid_of(rows * cols), dim3(256), 0, c->stream, rows, cols,
                           (const double*)src, lds, (bf16*)dst, ldd, ps);
    else if (dtype == CFMI_F32)
        hipLaunchKernelGGL((split3_kernel<float>), grid_of(rows * cols), dim3(256), 0, c->stream, rows, cols,
                           (const float*)src, lds, (bf16*)dst, ldd, ps);
    else
        return cfmi_fail(c, CFMI_EINVAL, "split3: unsupported dtype %d", dtype);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_join3(cfmi_ctx* c, int64_t count, const void* src, int64_t ps, double* dst)
{
    if (!c) return CFMI_EINVAL;
    if (count <= 0) return CFMI_OK;
    hipLaunchKernelGGL(join3_kernel, grid_of(count), dim3(256), 0, c->stream, count, (const bf16*)src, ps, dst);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_gemm_b3(cfmi_ctx* c, int out_mode, int64_t m, int64_t n, int64_t k, double alpha, const void* A,
                 int64_t lda, int64_t sA, int64_t pA, int64_t off_a, const void* B, int64_t ldb, int64_t sB,
                 int64_t pB, int64_t off_b, double beta, void* C, int64_t ldc, int64_t sC, int64_t pC,
                 int64_t batch, int flags, int64_t row0, int64_t col0, int64_t pbase, double pdiag)
{
    if (!c) return CFMI_EINVAL;
    if (m < 0 || n < 0 || k < 0) return cfmi_fail(c, CFMI_EINVAL, "gemm_b3: negative size");
    if (m == 0 || n == 0) return CFMI_OK;
    if (batch < 1 || batch > 65535) return cfmi_fail(c, CFMI_EINVAL, "gemm_b3: batch out of range");
    if (out_mode < 0 || out_mode > 2) return cfmi_fail(c, CFMI_EINVAL, "gemm_b3: bad out_mode %d", out_mode);
    if (lda < k || ldb < k) return cfmi_fail(c, CFMI_EINVAL, "gemm_b3: leading dimension below k");
    if (out_mode == 0 && ldc < m) return cfmi_fail(c, CFMI_EINVAL, "gemm_b3: ldc < m");
    if (out_mode == 1 && (ldc < n || pC < m * ldc)) return cfmi_fail(c, CFMI_EINVAL, "gemm_b3: output planes overlap");
    if (out_mode == 2 && (row0 < 0 || col0 < 0 || pbase < 0 || pbase > row0 * (row0 + 1) / 2))
        return cfmi_fail(c, CFMI_EINVAL, "gemm_b3: packed row band starts after row %lld", (long long)row0);
    B3Args g = {};
    g.M = m; g.N = n; g.K = k;
    g.alpha = (float)alpha; g.beta = (float)beta;
    g.A = (const bf16*)A; g.lda = lda; g.sA = sA; g.pA = pA;
    g.B = (const bf16*)B; g.ldb = ldb; g.sB = sB; g.pB = pB;
    g.C = C; g.ldc = ldc; g.sC = sC; g.pC = pC;
    g.tri = flags & 5; g.offA = off_a; g.offB = off_b;
    g.row0 = row0; g.col0 = col0; g.pbase = pbase; g.pdiag = (float)pdiag;
    // 256-tiles (a CU holds one) when there are enough of them to keep every
    // CU busy and either the product has no B-trim and a full output (any
    // A-trim: its rows are spread over the XCDs, see tile_order), or N and K
    // are large (the row count of a band may be small).  Measured: on
    // 3968-row lower-triangle band outputs with a B-trim the 256-cube
    // staircase computes 31% zeros against 10% for 128-cubes and the 256
    // configuration takes 2x as long; on the triangular 3968^3 products it is
    // 1.3x faster, and on the 8064 / 16128-row products of an nside-32
    // whitening 1.55x.  flags 256 / 512 force the 128 / 256 configuration.
    const int64_t t256 = ((m + 255) / 256) * ((n + 255) / 256) * batch;
    const bool plain = !(flags & 4) && out_mode != 2;
    const bool large = n >= 6144 && k >= 6144;
    const bool big = (flags & 512) ? true : (flags & 256) ? false : ((plain || large) && t256 >= 512);
    const int64_t bt = big ? 256 : 128;
    g.tm = (m + bt - 1) / bt; g.tn = (n + bt - 1) / bt;
    g.vecA = aligned16(A, lda, pA) && (sA % 8 == 0);
    g.vecB = aligned16(B, ldb, pB) && (sB % 8 == 0);
    g.tmp = ((flags & 1) || out_mode == 2) && g.tm % 2 == 0 ? g.tm + 1 : g.tm;
    // flag 1024: the column-by-column order of tile_order, batch over blockIdx.z
    g.grouped = !(flags & 1024);
    // group shape 2 x 8 tiles (both tile sizes); a value v > 0 in bits 12-15
    // / 16-19 gives 2^(v - 1) group rows / columns instead (1 .. 64).
    // Measured on the nside-16 TQU whitening: 157.7 ms at 2 x 8, within 1%
    // of it for 1-2 rows x 8-64 columns, 167.8 ms at 4 x 8 / 4 x 32, 188 ms
    // at 8 x 8 128-tiles with 4 x 8 256-tiles, 200 ms at 16 x 4.
    const int lgm = (flags >> 12) & 15, lgn = (flags >> 16) & 15;
    g.gm = lgm ? (1 << std::min(lgm - 1, 6)) : 2;
    g.gn = lgn ? (1 << std::min(lgn - 1, 6)) : 8;
    g.gtm = (g.tm + g.gm - 1) / g.gm;
    g.gtn = (g.tn + g.gn - 1) / g.gn;
    g.gtmp = g.gtm | 1;
    g.ngroups = g.gtmp * g.gtn * batch;
    const int64_t nwg = g.grouped ? (g.ngroups + 7) / 8 * 8 * g.gm * g.gn : g.tmp * g.tn;
    if (nwg > 0x7fffffffLL) return cfmi_fail(c, CFMI_EINVAL, "gemm_b3: too many tiles");
    dim3 grid((unsigned)nwg, 1, g.grouped ? 1u : (unsigned)batch);
#define CFMI_B3_GO(MODE_, BT_) \
    hipLaunchKernelGGL((gemm_b3_kernel<MODE_, BT_>), grid, dim3(Cfg<BT_>::NT), 0, c->stream, g)
    if (big) {
        if (out_mode == 0) CFMI_B3_GO(0, 256);
        else if (out_mode == 1) CFMI_B3_GO(1, 256);
        else CFMI_B3_GO(2, 256);
    } else {
        if (out_mode == 0) CFMI_B3_GO(0, 128);
        else if (out_mode == 1) CFMI_B3_GO(1, 128);
        else CFMI_B3_GO(2, 128);
    }
#undef CFMI_B3_GO
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

}  // extern "C"
