// Ensemble MLP surrogate of the exact log-likelihood (no reference
// counterpart).  E independent networks 3 -> 256 -> 256 -> 256 -> 1 with SiLU
// activations; weights used in BF16 (master copy and Adam moments in FP32),
// every product accumulated in FP32.
//
//   * the 256 x 256 layers run on BF16 MFMA (v_mfma_f32_16x16x16_bf16):
//     64 x 64 output tiles per workgroup, 2 x 2 waves of 32 x 32, operand
//     tiles staged in LDS in (row, k) order so that any storage transpose
//     of A or B is handled by the staging loads;
//   * forward epilogue: bias + SiLU, writing the pre-activation (FP32, for
//     the backward pass) and the activation (BF16, the next MFMA operand);
//   * backward-data epilogue: multiply by SiLU'(z) of the layer below;
//   * weight gradients dW = delta^T h reduce over the batch in fixed-size
//     chunks whose partial sums are added in a fixed order (deterministic);
//   * the 3-wide input layer and the 1-wide output layer are VALU kernels.
//
// Parameter layout of one member (FP32 master and BF16 copies alike):
//   W1 [256][3] @0, b1 [256] @768, W2 [256][256] @1024, b2 @66560,
//   W3 [256][256] @66816, b3 @132352, W4 [1][256] @132608, b4 @132864;
//   132865 values.  W[out][in] row-major.
#include "cfmi_internal.h"

#include <mutex>

#include <cmath>
#include <cstring>

namespace {

constexpr int H = 256;
constexpr int NIN = 3;
constexpr int64_t OFF_W1 = 0, OFF_B1 = 768, OFF_W2 = 1024, OFF_B2 = 66560, OFF_W3 = 66816, OFF_B3 = 132352,
                  OFF_W4 = 132608, OFF_B4 = 132864, NPAR = 132865;

typedef unsigned short bf16;
typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16 v)
{
    unsigned u = (unsigned)v << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
// Round to nearest even (NaN kept quiet).
__device__ __forceinline__ bf16 f2bf(float f)
{
    unsigned u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z)
{
    float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

// ---- BF16 MFMA GEMM: C[M x N] = op(A) op(B), A, B row-major BF16 ----------
// TA = 0: A stored M x K; TA = 1: stored K x M.  TB = 0: B stored K x N;
// TB = 1: stored N x K.  Batched over blockIdx.z (ensemble member), split-K
// over blockIdx.y (EPI_PART only).
enum { EPI_PART = 0, EPI_FWD = 1, EPI_BWD = 2 };

struct MG {
    int M, N, K, kchunk;
    const bf16* A;
    int64_t sA;
    const bf16* B;
    int64_t sB;
    float* C;          // EPI_PART: partials [member][split][M][N]; EPI_FWD: z out [M][N]
    int64_t sC;
    bf16* Hout;        // EPI_FWD: SiLU(z) ; EPI_BWD: delta out
    int64_t sH;
    const float* bias;  // EPI_FWD
    int64_t sbias;
    const float* Z;     // EPI_BWD: pre-activation of the layer below
    int64_t sZ;
    int nsplit;
};

constexpr int MT = 64, KT = 32, KP = KT + 8;

template <int TA, int TB, int EPI>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(MG g)
{
    __shared__ bf16 As[MT][KP];
    __shared__ bf16 Bs[MT][KP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tn = (g.N + MT - 1) / MT;
    const int m0 = (blockIdx.x / tn) * MT, n0 = (blockIdx.x % tn) * MT;
    const int e = blockIdx.z, sp = blockIdx.y;
    const bf16* A = g.A + e * g.sA;
    const bf16* B = g.B + e * g.sB;
    const int kbeg = sp * g.kchunk;
    const int kend = min(g.K, kbeg + g.kchunk);
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    f4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f4{0, 0, 0, 0};
    for (int k0 = kbeg; k0 < kend; k0 += KT) {
        __syncthreads();
        // A tile -> As[m][k]
        if (TA == 0) {
            const int r = tid >> 2, kc = (tid & 3) * 8;
            const int gm = m0 + r;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int gk = k0 + kc + u;
                As[r][kc + u] = load_or_zero(A, (int64_t)gm * g.K + gk, gm < g.M && gk < kend);
            }
        } else {
            const int kr = tid >> 3, mc = (tid & 7) * 8;
            const int gk = k0 + kr;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int gm = m0 + mc + u;
                As[mc + u][kr] = load_or_zero(A, (int64_t)gk * g.M + gm, gm < g.M && gk < kend);
            }
        }
        // B tile -> Bs[n][k]
        if (TB == 1) {
            const int r = tid >> 2, kc = (tid & 3) * 8;
            const int gn = n0 + r;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int gk = k0 + kc + u;
                Bs[r][kc + u] = load_or_zero(B, (int64_t)gn * g.K + gk, gn < g.N && gk < kend);
            }
        } else {
            const int kr = tid >> 3, nc = (tid & 7) * 8;
            const int gk = k0 + kr;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int gn = n0 + nc + u;
                Bs[nc + u][kr] = load_or_zero(B, (int64_t)gk * g.N + gn, gn < g.N && gk < kend);
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KT; kk += 16) {
            s4 av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    av[a][j] = (short)As[wm + 16 * a + (lane & 15)][kk + 4 * (lane >> 4) + j];
                    bv[a][j] = (short)Bs[wn + 16 * a + (lane & 15)][kk + 4 * (lane >> 4) + j];
                }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
    }
    // Output lane layout (measured on gfx950): row 4 (lane / 16) + r, column lane % 16.
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + 16 * a + 4 * (lane >> 4) + r;
                const int gn = n0 + wn + 16 * b + (lane & 15);
                if (gm >= g.M || gn >= g.N) continue;
                const int64_t o = (int64_t)gm * g.N + gn;
                const float v = acc[a][b][r];
                if (EPI == EPI_PART) {
                    g.C[e * g.sC + (int64_t)sp * g.M * g.N + o] = v;
                } else if (EPI == EPI_FWD) {
                    const float zz = v + g.bias[e * g.sbias + gn];
                    g.C[e * g.sC + o] = zz;
                    g.Hout[e * g.sH + o] = f2bf(silu(zz));
                } else {
                    g.Hout[e * g.sH + o] = f2bf(v * dsilu(g.Z[e * g.sZ + o]));
                }
            }
}

// ---- small kernels ---------------------------------------------------------

// z1 = x W1^T + b1, h1 = SiLU(z1); x row idx[e][b] of X (or b when idx NULL).
__global__ void l1_fwd_kernel(int B, const float* X, const int* idx, const bf16* wb, const float* w32, float* z1,
                              bf16* h1, int64_t sAct)
{
    const int e = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)B * H) return;
    const int b = (int)(t / H), j = (int)(t % H);
    const int row = idx ? idx[(int64_t)e * B + b] : b;
    const bf16* W1 = wb + e * NPAR + OFF_W1;
    float z = w32[e * NPAR + OFF_B1 + j];
#pragma unroll
    for (int k = 0; k < NIN; ++k) z += X[(int64_t)row * NIN + k] * bf2f(W1[j * NIN + k]);
    z1[e * sAct + t] = z;
    h1[e * sAct + t] = f2bf(silu(z));
}

// y[e][b] = h3[b] . W4 + b4 (one wave per sample, FP32 accumulation).
__global__ void l4_fwd_kernel(int B, const bf16* h3, int64_t sAct, const bf16* wb, const float* w32, float* y)
{
    const int e = blockIdx.y;
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (b >= B) return;
    const bf16* W4 = wb + e * NPAR + OFF_W4;
    const bf16* h = h3 + e * sAct + (int64_t)b * H;
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < H / 64; ++q) s += bf2f(h[lane + 64 * q]) * bf2f(W4[lane + 64 * q]);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    if (lane == 0) y[(int64_t)e * B + b] = s + w32[e * NPAR + OFF_B4];
}

// dy = 2 (y - t) / B; loss[e] partials via one block per member.
__global__ void loss_kernel(int B, const float* y, const float* T, const int* idx, float* dy, float* loss)
{
    const int e = blockIdx.x;
    __shared__ float red[256];
    float s = 0.0f;
    for (int b = threadIdx.x; b < B; b += 256) {
        const int row = idx ? idx[(int64_t)e * B + b] : b;
        const float r = y[(int64_t)e * B + b] - T[row];
        s += r * r;
        dy[(int64_t)e * B + b] = 2.0f * r / (float)B;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[e] = red[0] / (float)B;
}

// delta3 = dy W4 SiLU'(z3) (BF16) for batch rows of chunk blockIdx.x; the
// chunk's partial sums of gW4[j] = sum_b dy_b h3[b][j] and gb4 = sum_b dy_b
// go to part[e][chunk][0..256] (reduced in a fixed order afterwards).
// Chunks are short (RCH rows, B / RCH blocks per member) so that the batch
// spreads over the CUs; each chunk's dy (and, in colsum_kernel, its input
// rows) is staged in LDS first, so the row loop issues only independent,
// coalesced loads.
constexpr int RCH = 8;  // batch rows per reduction chunk

__global__ void l4_bwd_kernel(int B, const float* dy, const bf16* h3, const float* z3, int64_t sAct, const bf16* wb,
                              bf16* d3, float* part, int nchunk)
{
    __shared__ float g_s[RCH];
    const int e = blockIdx.y, c = blockIdx.x;
    const int j = threadIdx.x;  // column (256 threads)
    const int b0 = c * RCH, nb = min(RCH, B - b0);
    if (j < nb) g_s[j] = dy[(int64_t)e * B + b0 + j];
    __syncthreads();
    const float w4 = bf2f(wb[e * NPAR + OFF_W4 + j]);
    float gw = 0.0f, gb = 0.0f;
#pragma unroll
    for (int r = 0; r < RCH; ++r) {
        if (r >= nb) break;
        const float g = g_s[r];
        const int64_t o = e * sAct + (int64_t)(b0 + r) * H + j;
        gw += g * bf2f(h3[o]);
        gb += g;
        d3[o] = f2bf(g * w4 * dsilu(z3[o]));
    }
    float* pp = part + ((int64_t)e * nchunk + c) * (H + 1);
    pp[j] = gw;
    if (j == 0) pp[H] = gb;
}

// Chunk partials of db[j] = sum_b delta[b][j] and, for the first layer,
// gW1[j][k] = sum_b delta[b][j] x[b][k]: part[e][chunk][j * 4 + {0: db, 1..3: gW1}].
__global__ void colsum_kernel(int B, const bf16* d, int64_t sAct, const float* X, const int* idx, float* part,
                              int nchunk)
{
    __shared__ float x_s[RCH][NIN];
    const int e = blockIdx.y, c = blockIdx.x;
    const int j = threadIdx.x;
    const int b0 = c * RCH, nb = min(RCH, B - b0);
    if (X && j < nb * NIN) {
        const int r = j / NIN, k = j % NIN;
        const int row = idx ? idx[(int64_t)e * B + b0 + r] : b0 + r;
        x_s[r][k] = X[(int64_t)row * NIN + k];
    }
    __syncthreads();
    float sb = 0.0f, sw[NIN] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int r = 0; r < RCH; ++r) {
        if (r >= nb) break;
        const float v = bf2f(d[e * sAct + (int64_t)(b0 + r) * H + j]);
        sb += v;
        if (X) {
#pragma unroll
            for (int k = 0; k < NIN; ++k) sw[k] += v * x_s[r][k];
        }
    }
    float* pp = part + ((int64_t)e * nchunk + c) * (4 * H);
    pp[4 * j] = sb;
#pragma unroll
    for (int k = 0; k < NIN; ++k) pp[4 * j + 1 + k] = sw[k];
}

// Fixed-order sums of the chunk partials into the gradient.
__global__ void l4_reduce_kernel(int nchunk, const float* part, float* grad)
{
    const int e = blockIdx.y, j = threadIdx.x;
    float gw = 0.0f, gb = 0.0f;
    for (int c = 0; c < nchunk; ++c) {
        const float* pp = part + ((int64_t)e * nchunk + c) * (H + 1);
        gw += pp[j];
        if (j == 0) gb += pp[H];
    }
    grad[e * NPAR + OFF_W4 + j] = gw;
    if (j == 0) grad[e * NPAR + OFF_B4] = gb;
}

__global__ void colsum_reduce_kernel(int nchunk, const float* part, float* grad, int64_t off_b, int with_w1)
{
    const int e = blockIdx.y, j = threadIdx.x;
    float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int c = 0; c < nchunk; ++c) {
        const float* pp = part + ((int64_t)e * nchunk + c) * (4 * H) + 4 * j;
#pragma unroll
        for (int k = 0; k < 4; ++k) s[k] += pp[k];
    }
    grad[e * NPAR + off_b + j] = s[0];
    if (with_w1)
#pragma unroll
        for (int k = 0; k < NIN; ++k) grad[e * NPAR + OFF_W1 + j * NIN + k] = s[1 + k];
}

// grad[e][off + i] = sum over splits of part[e][s][i] (fixed order).
__global__ void split_sum_kernel(int nsplit, int64_t count, const float* part, int64_t spart, float* grad,
                                 int64_t off)
{
    const int e = blockIdx.y;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int q = 0; q < nsplit; ++q) s += part[e * spart + (int64_t)q * count + i];
        grad[e * NPAR + off + i] = s;
    }
}

__global__ void adam_kernel(int64_t total, float* w, const float* g, float* m, float* v, float lr, float b1, float b2,
                            float eps, float c1, float c2, bf16* wb)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        wb[i] = f2bf(wi);
    }
}

__global__ void to_bf16_kernel(int64_t total, const float* w, bf16* wb)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        wb[i] = f2bf(w[i]);
}

// mean and (population) standard deviation over the E members.
__global__ void stats_kernel(int E, int B, const float* y, float* mean, float* sd)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    float s = 0.0f;
    for (int e = 0; e < E; ++e) s += y[(int64_t)e * B + b];
    const float mu = s / (float)E;
    float q = 0.0f;
    for (int e = 0; e < E; ++e) {
        const float d = y[(int64_t)e * B + b] - mu;
        q += d * d;
    }
    mean[b] = mu;
    sd[b] = sqrtf(q / (float)E);
}

inline dim3 g1(int64_t tot) { return dim3((unsigned)((tot + 255) / 256)); }
inline dim3 gcap(int64_t tot)
{
    int64_t b = (tot + 255) / 256;
    return dim3((unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536));
}

template <int TA, int TB, int EPI>
int launch_gemm(cfmi_ctx* c, MG& g, int E)
{
    const int tiles = ((g.M + MT - 1) / MT) * ((g.N + MT - 1) / MT);
    dim3 grid((unsigned)tiles, (unsigned)g.nsplit, (unsigned)E);
    hipLaunchKernelGGL((mlp_gemm_kernel<TA, TB, EPI>), grid, dim3(256), 0, c->stream, g);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

// Activation workspace of one member: z1, z2, z3 (FP32) and h1, h2, h3, d1,
// d2, d3 (BF16), each B x 256.
struct Acts {
    float* z[3];
    bf16* h[3];
    bf16* d[3];
    int64_t s;  // per-member element stride of each array
};

Acts acts_of(void* ws, int E, int B)
{
    Acts a;
    const int64_t s = (int64_t)B * H;
    a.s = s;
    float* zf = (float*)ws;
    for (int l = 0; l < 3; ++l) a.z[l] = zf + (int64_t)l * E * s;
    bf16* hb = (bf16*)(zf + 3 * E * s);
    for (int l = 0; l < 3; ++l) a.h[l] = hb + (int64_t)l * E * s;
    for (int l = 0; l < 3; ++l) a.d[l] = hb + (int64_t)(3 + l) * E * s;
    return a;
}

// Scratch of cfmi_mlp_backward_noalloc (32 MiB; each GPU the library runs
// on holds its own copy until the process ends).
constexpr int64_t MLP_SCRATCH_FLOATS = (int64_t)8 << 20;
static __device__ __align__(256) float scratch[MLP_SCRATCH_FLOATS];

// Device address of `scratch` on the context's device: obtained once per
// device, with that device current, by hipGetSymbolAddress (the array's name
// in host code is not a device address).
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_ptr[MAX_DEVICES] = {};

int scratch_address(cfmi_ctx* c, float** out)
{
    if (c->device < 0 || c->device >= MAX_DEVICES) return cfmi_fail(c, CFMI_EINVAL, "device %d out of range", c->device);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[c->device]) {
        int cur = -1;
        CFMI_HIP(c, hipGetDevice(&cur));
        if (cur != c->device) CFMI_HIP(c, hipSetDevice(c->device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != c->device) CFMI_HIP(c, hipSetDevice(cur));
        if (e != hipSuccess || !p)
            return cfmi_fail(c, CFMI_EHIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        scratch_ptr[c->device] = p;
    }
    *out = static_cast<float*>(scratch_ptr[c->device]);
    return CFMI_OK;
}

static int mlp_backward_impl(cfmi_ctx* c, int E, int B, const float* X, const int* idx, const void* wb16,
                             const float* dy, void* ws, float* grad, float* part, int chunk)
{
    if (!c) return CFMI_EINVAL;
    if (chunk < 32 || chunk % 32) return cfmi_fail(c, CFMI_EINVAL, "mlp: gradient chunk must be a multiple of 32");
    const bf16* wb = (const bf16*)wb16;
    Acts a = acts_of(ws, E, B);
    const int nsplit = (B + chunk - 1) / chunk;
    // Reduction partials reuse `part` (it holds at least E * nsplit * 256 * 256 floats).
    const int nchunk = (B + RCH - 1) / RCH;
    if ((int64_t)nchunk * 4 * H > (int64_t)nsplit * H * H)
        return cfmi_fail(c, CFMI_EINVAL, "mlp: gradient chunk too large for the batch reductions");
    // output layer
    hipLaunchKernelGGL(l4_bwd_kernel, dim3((unsigned)nchunk, (unsigned)E), dim3(256), 0, c->stream, B, dy, a.h[2],
                       a.z[2], a.s, wb, a.d[2], part, nchunk);
    CFMI_LAUNCH_CHECK(c);
    hipLaunchKernelGGL(l4_reduce_kernel, dim3(1, (unsigned)E), dim3(256), 0, c->stream, nchunk, part, grad);
    CFMI_LAUNCH_CHECK(c);
    for (int l = 2; l >= 1; --l) {
        const int64_t offw = l == 2 ? OFF_W3 : OFF_W2, offb = l == 2 ? OFF_B3 : OFF_B2;
        // dW_l = delta_l^T h_{l-1}  (split over the batch, partials reduced in order)
        MG g = {};
        g.M = H; g.N = H; g.K = B; g.kchunk = chunk; g.nsplit = nsplit;
        g.A = a.d[l]; g.sA = a.s;          // stored B x 256 = K x M  (TA = 1)
        g.B = a.h[l - 1]; g.sB = a.s;      // stored B x 256 = K x N  (TB = 0)
        g.C = part; g.sC = (int64_t)nsplit * H * H;
        int rc = launch_gemm<1, 0, EPI_PART>(c, g, E);
        if (rc) return rc;
        hipLaunchKernelGGL(split_sum_kernel, dim3(gcap((int64_t)H * H).x, (unsigned)E), dim3(256), 0, c->stream,
                           nsplit, (int64_t)H * H, part, (int64_t)nsplit * H * H, grad, offw);
        CFMI_LAUNCH_CHECK(c);
        hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)nchunk, (unsigned)E), dim3(256), 0, c->stream, B, a.d[l],
                           a.s, (const float*)nullptr, (const int*)nullptr, part, nchunk);
        CFMI_LAUNCH_CHECK(c);
        hipLaunchKernelGGL(colsum_reduce_kernel, dim3(1, (unsigned)E), dim3(256), 0, c->stream, nchunk, part, grad, offb,
                           0);
        CFMI_LAUNCH_CHECK(c);
        // delta_{l-1} = (delta_l W_l) SiLU'(z_{l-1})
        MG d = {};
        d.M = B; d.N = H; d.K = H; d.kchunk = H; d.nsplit = 1;
        d.A = a.d[l]; d.sA = a.s;          // B x 256 = M x K (TA = 0)
        d.B = wb + offw; d.sB = NPAR;      // W [out][in] = K x N (TB = 0)
        d.Hout = a.d[l - 1]; d.sH = a.s;
        d.Z = a.z[l - 1]; d.sZ = a.s;
        rc = launch_gemm<0, 0, EPI_BWD>(c, d, E);
        if (rc) return rc;
    }
    // first layer: gb1 and gW1 = delta_1^T x
    hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)nchunk, (unsigned)E), dim3(256), 0, c->stream, B, a.d[0], a.s, X,
                       idx, part, nchunk);
    CFMI_LAUNCH_CHECK(c);
    hipLaunchKernelGGL(colsum_reduce_kernel, dim3(1, (unsigned)E), dim3(256), 0, c->stream, nchunk, part, grad, OFF_B1,
                       1);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

}  // namespace

extern "C" {

int64_t cfmi_mlp_param_count(void) { return NPAR; }

int64_t cfmi_mlp_workspace_bytes(int E, int B)
{
    const int64_t s = (int64_t)B * H * E;
    return 3 * s * 4 + 6 * s * 2;
}

int cfmi_mlp_forward(cfmi_ctx* c, int E, int B, const float* X, const int* idx, const float* w32, const void* wb16,
                     float* y, void* ws)
{
    if (!c) return CFMI_EINVAL;
    if (E < 1 || E > 64 || B < 1) return cfmi_fail(c, CFMI_EINVAL, "mlp: bad ensemble %d / batch %d", E, B);
    const bf16* wb = (const bf16*)wb16;
    Acts a = acts_of(ws, E, B);
    hipLaunchKernelGGL(l1_fwd_kernel, dim3(g1((int64_t)B * H).x, (unsigned)E), dim3(256), 0, c->stream, B, X, idx,
                       wb, w32, a.z[0], a.h[0], a.s);
    CFMI_LAUNCH_CHECK(c);
    for (int l = 1; l < 3; ++l) {
        MG g = {};
        g.M = B; g.N = H; g.K = H; g.kchunk = H; g.nsplit = 1;
        g.A = a.h[l - 1]; g.sA = a.s;
        g.B = wb + (l == 1 ? OFF_W2 : OFF_W3); g.sB = NPAR;
        g.C = a.z[l]; g.sC = a.s;
        g.Hout = a.h[l]; g.sH = a.s;
        g.bias = w32 + (l == 1 ? OFF_B2 : OFF_B3); g.sbias = NPAR;
        int rc = launch_gemm<0, 1, EPI_FWD>(c, g, E);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(l4_fwd_kernel, dim3((unsigned)((B + 3) / 4), (unsigned)E), dim3(256), 0, c->stream, B, a.h[2],
                       a.s, wb, w32, y);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_mlp_loss(cfmi_ctx* c, int E, int B, const float* y, const float* T, const int* idx, float* dy, float* loss)
{
    if (!c) return CFMI_EINVAL;
    hipLaunchKernelGGL(loss_kernel, dim3((unsigned)E), dim3(256), 0, c->stream, B, y, T, idx, dy, loss);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_mlp_backward(cfmi_ctx* c, int E, int B, const float* X, const int* idx, const void* wb16, const float* dy,
                      void* ws, float* grad, float* part, int chunk)
{
    return mlp_backward_impl(c, E, B, X, idx, wb16, dy, ws, grad, part, chunk);
}

// The backward without a workspace argument: its reduction partials live in
// the file-scope array `scratch` (MLP_SCRATCH_FLOATS floats, E * ceil(B /
// chunk) * 256 * 256 of them used).  No allocation happens here; host code
// reaches the array only through the per-device address that
// scratch_address() obtains once with hipGetSymbolAddress.  Every partial is
// written before it is read, so the array needs no clearing; calls that
// overlap on one GPU share it and race.
int cfmi_mlp_backward_noalloc(cfmi_ctx* c, int E, int B, const float* X, const int* idx, const void* wb16,
                              const float* dy, void* ws, float* grad, int chunk)
{
    if (!c) return CFMI_EINVAL;
    if (E <= 0 || B <= 0) return cfmi_fail(c, CFMI_EINVAL, "mlp_backward_noalloc: E and B must be positive");
    if (chunk < 32 || chunk % 32) return cfmi_fail(c, CFMI_EINVAL, "mlp: gradient chunk must be a multiple of 32");
    const int64_t need = (int64_t)E * ((B + chunk - 1) / chunk) * H * H;
    if (need > MLP_SCRATCH_FLOATS)
        return cfmi_fail(c, CFMI_EINVAL,
                         "mlp_backward_noalloc: E * ceil(B / chunk) * 256 * 256 = %lld floats exceed the %lld of the "
                         "static scratch (E * ceil(B / chunk) <= %lld); use cfmi_mlp_backward with a workspace",
                         (long long)need, (long long)MLP_SCRATCH_FLOATS, (long long)(MLP_SCRATCH_FLOATS / (H * H)));
    float* part = nullptr;
    int rc = scratch_address(c, &part);
    if (rc) return rc;
    return mlp_backward_impl(c, E, B, X, idx, wb16, dy, ws, grad, part, chunk);
}

int cfmi_mlp_adam(cfmi_ctx* c, int64_t total, float* w, const float* g, float* m, float* v, double lr, double beta1,
                  double beta2, double eps, int step, void* wb16)
{
    if (!c) return CFMI_EINVAL;
    if (step < 1) return cfmi_fail(c, CFMI_EINVAL, "mlp: Adam step counts from 1");
    const float c1 = (float)(1.0 - std::pow(beta1, step)), c2 = (float)(1.0 - std::pow(beta2, step));
    hipLaunchKernelGGL(adam_kernel, gcap(total), dim3(256), 0, c->stream, total, w, g, m, v, (float)lr, (float)beta1,
                       (float)beta2, (float)eps, c1, c2, (bf16*)wb16);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_mlp_to_bf16(cfmi_ctx* c, int64_t total, const float* w, void* wb16)
{
    if (!c) return CFMI_EINVAL;
    hipLaunchKernelGGL(to_bf16_kernel, gcap(total), dim3(256), 0, c->stream, total, w, (bf16*)wb16);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_mlp_ensemble_stats(cfmi_ctx* c, int E, int B, const float* y, float* mean, float* sd)
{
    if (!c) return CFMI_EINVAL;
    hipLaunchKernelGGL(stats_kernel, g1(B), dim3(256), 0, c->stream, E, B, y, mean, sd);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

}  // extern "C"
