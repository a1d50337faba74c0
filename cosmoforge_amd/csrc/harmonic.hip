// Harmonic-mode form of the Fisher and QML contractions.
//
// By the addition theorems the pixel-space derivative of every spectrum at
// one multipole is a sum of 2l+1 rank-one terms over real harmonic modes:
// dS/dC_l^XY = sum_m (x_lm y_lm^T [+ y_lm x_lm^T when X != Y]), with x, y the
// T (spin-0) or E / B (spin-2, over the Q and U rows) mode vectors.  The
// Fisher trace then only needs M = V^T C^-1 V:
//   Tr[C^-1 u v^T C^-1 u' v'^T] = M[v, u'] M[u, v']   (M symmetric),
// and the QML quadratic forms only u^T y and u^T G v.  This file builds V
// (cfmi_harmonic_modes: one thread per pixel and m, the Wigner d^l_{m,+-2}
// (or d^l_{m,0}) upward recurrence in l held in registers) and reduces the
// products over the modes of each bin pair (cfmi_harmonic_fisher, one block
// per F entry, reading M column-contiguously) or each bin
// (cfmi_harmonic_pairsum).  The modes are l-major (l in [lmin, lmax]; m = 0,
// then the cos and sin modes of m = 1..l), so a bin is a contiguous mode
// range.  Mode conventions follow src/cosmoforge.cosmocore/cosmocore/basis/
// harmonic_basis.py:19-117 and basics/wigner.py:12-160, with the E and B
// modes negated so that the T x pol. blocks carry the pixel-space sign of
// signal_kernels.py (Q + iU = -sum (a_E + i a_B) 2Y).
#include "cfmi_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr int HT = 128;   // threads per block, modes kernel
constexpr int RT = 256;   // threads per block, reductions

// d^p_{p,q}(theta) = sqrt((2p)! / ((p+q)! (p-q)!)) c^{p+q} s^{p-q}, p >= |q|,
// c = cos(theta/2), s = sin(theta/2).
__device__ double dstart(int p, int q, double c, double s)
{
    const int a = p + q, b = p - q;
    if ((a > 0 && c == 0.0) || (b > 0 && s == 0.0)) return 0.0;
    double lg = 0.5 * (lgamma(2.0 * p + 1.0) - lgamma(a + 1.0) - lgamma(b + 1.0));
    if (a > 0) lg += a * log(c);
    if (b > 0) lg += b * log(s);
    return exp(lg);
}

// Upward recurrence of d^l_{m,s} in l (fixed m, s), from l0 = max(|m|, |s|).
struct DRec {
    double prev, cur;
    int l, m, s;

    __device__ void init(int m_, int s_, double c, double sn)
    {
        m = m_;
        s = s_;
        int p = m, q = s;
        double sg = 1.0;
        if (abs(p) < abs(q)) {          // d_{m,s} = (-1)^{m-s} d_{s,m}
            p = s;
            q = m;
            if ((m - s) & 1) sg = -sg;
        }
        if (p < 0) {                    // d_{p,q} = (-1)^{p-q} d_{-p,-q}
            if ((p - q) & 1) sg = -sg;
            p = -p;
            q = -q;
        }
        l = p;
        prev = 0.0;
        cur = sg * dstart(p, q, c, sn);
    }

    __device__ void step(double z)
    {
        double next;
        if (l == 0) {
            next = z * cur;             // d^1_{00} = cos theta
        } else {
            const double L = l, L1 = l + 1.0, mm = (double)m * m, ss = (double)s * s;
            const double fn = sqrt((L1 * L1 - mm) * (L1 * L1 - ss));
            const double fl = sqrt((L * L - mm) * (L * L - ss));
            next = ((2.0 * L + 1.0) * (L * L1 * z - (double)m * s) * cur - L1 * fl * prev) / (L * fn);
        }
        prev = cur;
        cur = next;
        ++l;
    }
};

__global__ __launch_bounds__(HT) void modes_kernel(int spin, int64_t npix, const double* __restrict__ vec, int lmin,
                                                   int lmax, double* __restrict__ V, int64_t ldv, int64_t ra,
                                                   int64_t rb, int64_t ca, int64_t cb)
{
    const int64_t i = (int64_t)blockIdx.x * HT + threadIdx.x;
    const int m = blockIdx.y;
    if (i >= npix) return;
    const double x = vec[3 * i], y = vec[3 * i + 1];
    const double z = fmin(1.0, fmax(-1.0, vec[3 * i + 2]));
    const double c = sqrt(0.5 * (1.0 + z)), sn = sqrt(0.5 * (1.0 - z));
    const double phi = atan2(y, x);
    const double cm = cos(m * phi), sm = sin(m * phi);
    const double inv4pi = 0.25 / M_PI, r2 = M_SQRT2;
    const int64_t lmin2 = (int64_t)lmin * lmin;
    if (spin == 0) {
        DRec r;
        r.init(m, 0, c, sn);
        for (int l = r.l; l <= lmax; ++l) {
            if (l >= lmin) {
                const double sc = sqrt((2.0 * l + 1.0) * inv4pi);
                const int64_t base = ca + (int64_t)l * l - lmin2;
                if (m == 0) {
                    V[ra + i + base * ldv] = sc * r.cur;
                } else {
                    const double b = r2 * sc * r.cur;
                    V[ra + i + (base + 2 * m - 1) * ldv] = b * cm;
                    V[ra + i + (base + 2 * m) * ldv] = b * sm;
                }
            }
            if (l < lmax) r.step(z);
        }
        return;
    }
    DRec rp, rm;
    rp.init(m, 2, c, sn);
    rm.init(m, -2, c, sn);
    for (int l = rp.l; l <= lmax; ++l) {
        if (l >= lmin) {
            const double sc = sqrt((2.0 * l + 1.0) * inv4pi);
            const double dp = rm.cur + rp.cur, dm = rm.cur - rp.cur;
            const int64_t base = (int64_t)l * l - lmin2;
            if (m == 0) {
                const double h = 0.5 * sc;
                const int64_t ce = ca + base, cbb = cb + base;
                V[ra + i + ce * ldv] = -h * dp;
                V[rb + i + ce * ldv] = 0.0;
                V[ra + i + cbb * ldv] = 0.0;
                V[rb + i + cbb * ldv] = -h * dp;
            } else {
                const double h = 0.5 * r2 * sc;
                const int64_t ec = ca + base + 2 * m - 1, es = ec + 1;
                const int64_t bc = cb + base + 2 * m - 1, bs = bc + 1;
                V[ra + i + ec * ldv] = -h * dp * cm;
                V[rb + i + ec * ldv] = -h * dm * sm;
                V[ra + i + bc * ldv] = h * dm * sm;
                V[rb + i + bc * ldv] = -h * dp * cm;
                V[ra + i + es * ldv] = -h * dp * sm;
                V[rb + i + es * ldv] = h * dm * cm;
                V[ra + i + bs * ldv] = -h * dm * cm;
                V[rb + i + bs * ldv] = -h * dp * sm;
            }
        }
        if (l < lmax) {
            rp.step(z);
            rm.step(z);
        }
    }
}

// Bin / spectrum tables (device pointers) of the two reductions.
struct HarmTables {
    const int* lo;        // [nbins] first multipole of each bin
    const int* hi;        // [nbins] last multipole
    const int* pairs;     // [nspec][4]: groups (u0, v0, u1, v1); u1 < 0 for an auto spectrum
    const int64_t* goff;  // [ngroups] first column of each mode group
    const double* w;      // [nspec][lw] l-indexed weights
    int lw, nbins, lmin;
};

__device__ __forceinline__ int ell_of(int64_t k, int lmin)
{
    const int64_t t = k + (int64_t)lmin * lmin;
    int64_t l = (int64_t)sqrt((double)t);
    while ((l + 1) * (l + 1) <= t) ++l;
    while (l * l > t) --l;
    return (int)l;
}

__device__ double block_sum(double v, double* red)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < RT / 64; ++q) s += red[q];
    return s;
}

// Column dot products of every mode for every pair of mode groups:
// dots[(k G + a) G + b] = sum_i wrow_i X[i, goff[a] + k] Y[i, goff[b] + k]
// (wrow = 1 when null), a, b < G <= 3.  One block per mode k reads each of
// its G (or 2 G) columns once, however many group pairs use them; the sums
// are reduced in a fixed order.
__global__ __launch_bounds__(RT) void harm_coldots_kernel(int64_t ni, const double* __restrict__ X, int64_t ldx,
                                                          const double* __restrict__ Y, int64_t ldy,
                                                          const double* __restrict__ wrow,
                                                          const int64_t* __restrict__ goff, int G,
                                                          double* __restrict__ dots)
{
    __shared__ double red[9][RT / 64];
    const int64_t k = blockIdx.x;
    const bool same = X == Y && ldx == ldy;
    const double* xc[3];
    const double* yc[3];
#pragma unroll
    for (int g = 0; g < 3; ++g) {
        const int64_t col = g < G ? goff[g] + k : goff[0] + k;
        xc[g] = X + col * ldx;
        yc[g] = Y + col * ldy;
    }
    double acc[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) acc[a][b] = 0.0;
    for (int64_t i = threadIdx.x; i < ni; i += RT) {
        const double wi = wrow ? wrow[i] : 1.0;
        double x[3], y[3];
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            if (g < G) {
                const double xv = xc[g][i];
                x[g] = xv * wi;
                y[g] = same ? xv : yc[g][i];
            } else {
                x[g] = 0.0;
                y[g] = 0.0;
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) acc[a][b] += x[a] * y[b];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            double v = acc[a][b];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if ((threadIdx.x & 63) == 0) red[3 * a + b][threadIdx.x >> 6] = v;
        }
    __syncthreads();
    if (threadIdx.x < 9) {
        const int a = threadIdx.x / 3, b = threadIdx.x % 3;
        if (a < G && b < G) {
            double t = 0.0;
            for (int q = 0; q < RT / 64; ++q) t += red[threadIdx.x][q];
            dots[(k * G + a) * G + b] = t;
        }
    }
}

// out[p] = alpha sum_{k in bin(p)} w(s, l(k)) sum_pairs dots[k, u-group, v-group], one thread per parameter.
__global__ void harm_binsum_kernel(const double* __restrict__ dots, int G, HarmTables t, int npar, double alpha,
                                   double* __restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npar) return;
    const int s = p / t.nbins, b = p % t.nbins;
    const int64_t lmin2 = (int64_t)t.lmin * t.lmin;
    const int64_t ka = (int64_t)t.lo[b] * t.lo[b] - lmin2, K = (int64_t)(t.hi[b] + 1) * (t.hi[b] + 1) - lmin2 - ka;
    const int* Pp = t.pairs + 4 * s;
    const int np = Pp[2] < 0 ? 1 : 2;
    double acc = 0.0;
    for (int64_t k = ka; k < ka + K; ++k) {
        const double a = t.w[s * t.lw + ell_of(k, t.lmin)];
        if (a == 0.0) continue;
        double term = 0.0;
        for (int e = 0; e < np; ++e) term += dots[(k * G + Pp[2 * e]) * G + Pp[2 * e + 1]];
        acc += a * term;
    }
    out[p] = alpha * acc;
}

// One block per lower-triangle entry (p1 >= p2) of F.
// M2: V^T C2^-1 V of a cross analysis (M itself otherwise):
// Tr[C1^-1 u' v'^T C2^-1 u v^T] = M1[v, u'] M2[u, v'].
__global__ __launch_bounds__(RT) void harm_fisher_kernel(const double* __restrict__ M,
                                                         const double* __restrict__ M2, int64_t ldm, HarmTables t,
                                                         int npar, double* __restrict__ F)
{
    const int64_t idx = blockIdx.x;
    int64_t p1 = (int64_t)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
    while ((p1 + 1) * (p1 + 2) / 2 <= idx) ++p1;
    while (p1 * (p1 + 1) / 2 > idx) --p1;
    const int64_t p2 = idx - p1 * (p1 + 1) / 2;
    const int s1 = (int)(p1 / t.nbins), b1 = (int)(p1 % t.nbins);
    const int s2 = (int)(p2 / t.nbins), b2 = (int)(p2 % t.nbins);
    const int64_t lmin2 = (int64_t)t.lmin * t.lmin;
    const int64_t k1a = (int64_t)t.lo[b1] * t.lo[b1] - lmin2, K1 = (int64_t)(t.hi[b1] + 1) * (t.hi[b1] + 1) - lmin2 - k1a;
    const int64_t k2a = (int64_t)t.lo[b2] * t.lo[b2] - lmin2, K2 = (int64_t)(t.hi[b2] + 1) * (t.hi[b2] + 1) - lmin2 - k2a;
    const int* P1 = t.pairs + 4 * s1;
    const int* P2 = t.pairs + 4 * s2;
    const int n1 = P1[2] < 0 ? 1 : 2, n2 = P2[2] < 0 ? 1 : 2;
    double acc = 0.0;
    for (int64_t q = threadIdx.x; q < K1 * K2; q += RT) {
        const int64_t k1 = k1a + q % K1, k2 = k2a + q / K1;
        const double a = t.w[s1 * t.lw + ell_of(k1, t.lmin)] * t.w[s2 * t.lw + ell_of(k2, t.lmin)];
        if (a == 0.0) continue;
        double term = 0.0;
        for (int e = 0; e < n1; ++e) {
            const int64_t u = t.goff[P1[2 * e]] + k1, v = t.goff[P1[2 * e + 1]] + k1;
            for (int f = 0; f < n2; ++f) {
                const int64_t u2 = t.goff[P2[2 * f]] + k2, v2 = t.goff[P2[2 * f + 1]] + k2;
                term += M[v + u2 * ldm] * M2[u + v2 * ldm];
            }
        }
        acc += a * term;
    }
    __shared__ double red[RT / 64];
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) {
        F[p1 + p2 * npar] = 0.5 * s;
        F[p2 + p1 * npar] = 0.5 * s;
    }
}

// One block per bin pair (b1 >= b2), all spectrum pairs at once: each mode
// pair (k1 in b1, k2 in b2) reads the G x G entries M[g k1, g' k2] (and
// those of M2) once, into thread-private LDS slots, and every (s1, s2) term
// M[v_e, u_f] M2[u_e, v_f] picks its two entries there by the pair table of
// the spectra (uniform indices, no bank conflicts).  The G^2 reads per mode
// pair replace the sum over spectrum pairs of their 2 n1 n2 reads in
// harm_fisher_kernel (9 against 2 x 81 for T, E, B).  Fixed-order sums.
constexpr int HFG = 3;    // mode groups (T, E, B)
constexpr int HFS = 6;    // spectra
template <int G>
__global__ __launch_bounds__(RT) void harm_fisher_bins_kernel(const double* __restrict__ M,
                                                              const double* __restrict__ M2, int64_t ldm,
                                                              HarmTables t, int nspec, int npar,
                                                              double* __restrict__ F)
{
    __shared__ double mv[HFG * HFG][RT];
    __shared__ double mv2[HFG * HFG][RT];
    __shared__ int ent[HFS * HFS][4][2];     // per (s1, s2): up to 4 (index into mv, index into mv2)
    __shared__ int nent[HFS * HFS];
    __shared__ double red[HFS * HFS][RT / 64];
    const int64_t idx = blockIdx.x;
    int b1 = (int)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
    while ((int64_t)(b1 + 1) * (b1 + 2) / 2 <= idx) ++b1;
    while ((int64_t)b1 * (b1 + 1) / 2 > idx) --b1;
    const int b2 = (int)(idx - (int64_t)b1 * (b1 + 1) / 2);
    const int tid = threadIdx.x;
    if (tid < nspec * nspec) {
        const int s1 = tid / nspec, s2 = tid % nspec;
        const int* P1 = t.pairs + 4 * s1;
        const int* P2 = t.pairs + 4 * s2;
        const int n1 = P1[2] < 0 ? 1 : 2, n2 = P2[2] < 0 ? 1 : 2;
        int c = 0;
        for (int e = 0; e < n1; ++e)
            for (int f = 0; f < n2; ++f) {
                // M[v_e(k1), u_f(k2)] M2[u_e(k1), v_f(k2)]
                ent[tid][c][0] = P1[2 * e + 1] * HFG + P2[2 * f];
                ent[tid][c][1] = P1[2 * e] * HFG + P2[2 * f + 1];
                ++c;
            }
        nent[tid] = c;
    }
    __syncthreads();
    const int64_t lmin2 = (int64_t)t.lmin * t.lmin;
    const int64_t k1a = (int64_t)t.lo[b1] * t.lo[b1] - lmin2, K1 = (int64_t)(t.hi[b1] + 1) * (t.hi[b1] + 1) - lmin2 - k1a;
    const int64_t k2a = (int64_t)t.lo[b2] * t.lo[b2] - lmin2, K2 = (int64_t)(t.hi[b2] + 1) * (t.hi[b2] + 1) - lmin2 - k2a;
    double acc[HFS * HFS];
#pragma unroll
    for (int i = 0; i < HFS * HFS; ++i) acc[i] = 0.0;
    for (int64_t q = tid; q < K1 * K2; q += RT) {
        const int64_t k1 = k1a + q % K1, k2 = k2a + q / K1;
        const int l1 = ell_of(k1, t.lmin), l2 = ell_of(k2, t.lmin);
        // the G^2 entries of M and M2 for this mode pair, all loads in flight together
        double v1[G][G], v2[G][G];
#pragma unroll
        for (int g1 = 0; g1 < G; ++g1)
#pragma unroll
            for (int g2 = 0; g2 < G; ++g2) {
                const int64_t o = (t.goff[g1] + k1) + (t.goff[g2] + k2) * ldm;
                v1[g1][g2] = M[o];
                v2[g1][g2] = M2[o];
            }
#pragma unroll
        for (int g1 = 0; g1 < G; ++g1)
#pragma unroll
            for (int g2 = 0; g2 < G; ++g2) {
                mv[g1 * HFG + g2][tid] = v1[g1][g2];
                mv2[g1 * HFG + g2][tid] = v2[g1][g2];
            }
#pragma unroll
        for (int s1 = 0; s1 < HFS; ++s1) {
            if (s1 >= nspec) break;
            const double w1 = t.w[s1 * t.lw + l1];
#pragma unroll
            for (int s2 = 0; s2 < HFS; ++s2) {
                if (s2 >= nspec) break;
                const int pq = s1 * nspec + s2;
                double term = 0.0;
                for (int c = 0; c < nent[pq]; ++c) term += mv[ent[pq][c][0]][tid] * mv2[ent[pq][c][1]][tid];
                acc[s1 * HFS + s2] += w1 * t.w[s2 * t.lw + l2] * term;
            }
        }
    }
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int i = 0; i < HFS * HFS; ++i) {
        double v = acc[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[i][wave] = v;
    }
    __syncthreads();
    if (tid < HFS * HFS) {
        const int s1 = tid / HFS, s2 = tid % HFS;
        if (s1 < nspec && s2 < nspec) {
            double v = 0.0;
            for (int w = 0; w < RT / 64; ++w) v += red[tid][w];
            const int p1 = s1 * t.nbins + b1, p2 = s2 * t.nbins + b2;
            // a diagonal bin pair gives each entry twice, (s1, s2) and (s2, s1): the lower one writes it
            if (b1 != b2 || p1 >= p2) {
                F[p1 + (int64_t)p2 * npar] = 0.5 * v;
                F[p2 + (int64_t)p1 * npar] = 0.5 * v;
            }
        }
    }
}

// out[p + j ldo] = alpha sum_{k in bin(p)} w sum_pairs sum_{i < ni} X[i + u_k ldx + j sxj] Y[i + v_k ldy + j syj]
__global__ __launch_bounds__(RT) void harm_pairsum_kernel(int64_t ni, const double* __restrict__ X, int64_t ldx,
                                                          int64_t sxj, const double* __restrict__ Y, int64_t ldy,
                                                          int64_t syj, HarmTables t, double alpha,
                                                          double* __restrict__ out, int64_t ldo)
{
    const int p = blockIdx.x;
    const int64_t j = blockIdx.y;
    const int s = p / t.nbins, b = p % t.nbins;
    const int64_t lmin2 = (int64_t)t.lmin * t.lmin;
    const int64_t ka = (int64_t)t.lo[b] * t.lo[b] - lmin2, K = (int64_t)(t.hi[b] + 1) * (t.hi[b] + 1) - lmin2 - ka;
    const int* Pp = t.pairs + 4 * s;
    const int np = Pp[2] < 0 ? 1 : 2;
    const double* Xj = X + j * sxj;
    const double* Yj = Y + j * syj;
    double acc = 0.0;
    if (ni == 1) {
        // one value per mode: threads over the modes
        for (int64_t kk = threadIdx.x; kk < K; kk += RT) {
            const int64_t k = ka + kk;
            const double a = t.w[s * t.lw + ell_of(k, t.lmin)];
            if (a == 0.0) continue;
            double term = 0.0;
            for (int e = 0; e < np; ++e)
                term += Xj[(t.goff[Pp[2 * e]] + k) * ldx] * Yj[(t.goff[Pp[2 * e + 1]] + k) * ldy];
            acc += a * term;
        }
    } else {
        // modes in turn (weight looked up once per mode), threads over the
        // contiguous rows of the two columns of each pair
        for (int64_t k = ka; k < ka + K; ++k) {
            const double a = t.w[s * t.lw + ell_of(k, t.lmin)];
            if (a == 0.0) continue;
            double term = 0.0;
            for (int e = 0; e < np; ++e) {
                const double* xc = Xj + (t.goff[Pp[2 * e]] + k) * ldx;
                const double* yc = Yj + (t.goff[Pp[2 * e + 1]] + k) * ldy;
                for (int64_t i = threadIdx.x; i < ni; i += RT) term += xc[i] * yc[i];
            }
            acc += a * term;
        }
    }
    __shared__ double red[RT / 64];
    const double r = block_sum(acc, red);
    if (threadIdx.x == 0) out[p + j * ldo] = alpha * r;
}

int check_tables(cfmi_ctx* c, int lmin, int nbins, const int* lo, const int* hi, int nspec, const int* pairs,
                 const int64_t* goff, const double* w, int lw)
{
    if (lmin < 0 || nbins <= 0 || nspec <= 0 || lw <= 0)
        return cfmi_fail(c, CFMI_EINVAL, "harmonic: lmin >= 0, nbins, nspec, lw > 0 required");
    if (!lo || !hi || !pairs || !goff || !w) return cfmi_fail(c, CFMI_EINVAL, "harmonic: null table");
    if ((int64_t)nspec * nbins > 65535 * 64) return cfmi_fail(c, CFMI_EINVAL, "harmonic: too many parameters");
    return CFMI_OK;
}

}  // namespace

extern "C" {

int cfmi_harmonic_modes(cfmi_ctx* c, int spin, int64_t npix, const double* vec, int lmin, int lmax, double* V,
                        int64_t ldv, int64_t ncols, int64_t row_a, int64_t row_b, int64_t col_a, int64_t col_b)
{
    if (!c) return CFMI_EINVAL;
    if (spin != 0 && spin != 2) return cfmi_fail(c, CFMI_EINVAL, "harmonic_modes: spin must be 0 or 2");
    if (npix < 0 || lmin < 0 || lmax < lmin) return cfmi_fail(c, CFMI_EINVAL, "harmonic_modes: bad sizes");
    if (spin == 2 && lmin < 2) return cfmi_fail(c, CFMI_EINVAL, "harmonic_modes: spin-2 modes start at l = 2");
    if (lmax > 8192) return cfmi_fail(c, CFMI_EINVAL, "harmonic_modes: lmax > 8192");
    if (npix == 0) return CFMI_OK;
    if (!vec || !V) return cfmi_fail(c, CFMI_EINVAL, "harmonic_modes: null pointer");
    if (row_a < 0 || row_b < 0 || col_a < 0 || col_b < 0 || ldv < row_a + npix || (spin == 2 && ldv < row_b + npix))
        return cfmi_fail(c, CFMI_EINVAL, "harmonic_modes: rows exceed ldv");
    const int64_t nm = (int64_t)(lmax + 1) * (lmax + 1) - (int64_t)lmin * lmin;
    if (col_a + nm > ncols || (spin == 2 && col_b + nm > ncols))
        return cfmi_fail(c, CFMI_EINVAL, "harmonic_modes: mode columns exceed ncols");
    dim3 grid((unsigned)((npix + HT - 1) / HT), (unsigned)(lmax + 1));
    hipLaunchKernelGGL(modes_kernel, grid, dim3(HT), 0, c->stream, spin, npix, vec, lmin, lmax, V, ldv, row_a,
                       row_b, col_a, col_b);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_harmonic_fisher(cfmi_ctx* c, const double* M, const double* M2, int64_t ldm, int lmin, int nbins,
                         const int* bin_lo, const int* bin_hi, int nspec, const int* pairs, const int64_t* goff,
                         const double* w, int lw, double* F)
{
    if (!c) return CFMI_EINVAL;
    int rc = check_tables(c, lmin, nbins, bin_lo, bin_hi, nspec, pairs, goff, w, lw);
    if (rc) return rc;
    if (!M || !F) return cfmi_fail(c, CFMI_EINVAL, "harmonic_fisher: null pointer");
    const int64_t npar = (int64_t)nspec * nbins;
    const int64_t nent = npar * (npar + 1) / 2;
    if (nent > 0x7fffffffLL) return cfmi_fail(c, CFMI_EINVAL, "harmonic_fisher: too many entries");
    HarmTables t{bin_lo, bin_hi, pairs, goff, w, lw, nbins, lmin};
    int G = 0;
    if (nspec <= HFS) {
        // the mode groups the spectra pair (a small device table, read once per Fisher matrix)
        std::vector<int> ph(4 * (size_t)nspec);
        CFMI_HIP(c, hipMemcpyAsync(ph.data(), pairs, ph.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        CFMI_HIP(c, hipStreamSynchronize(c->stream));
        for (int v : ph) G = std::max(G, v + 1);
    }
    if (nspec <= HFS && G <= HFG) {
        const dim3 grid((unsigned)((int64_t)nbins * (nbins + 1) / 2));
        if (G == 3)
            hipLaunchKernelGGL(harm_fisher_bins_kernel<3>, grid, dim3(RT), 0, c->stream, M, M2 ? M2 : M, ldm, t, nspec,
                               (int)npar, F);
        else if (G == 2)
            hipLaunchKernelGGL(harm_fisher_bins_kernel<2>, grid, dim3(RT), 0, c->stream, M, M2 ? M2 : M, ldm, t, nspec,
                               (int)npar, F);
        else
            hipLaunchKernelGGL(harm_fisher_bins_kernel<1>, grid, dim3(RT), 0, c->stream, M, M2 ? M2 : M, ldm, t, nspec,
                               (int)npar, F);
        CFMI_LAUNCH_CHECK(c);
        return CFMI_OK;
    }
    hipLaunchKernelGGL(harm_fisher_kernel, dim3((unsigned)nent), dim3(RT), 0, c->stream, M, M2 ? M2 : M, ldm, t,
                       (int)npar, F);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_harmonic_pairsum(cfmi_ctx* c, int64_t ni, const double* X, int64_t ldx, int64_t sxj, const double* Y,
                          int64_t ldy, int64_t syj, int64_t ncols, int lmin, int nbins, const int* bin_lo,
                          const int* bin_hi, int nspec, const int* pairs, const int64_t* goff, const double* w,
                          int lw, double alpha, double* out, int64_t ldo)
{
    if (!c) return CFMI_EINVAL;
    int rc = check_tables(c, lmin, nbins, bin_lo, bin_hi, nspec, pairs, goff, w, lw);
    if (rc) return rc;
    if (ni <= 0 || ncols <= 0 || ncols > 65535) return cfmi_fail(c, CFMI_EINVAL, "harmonic_pairsum: bad sizes");
    if (!X || !Y || !out) return cfmi_fail(c, CFMI_EINVAL, "harmonic_pairsum: null pointer");
    const int64_t npar = (int64_t)nspec * nbins;
    if (ldo < npar) return cfmi_fail(c, CFMI_EINVAL, "harmonic_pairsum: ldo < nspec * nbins");
    HarmTables t{bin_lo, bin_hi, pairs, goff, w, lw, nbins, lmin};
    hipLaunchKernelGGL(harm_pairsum_kernel, dim3((unsigned)npar, (unsigned)ncols), dim3(RT), 0, c->stream, ni, X,
                       ldx, sxj, Y, ldy, syj, t, alpha, out, ldo);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_harmonic_coldots(cfmi_ctx* c, int64_t ni, const double* X, int64_t ldx, const double* Y, int64_t ldy,
                          const double* wrow, int ngroups, int64_t nmodes, int lmin, int nbins, const int* bin_lo,
                          const int* bin_hi, int nspec, const int* pairs, const int64_t* goff, const double* w,
                          int lw, double alpha, double* out)
{
    if (!c) return CFMI_EINVAL;
    int rc = check_tables(c, lmin, nbins, bin_lo, bin_hi, nspec, pairs, goff, w, lw);
    if (rc) return rc;
    if (ni <= 0 || nmodes <= 0 || ngroups < 1 || ngroups > 3)
        return cfmi_fail(c, CFMI_EINVAL, "harmonic_coldots: ni, nmodes > 0 and 1 <= ngroups <= 3 required");
    if (nmodes > 0x7fffffffLL) return cfmi_fail(c, CFMI_EINVAL, "harmonic_coldots: too many modes");
    if (!X || !Y || !out) return cfmi_fail(c, CFMI_EINVAL, "harmonic_coldots: null pointer");
    if (ldx < ni || ldy < ni) return cfmi_fail(c, CFMI_EINVAL, "harmonic_coldots: leading dimension < ni");
    double* dots = (double*)cfmi_workspace(c, (size_t)nmodes * ngroups * ngroups * sizeof(double));
    if (!dots) return cfmi_fail(c, CFMI_ENOMEM, "harmonic_coldots: workspace");
    hipLaunchKernelGGL(harm_coldots_kernel, dim3((unsigned)nmodes), dim3(RT), 0, c->stream, ni, X, ldx, Y, ldy, wrow,
                       goff, ngroups, dots);
    CFMI_LAUNCH_CHECK(c);
    const int npar = nspec * nbins;
    HarmTables t{bin_lo, bin_hi, pairs, goff, w, lw, nbins, lmin};
    hipLaunchKernelGGL(harm_binsum_kernel, dim3((unsigned)((npar + 255) / 256)), dim3(256), 0, c->stream, dots,
                       ngroups, t, npar, alpha, out);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

}  // extern "C"
