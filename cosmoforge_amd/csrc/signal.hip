// Pixel-pair recurrence kernels (FP64 VALU): pointings, signal covariance
// blocks and binned derivative stacks.
//
// Reference algorithms (ggalloni/CosmoForge 1.3.0, src/cosmoforge.cosmocore/cosmocore):
//   recurrences      basics/legendre.py:10-149  (P_l, P^22_l with F1/F2, P^02_l)
//   rotation angles  basics/geometry.py:14-110
//   pair loops       signal_kernels.py:164-387
//   pointings        signal_kernels.py:113-161 + HEALPix pix2ang (Gorski et al. 2005)
//
// Design: one workgroup owns a 32 x 32 tile of pixel pairs; each thread runs
// the l-recurrence for four pairs at once (ILP across independent chains),
// with the recurrence coefficients and the normalised weights read as
// wave-uniform values.  Rotation angles enter only through cos/sin of
// 2 acos(d), evaluated algebraically (2d^2 - 1, 2 d sqrt((1-d)(1+d))), so no
// transcendental is issued per pair.  Each finished bin is written as a
// coalesced tile and, through an LDS transpose, as its mirror tile.
#include "cfmi_internal.h"

#include <cmath>
#include <cstring>

namespace {

constexpr int TILE = 32;
constexpr int PPT = 4;  // pairs per thread (TILE*TILE / 256)

enum Kind { K00S = 0, K00G = 1, K22 = 2, K02 = 3 };

// Coefficient / weight table layout in ctx->d_coef (length 5 * (L+1)):
//   [0]      cA_l   recurrence factor of x P_{l-1}
//   [L+1]    cB_l   recurrence factor of P_{l-2}
//   [2(L+1)] w0_l   normalised weights (up to three spectra)
//   [3(L+1)] w1_l
//   [4(L+1)] w2_l
__global__ void prep_table_kernel(int kind, int L, const double* w0, const double* w1, const double* w2,
                                  double* tab)
{
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l > L) return;
    const double inv4pi = 1.0 / (4.0 * 3.14159265358979323846);
    double cA = 0.0, cB = 0.0, norm = 0.0;
    if (kind == K00S || kind == K00G) {
        if (l >= 3) {
            cA = (2.0 * l - 1.0) / l;
            cB = (l - 1.0) / l;
        }
        norm = (2.0 * l + 1.0) * inv4pi;
    } else {
        if (l >= 3) {
            cA = (2.0 * l - 1.0) / (l - 2.0);
            cB = (l + 1.0) / (l - 2.0);
        }
        if (l >= 2) {
            double f = (double)(l + 2) * (l + 1) * l * (l - 1);
            norm = (2.0 * l + 1.0) * inv4pi / (kind == K22 ? f : sqrt(f));
        }
    }
    tab[l] = cA;
    tab[(L + 1) + l] = cB;
    tab[2 * (L + 1) + l] = w0 ? w0[l] * norm : 0.0;
    tab[3 * (L + 1) + l] = w1 ? w1[l] * norm : 0.0;
    tab[4 * (L + 1) + l] = w2 ? w2[l] * norm : 0.0;
}

struct PairArgs {
    const double* vr;  // rows j (n_r x 3)
    const double* vc;  // cols i (n_c x 3)
    int64_t nr, nc;
    const double* tab;
    int L;
    int lrun;          // last l the recurrence must reach
    const int* blo;
    const int* bhi;
    int nbins;
    int sum_mode;      // 1: single summed output (signal), 0: per-spectrum stacks
    void* out[3];      // double*, or float* for FP32 derivative stacks
    void* mir[3];
    int64_t ld, stride, ldm, stride_m;
    int64_t np;        // plane offset: n (22) or n_p (02)
    int64_t ntile;     // tiles per side (symmetric kinds)
    double* state;     // per-pair recurrence state (P_l, P_{l-1}) after the last bin, or NULL
    int lresume;       // > 0: start at l = lresume from the saved state
    int64_t ps, ps_m;  // three-piece BF16 output: plane strides of out / mir
};

// Element type of a three-piece BF16 output (x = x0 + x1 + x2, planes ps apart).
struct Bf3 {
    unsigned short v;
};

__device__ __forceinline__ unsigned short f2bf16(float f)
{
    unsigned u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ double bf16f(unsigned short v)
{
    unsigned u = (unsigned)v << 16;
    float f;
    memcpy(&f, &u, 4);
    return (double)f;
}

template <typename TO>
__device__ __forceinline__ void put(TO* p, int64_t i, double v, int64_t) { p[i] = (TO)v; }
template <>
__device__ __forceinline__ void put<Bf3>(Bf3* p, int64_t i, double v, int64_t ps)
{
    unsigned short* q = reinterpret_cast<unsigned short*>(p) + i;
    const unsigned short h = f2bf16((float)v);
    double r = v - bf16f(h);
    const unsigned short m = f2bf16((float)r);
    r -= bf16f(m);
    q[0] = h;
    q[ps] = m;
    q[2 * ps] = f2bf16((float)r);
}

// cos/sin of the rotation angles of geometry.py:49-110, trig-free.
// For pairs on a common meridian d is +-1 up to rounding and sin(2 acos d)
// is rounding noise of order sqrt(eps); the reference's values carry that
// noise, so d must be formed with exactly the reference's operation order:
// no FMA contraction here.
__device__ __forceinline__ void rot_angles(const double* r1, const double* r2, double& c12, double& s12,
                                           double& c21, double& s21)
{
#pragma clang fp contract(off)
    double r12x = r1[1] * r2[2] - r1[2] * r2[1];
    double r12y = r1[2] * r2[0] - r1[0] * r2[2];
    double r12z = r1[0] * r2[1] - r1[1] * r2[0];
    double mod = sqrt(r12x * r12x + r12y * r12y + r12z * r12z);
    if (mod < 1e-8) {
        c12 = 1.0; s12 = 0.0; c21 = 1.0; s21 = 0.0;
        return;
    }
    r12x /= mod; r12y /= mod; r12z /= mod;
    const double eps = 3.14159265358979311599796346854 / 180.0 / 3600.0 / 100.0;
    double p1x = -r1[1], p1y = r1[0];
    double nm = sqrt(p1x * p1x + p1y * p1y);
    if (nm < 1e-8) { p1x += eps; nm = sqrt(p1x * p1x + p1y * p1y); }
    p1x /= nm; p1y /= nm;
    double p2x = -r2[1], p2y = r2[0];
    nm = sqrt(p2x * p2x + p2y * p2y);
    if (nm < 1e-8) { p2x += eps; nm = sqrt(p2x * p2x + p2y * p2y); }
    p2x /= nm; p2y /= nm;
    double d1 = r12x * p1x + r12y * p1y;
    d1 = d1 > 1.0 ? 1.0 : (d1 < -1.0 ? -1.0 : d1);
    double sg1 = r12z > 0.0 ? 1.0 : -1.0;
    double d2 = -r12x * p2x - r12y * p2y;
    d2 = d2 > 1.0 ? 1.0 : (d2 < -1.0 ? -1.0 : d2);
    double sg2 = -r12z > 0.0 ? 1.0 : -1.0;
    // a = 2 acos(d) * sg  =>  cos a = 2 d^2 - 1, sin a = sg * 2 d sqrt(1 - d^2)
    c12 = 2.0 * d1 * d1 - 1.0;
    s12 = sg1 * 2.0 * d1 * sqrt((1.0 - d1) * (1.0 + d1));
    c21 = 2.0 * d2 * d2 - 1.0;
    s21 = sg2 * 2.0 * d2 * sqrt((1.0 - d2) * (1.0 + d2));
}

// x_ij = v_j . v_i summed left to right without contraction, as the
// reference's scalar loop does.
__device__ __forceinline__ double dot3_exact(const double* a, const double* b)
{
#pragma clang fp contract(off)
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Write a tile plane: value v[q] belongs to pair (j0 + tj, i0 + ti0 + 8q).
// direct: dst[j + i*ld]; transposed (if dstT): dstT[i + j*ldT].
// diag_tile: only pairs with j >= i are valid (others skipped).
template <typename TO>
__device__ __forceinline__ void write_plane(double (*sT)[TILE + 1], const double* v, TO* dst, int64_t ld,
                                            TO* dstT, int64_t ldT, int64_t j0, int64_t i0, int64_t nr,
                                            int64_t nc, bool diag_tile, int64_t ps = 0, int64_t psT = 0)
{
    const int t = threadIdx.x;
    const int tj = t & (TILE - 1), ti0 = t >> 5;
    if (dst) {
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            int ti = ti0 + 8 * q;
            int64_t j = j0 + tj, i = i0 + ti;
            if (j < nr && i < nc && (!diag_tile || tj >= ti)) put(dst, j + i * ld, v[q], ps);
        }
    }
    if (dstT) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PPT; ++q) sT[ti0 + 8 * q][tj] = v[q];
        __syncthreads();
        const int a = t & (TILE - 1), b0 = t >> 5;
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            int b = b0 + 8 * q;  // local row index j
            int64_t i = i0 + a, j = j0 + b;
            if (j < nr && i < nc && (!diag_tile || b >= a)) put(dstT, i + j * ldT, sT[a][b], psT);
        }
    }
}

template <int KIND, typename TO>
__global__ __launch_bounds__(256) void pair_kernel(PairArgs p)
{
    __shared__ double sT[TILE][TILE + 1];
    int64_t tj_, ti_;
    bool diag_tile = false;
    if (KIND == K00S || KIND == K22) {
        int64_t k = blockIdx.x;
        int64_t r = (int64_t)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
        while ((r + 1) * (r + 2) / 2 <= k) ++r;
        while (r * (r + 1) / 2 > k) --r;
        tj_ = r;
        ti_ = k - r * (r + 1) / 2;
        diag_tile = tj_ == ti_;
    } else {
        tj_ = blockIdx.x;
        ti_ = blockIdx.y;
    }
    const int64_t j0 = tj_ * TILE, i0 = ti_ * TILE;
    const int t = threadIdx.x;
    const int tj = t & (TILE - 1), ti0 = t >> 5;

    // Per-pair argument x and rotation cos/sin.
    double x[PPT], y1[PPT];
    double c1[PPT], s1[PPT], c2[PPT], s2[PPT];
    const int64_t j = j0 + tj;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        int64_t i = i0 + ti0 + 8 * q;
        bool valid = j < p.nr && i < p.nc;
        bool self = (KIND == K00S || KIND == K22) && j == i;
        double xv = 1.0;
        c1[q] = 1.0; s1[q] = 0.0; c2[q] = 1.0; s2[q] = 0.0;
        if (valid && !self) {
            const double* vj = p.vr + 3 * j;
            const double* vi = p.vc + 3 * i;
            xv = dot3_exact(vj, vi);
            if (KIND == K22) {
                double a, b, c, d;
                rot_angles(vj, vi, a, b, c, d);
                c1[q] = a; s1[q] = -b; c2[q] = c; s2[q] = -d;
            } else if (KIND == K02) {
                double a, b, c, d;
                rot_angles(vj, vi, a, b, c, d);
                c1[q] = a; s1[q] = b;
            }
        }
        x[q] = xv;
        y1[q] = 1.0 - xv * xv;
    }

    // The recurrence and weight tables, read through the constant address
    // space: with l uniform these become scalar loads (one per step for the
    // whole wave) instead of a vector load per lane.
    typedef const __attribute__((address_space(4))) double ctab;
    const int L1 = p.L + 1;
    ctab* cA = (ctab*)p.tab;
    ctab* cB = (ctab*)p.tab + L1;
    ctab* w0 = (ctab*)p.tab + 2 * L1;
    ctab* w1 = (ctab*)p.tab + 3 * L1;
    ctab* w2 = (ctab*)p.tab + 4 * L1;

    // Accumulators: 00 -> acc0; 22 -> A1,A2 (w0), B1,B2 (w1), E (w2); 02 -> tq (w0), tu (w1)
    double acc[5][PPT];
#pragma unroll
    for (int a = 0; a < 5; ++a)
#pragma unroll
        for (int q = 0; q < PPT; ++q) acc[a][q] = 0.0;
    double pm1[PPT], pm2[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        pm1[q] = 0.0;
        pm2[q] = 0.0;
    }

    // Recurrence state slot of this thread's pairs (tile-major).
    double* st = p.state ? p.state + ((int64_t)(blockIdx.x + (int64_t)blockIdx.y * gridDim.x) * 256 + t) * (2 * PPT)
                         : nullptr;
    // l, and with it the table reads and the bin tests below, is the same
    // for every lane: decided from kernel arguments only, so that it lives
    // in a scalar register (a per-lane l turns every table read into a
    // vector load and every branch into an exec-masked one)
    int lstart = (KIND == K00S || KIND == K00G) ? 0 : 2;
    if (p.state && p.lresume > lstart) {
        lstart = __builtin_amdgcn_readfirstlane(p.lresume);
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            pm1[q] = st[2 * q];
            pm2[q] = st[2 * q + 1];
        }
    }

    // A null bin table means one bin covering l = 0..L (the signal matrix).
    int b = 0;
    int lo = p.blo ? p.blo[0] : 0, hi = p.bhi ? p.bhi[0] : p.L;
    for (int l = lstart; l <= p.lrun; ++l) {
        double P[PPT], F1[PPT], F2[PPT];
        if (KIND == K00S || KIND == K00G) {
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                if (l == 0) P[q] = 1.0;
                else if (l == 1) P[q] = x[q];
                else if (l == 2) P[q] = 1.5 * x[q] * x[q] - 0.5;
                else P[q] = cA[l] * x[q] * pm1[q] - cB[l] * pm2[q];
            }
        } else {
            const double fl = (double)l;
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                if (l == 2) {
                    P[q] = KIND == K22 ? 3.0 : 3.0 * y1[q];
                    pm1[q] = 0.0;
                } else {
                    P[q] = cA[l] * x[q] * pm1[q] - cB[l] * pm2[q];
                }
                if (KIND == K22) {
                    if (l == 2) {
                        F1[q] = 6.0 * (1.0 + x[q] * x[q]);
                        F2[q] = -12.0 * x[q];
                    } else {
                        F1[q] = -(2.0 * fl - 8.0 + fl * (fl - 1.0) * y1[q]) * P[q] + (2.0 * fl + 4.0) * x[q] * pm1[q];
                        F2[q] = 4.0 * (-(fl - 1.0) * x[q] * P[q] + (fl + 2.0) * pm1[q]);
                    }
                }
            }
        }
        if (l >= lo && l <= hi) {
            const double a0 = w0[l], a1 = w1[l], a2 = w2[l];
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                if (KIND == K00S || KIND == K00G) {
                    acc[0][q] += a0 * P[q];
                } else if (KIND == K22) {
                    acc[0][q] += a0 * F1[q];
                    acc[1][q] += a0 * F2[q];
                    acc[2][q] += a1 * F1[q];
                    acc[3][q] += a1 * F2[q];
                    acc[4][q] += a2 * (F1[q] + F2[q]);
                } else {
                    acc[0][q] += a0 * P[q];
                    acc[1][q] += a1 * P[q];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            pm2[q] = pm1[q];
            pm1[q] = P[q];
        }
        if (l == hi) {
            // ---- emit bin b ----
            if (KIND == K00S || KIND == K00G) {
                TO* o = p.out[0] ? (TO*)p.out[0] + b * p.stride : nullptr;
                TO* m = KIND == K00S ? (o) : (p.mir[0] ? (TO*)p.mir[0] + b * p.stride_m : nullptr);
                int64_t ldm = KIND == K00S ? p.ld : p.ldm;
                write_plane(sT, acc[0], o, p.ld, m, ldm, j0, i0, p.nr, p.nc, diag_tile, p.ps,
                            KIND == K00S ? p.ps : p.ps_m);
            } else if (KIND == K22) {
                const int nspec = p.sum_mode ? 1 : 3;
                for (int s = 0; s < nspec; ++s) {
                    TO* o = (TO*)p.out[s];
                    if (!o) continue;
                    o += b * p.stride;
                    double va[PPT], vb[PPT], vc[PPT], vd[PPT];
#pragma unroll
                    for (int q = 0; q < PPT; ++q) {
                        double qq, uu, qu;
                        if (p.sum_mode) {
                            qq = acc[0][q] - acc[3][q];
                            uu = acc[2][q] - acc[1][q];
                            qu = acc[4][q];
                        } else if (s == 0) {
                            qq = acc[0][q]; uu = -acc[1][q]; qu = 0.0;
                        } else if (s == 1) {
                            qq = -acc[3][q]; uu = acc[2][q]; qu = 0.0;
                        } else {
                            qq = 0.0; uu = 0.0; qu = acc[4][q];
                        }
                        double c1c2 = c1[q] * c2[q], s1s2 = s1[q] * s2[q];
                        double c1s2 = c1[q] * s2[q], s1c2 = s1[q] * c2[q];
                        va[q] = qq * c1c2 + uu * s1s2 + qu * (c1s2 + s1c2);
                        vb[q] = qq * s1s2 + uu * c1c2 - qu * (c1s2 + s1c2);
                        vc[q] = -qq * s1c2 + uu * c1s2 + qu * (c1c2 - s1s2);
                        vd[q] = -qq * c1s2 + uu * s1c2 + qu * (c1c2 - s1s2);
                    }
                    const int64_t n = p.np, ld = p.ld;
                    TO* QQ = o;
                    TO* UU = o + n + n * ld;
                    TO* UQ = o + n;
                    TO* QU = o + n * ld;
                    write_plane(sT, va, QQ, ld, QQ, ld, j0, i0, p.nr, p.nc, diag_tile, p.ps, p.ps);
                    write_plane(sT, vb, UU, ld, UU, ld, j0, i0, p.nr, p.nc, diag_tile, p.ps, p.ps);
                    write_plane(sT, vc, UQ, ld, QU, ld, j0, i0, p.nr, p.nc, diag_tile, p.ps, p.ps);
                    write_plane(sT, vd, QU, ld, UQ, ld, j0, i0, p.nr, p.nc, diag_tile, p.ps, p.ps);
                }
            } else {  // K02: rows j are spin-2 pixels, cols i spin-0
                const int nspec = p.sum_mode ? 1 : 2;
                for (int s = 0; s < nspec; ++s) {
                    TO* o = p.out[s] ? (TO*)p.out[s] + b * p.stride : nullptr;
                    TO* m = p.mir[s] ? (TO*)p.mir[s] + b * p.stride_m : nullptr;
                    if (!o && !m) continue;
                    double vq[PPT], vu[PPT];
#pragma unroll
                    for (int q = 0; q < PPT; ++q) {
                        double tq, tu;
                        if (p.sum_mode) { tq = -acc[0][q]; tu = -acc[1][q]; }
                        else if (s == 0) { tq = -acc[0][q]; tu = 0.0; }
                        else { tq = 0.0; tu = -acc[1][q]; }
                        vq[q] = tq * c1[q] - tu * s1[q];
                        vu[q] = tq * s1[q] + tu * c1[q];
                    }
                    const int64_t n = p.np;
                    write_plane(sT, vq, o, p.ld, m, p.ldm, j0, i0, p.nr, p.nc, false, p.ps, p.ps_m);
                    write_plane(sT, vu, o ? o + n : nullptr, p.ld, m ? m + n * p.ldm : nullptr, p.ldm, j0, i0,
                                p.nr, p.nc, false, p.ps, p.ps_m);
                }
            }
#pragma unroll
            for (int a = 0; a < 5; ++a)
#pragma unroll
                for (int q = 0; q < PPT; ++q) acc[a][q] = 0.0;
            ++b;
            if (b >= p.nbins || !p.blo) break;
            lo = p.blo[b];
            hi = p.bhi[b];
        }
    }
    if (st) {
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            st[2 * q] = pm1[q];
            st[2 * q + 1] = pm2[q];
        }
    }
}


int launch_pairs(cfmi_ctx* c, int kind, PairArgs& a, const double* w0, const double* w1, const double* w2,
                 int dtype = CFMI_F64)
{
    if (dtype != CFMI_F64 && dtype != CFMI_F32 && dtype != CFMI_BF16X3)
        return cfmi_fail(c, CFMI_EINVAL, "unsupported output dtype %d", dtype);
    if (dtype == CFMI_BF16X3) {
        a.ps = a.stride * a.nbins;
        a.ps_m = a.stride_m * a.nbins;
    }
    if (a.L < 0 || a.L > CFMI_MAX_ELL) return cfmi_fail(c, CFMI_EINVAL, "lmax %d out of range", a.L);
    if (a.nr <= 0 || a.nc <= 0 || a.nbins <= 0) return CFMI_OK;
    if (!c->d_coef) CFMI_HIP(c, hipMalloc((void**)&c->d_coef, 5 * (CFMI_MAX_ELL + 1) * sizeof(double)));
    hipLaunchKernelGGL(prep_table_kernel, dim3((a.L + 256) / 256), dim3(256), 0, c->stream, kind, a.L, w0, w1,
                       w2, c->d_coef);
    CFMI_LAUNCH_CHECK(c);
    a.tab = c->d_coef;
    int64_t tr = (a.nr + TILE - 1) / TILE, tc = (a.nc + TILE - 1) / TILE;
#define CFMI_PAIRS(K_, grid_)                                                                           \
    do {                                                                                               \
        if (dtype == CFMI_F32) hipLaunchKernelGGL((pair_kernel<K_, float>), grid_, dim3(256), 0, c->stream, a); \
        else if (dtype == CFMI_BF16X3)                                                                 \
            hipLaunchKernelGGL((pair_kernel<K_, Bf3>), grid_, dim3(256), 0, c->stream, a);             \
        else hipLaunchKernelGGL((pair_kernel<K_, double>), grid_, dim3(256), 0, c->stream, a);          \
    } while (0)
    if (kind == K00S || kind == K22) {
        dim3 g((unsigned)(tr * (tr + 1) / 2));
        if (kind == K00S) CFMI_PAIRS(K00S, g);
        else CFMI_PAIRS(K22, g);
    } else {
        dim3 g((unsigned)tr, (unsigned)tc);
        if (kind == K00G) CFMI_PAIRS(K00G, g);
        else CFMI_PAIRS(K02, g);
    }
#undef CFMI_PAIRS
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

// ---- pointings ----------------------------------------------------------
__device__ __forceinline__ int64_t isqrt64(int64_t v)
{
    int64_t r = (int64_t)sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

__device__ __forceinline__ int64_t compress_bits(int64_t v)
{
    int64_t out = 0;
    for (int bit = 0; v; ++bit, v >>= 2) out |= (v & 1) << bit;
    return out;
}

__global__ void pointings_kernel(int64_t nside, int nest, const int64_t* pix, int64_t n, double* vec)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double PI = 3.14159265358979323846;
    const double HALFPI = 0.5 * PI;
    int64_t p = pix[k];
    int64_t npix = 12 * nside * nside;
    double fact2 = 4.0 / (double)npix;
    double fact1 = (double)(nside << 1) * fact2;
    double z, phi, sth = 0.0;
    bool have_sth = false;
    if (!nest) {
        int64_t ncap = 2 * nside * (nside - 1);
        if (p < ncap) {
            int64_t iring = (1 + isqrt64(1 + 2 * p)) >> 1;
            int64_t iphi = (p + 1) - 2 * iring * (iring - 1);
            double tmp = (double)(iring * iring) * fact2;
            z = 1.0 - tmp;
            if (z > 0.99) { sth = sqrt(tmp * (2.0 - tmp)); have_sth = true; }
            phi = ((double)iphi - 0.5) * HALFPI / (double)iring;
        } else if (p < npix - ncap) {
            int64_t nl4 = 4 * nside;
            int64_t ip = p - ncap;
            int64_t tmp = ip / nl4;
            int64_t iring = tmp + nside;
            int64_t iphi = ip - nl4 * tmp + 1;
            double fodd = ((iring + nside) & 1) ? 1.0 : 0.5;
            z = (double)(2 * nside - iring) * fact1;
            phi = ((double)iphi - fodd) * PI * 0.75 * fact1;
        } else {
            int64_t ip = npix - p;
            int64_t iring = (1 + isqrt64(2 * ip - 1)) >> 1;
            int64_t iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
            double tmp = (double)(iring * iring) * fact2;
            z = tmp - 1.0;
            if (z < -0.99) { sth = sqrt(tmp * (2.0 - tmp)); have_sth = true; }
            phi = ((double)iphi - 0.5) * HALFPI / (double)iring;
        }
    } else {
        const int jrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
        const int jpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};
        int order = 0;
        while ((1LL << order) < nside) ++order;
        int64_t npface = nside * nside;
        int64_t face = p / npface;
        int64_t ipf = p & (npface - 1);
        int64_t ix = compress_bits(ipf), iy = compress_bits(ipf >> 1);
        int64_t jr = ((int64_t)jrll[face] << order) - ix - iy - 1;
        int64_t nr;
        if (jr < nside) {
            nr = jr;
            double tmp = (double)(nr * nr) * fact2;
            z = 1.0 - tmp;
            if (z > 0.99) { sth = sqrt(tmp * (2.0 - tmp)); have_sth = true; }
        } else if (jr > 3 * nside) {
            nr = 4 * nside - jr;
            double tmp = (double)(nr * nr) * fact2;
            z = tmp - 1.0;
            if (z < -0.99) { sth = sqrt(tmp * (2.0 - tmp)); have_sth = true; }
        } else {
            nr = nside;
            z = (double)(2 * nside - jr) * fact1;
        }
        int64_t t = (int64_t)jpll[face] * nr + ix - iy;
        if (t < 0) t += 8 * nr;
        phi = nr == nside ? 0.75 * HALFPI * (double)t * fact1 : (0.5 * HALFPI * (double)t) / (double)nr;
    }
    double theta = have_sth ? atan2(sth, z) : acos(z);
    double st = sin(theta);
    double xx = st * cos(phi), yy = st * sin(phi), zz = cos(theta);
    double norm = sqrt(xx * xx + yy * yy + zz * zz);
    vec[3 * k + 0] = xx / norm;
    vec[3 * k + 1] = yy / norm;
    vec[3 * k + 2] = zz / norm;
}

// Host evaluation with the platform libm (the reference's scalar loop).
void pointings_host_impl(int64_t nside, int nest, const int64_t* pix, int64_t n, double* vec)
{
#pragma clang fp contract(off)
    const double PI = 3.14159265358979323846;
    const double HALFPI = 0.5 * PI;
    const int jrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
    const int jpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};
    int64_t npix = 12 * nside * nside;
    double fact2 = 4.0 / (double)npix;
    double fact1 = (double)(nside << 1) * fact2;
    int order = 0;
    while ((1LL << order) < nside) ++order;
    for (int64_t k = 0; k < n; ++k) {
        int64_t p = pix[k];
        double z, phi, sth = 0.0;
        bool have_sth = false;
        if (!nest) {
            int64_t ncap = 2 * nside * (nside - 1);
            if (p < ncap) {
                int64_t v = 1 + 2 * p, r = (int64_t)std::sqrt((double)v);
                while (r * r > v) --r;
                while ((r + 1) * (r + 1) <= v) ++r;
                int64_t iring = (1 + r) >> 1;
                int64_t iphi = (p + 1) - 2 * iring * (iring - 1);
                double tmp = (double)(iring * iring) * fact2;
                z = 1.0 - tmp;
                if (z > 0.99) { sth = std::sqrt(tmp * (2.0 - tmp)); have_sth = true; }
                phi = ((double)iphi - 0.5) * HALFPI / (double)iring;
            } else if (p < npix - ncap) {
                int64_t nl4 = 4 * nside;
                int64_t ip = p - ncap;
                int64_t tmp = ip / nl4;
                int64_t iring = tmp + nside;
                int64_t iphi = ip - nl4 * tmp + 1;
                double fodd = ((iring + nside) & 1) ? 1.0 : 0.5;
                z = (double)(2 * nside - iring) * fact1;
                phi = ((double)iphi - fodd) * PI * 0.75 * fact1;
            } else {
                int64_t ip = npix - p;
                int64_t v = 2 * ip - 1, r = (int64_t)std::sqrt((double)v);
                while (r * r > v) --r;
                while ((r + 1) * (r + 1) <= v) ++r;
                int64_t iring = (1 + r) >> 1;
                int64_t iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
                double tmp = (double)(iring * iring) * fact2;
                z = tmp - 1.0;
                if (z < -0.99) { sth = std::sqrt(tmp * (2.0 - tmp)); have_sth = true; }
                phi = ((double)iphi - 0.5) * HALFPI / (double)iring;
            }
        } else {
            int64_t npface = nside * nside;
            int64_t face = p / npface;
            int64_t ipf = p & (npface - 1);
            int64_t ix = 0, iy = 0;
            for (int bit = 0, v = 0; bit < 2 * order + 2; bit += 2, ++v) {
                ix |= ((ipf >> bit) & 1) << v;
                iy |= ((ipf >> (bit + 1)) & 1) << v;
            }
            int64_t jr = ((int64_t)jrll[face] << order) - ix - iy - 1;
            int64_t nr;
            if (jr < nside) {
                nr = jr;
                double tmp = (double)(nr * nr) * fact2;
                z = 1.0 - tmp;
                if (z > 0.99) { sth = std::sqrt(tmp * (2.0 - tmp)); have_sth = true; }
            } else if (jr > 3 * nside) {
                nr = 4 * nside - jr;
                double tmp = (double)(nr * nr) * fact2;
                z = tmp - 1.0;
                if (z < -0.99) { sth = std::sqrt(tmp * (2.0 - tmp)); have_sth = true; }
            } else {
                nr = nside;
                z = (double)(2 * nside - jr) * fact1;
            }
            int64_t t = (int64_t)jpll[face] * nr + ix - iy;
            if (t < 0) t += 8 * nr;
            phi = nr == nside ? 0.75 * HALFPI * (double)t * fact1 : (0.5 * HALFPI * (double)t) / (double)nr;
        }
        double theta = have_sth ? std::atan2(sth, z) : std::acos(z);
        double st = std::sin(theta);
        double xx = st * std::cos(phi), yy = st * std::sin(phi), zz = std::cos(theta);
        double norm = std::sqrt(xx * xx + yy * yy + zz * zz);
        vec[3 * k + 0] = xx / norm;
        vec[3 * k + 1] = yy / norm;
        vec[3 * k + 2] = zz / norm;
    }
}

}  // namespace

extern "C" {

int cfmi_pointings_host(int64_t nside, int nest, const int64_t* pix, int64_t n, double* vec)
{
    if (nside < 1 || (nest && (nside & (nside - 1)))) return CFMI_EINVAL;
    int64_t npix = 12 * nside * nside;
    for (int64_t k = 0; k < n; ++k)
        if (pix[k] < 0 || pix[k] >= npix) return CFMI_EINVAL;
    pointings_host_impl(nside, nest, pix, n, vec);
    return CFMI_OK;
}

int cfmi_pointings(cfmi_ctx* c, int64_t nside, int nest, const int64_t* pix, int64_t n, double* vec)
{
    if (!c) return CFMI_EINVAL;
    if (nside < 1 || (nest && (nside & (nside - 1))))
        return cfmi_fail(c, CFMI_EINVAL, "pointings: invalid nside %lld", (long long)nside);
    if (n <= 0) return CFMI_OK;
    hipLaunchKernelGGL(pointings_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, nside, nest,
                       pix, n, vec);
    CFMI_LAUNCH_CHECK(c);
    return CFMI_OK;
}

int cfmi_signal_00(cfmi_ctx* c, const double* cl, int lmax, const double* vec_r, int64_t n_r,
                   const double* vec_c, int64_t n_c, int symmetric, double* out, int64_t ld, double* mirror,
                   int64_t ld_mirror)
{
    if (!c) return CFMI_EINVAL;
    if (symmetric && (n_r != n_c || mirror))
        return cfmi_fail(c, CFMI_EINVAL, "signal_00: symmetric block needs n_r == n_c and no mirror");
    PairArgs a = {};
    a.vr = vec_r; a.vc = vec_c; a.nr = n_r; a.nc = n_c; a.L = lmax; a.lrun = lmax;
    a.blo = nullptr; a.bhi = nullptr; a.nbins = 1; a.sum_mode = 1;
    a.out[0] = out; a.mir[0] = mirror; a.ld = ld; a.ldm = ld_mirror;
    return launch_pairs(c, symmetric ? K00S : K00G, a, cl, nullptr, nullptr);
}

int cfmi_signal_22(cfmi_ctx* c, const double* cl_ee, const double* cl_bb, const double* cl_eb, int lmax,
                   const double* vec, int64_t n, double* out, int64_t ld)
{
    if (!c) return CFMI_EINVAL;
    if (lmax < 2) return cfmi_fill(c, CFMI_F64, 2 * n, 2 * n, out, ld, 0.0);
    PairArgs a = {};
    a.vr = vec; a.vc = vec; a.nr = n; a.nc = n; a.L = lmax; a.lrun = lmax;
    a.blo = nullptr; a.bhi = nullptr; a.nbins = 1; a.sum_mode = 1;
    a.out[0] = out; a.ld = ld; a.np = n;
    return launch_pairs(c, K22, a, cl_ee, cl_bb, cl_eb);
}

int cfmi_signal_02(cfmi_ctx* c, const double* cl_te, const double* cl_tb, int lmax, const double* vec_t,
                   int64_t n_t, const double* vec_p, int64_t n_p, double* out, int64_t ld, double* mirror,
                   int64_t ld_mirror)
{
    if (!c) return CFMI_EINVAL;
    if (lmax < 2) {
        int rc = cfmi_fill(c, CFMI_F64, 2 * n_p, n_t, out, ld, 0.0);
        if (!rc && mirror) rc = cfmi_fill(c, CFMI_F64, n_t, 2 * n_p, mirror, ld_mirror, 0.0);
        return rc;
    }
    PairArgs a = {};
    a.vr = vec_p; a.vc = vec_t; a.nr = n_p; a.nc = n_t; a.L = lmax; a.lrun = lmax;
    a.blo = nullptr; a.bhi = nullptr; a.nbins = 1; a.sum_mode = 1;
    a.out[0] = out; a.mir[0] = mirror; a.ld = ld; a.ldm = ld_mirror; a.np = n_p;
    return launch_pairs(c, K02, a, cl_te, cl_tb, nullptr);
}

int64_t cfmi_deriv_state_doubles(int64_t n_r, int64_t n_c, int symmetric)
{
    if (n_r <= 0 || n_c <= 0) return 0;
    int64_t tr = (n_r + TILE - 1) / TILE, tc = (n_c + TILE - 1) / TILE;
    int64_t tiles = symmetric ? tr * (tr + 1) / 2 : tr * tc;
    return tiles * 256 * 2 * PPT;
}

int cfmi_deriv_stack_00(cfmi_ctx* c, int dtype, const double* w, int lmax, const int* bin_lo, const int* bin_hi,
                        int nbins, const double* vec_r, int64_t n_r, const double* vec_c, int64_t n_c, int symmetric,
                        void* out, int64_t ld, int64_t stride, void* mirror, int64_t ld_mirror,
                        int64_t stride_mirror, double* state, int lresume)
{
    if (!c) return CFMI_EINVAL;
    if (nbins <= 0) return CFMI_OK;
    if (symmetric && (n_r != n_c || mirror))
        return cfmi_fail(c, CFMI_EINVAL, "deriv_stack_00: symmetric block needs n_r == n_c and no mirror");
    PairArgs a = {};
    a.vr = vec_r; a.vc = vec_c; a.nr = n_r; a.nc = n_c; a.L = lmax; a.lrun = lmax;
    a.blo = bin_lo; a.bhi = bin_hi; a.nbins = nbins; a.sum_mode = 0;
    a.out[0] = out; a.mir[0] = mirror; a.ld = ld; a.stride = stride; a.ldm = ld_mirror; a.stride_m = stride_mirror;
    a.state = state; a.lresume = lresume;
    return launch_pairs(c, symmetric ? K00S : K00G, a, w, nullptr, nullptr, dtype);
}

int cfmi_deriv_stack_22(cfmi_ctx* c, int dtype, const double* w_ee, const double* w_bb, const double* w_eb,
                        int lmax, const int* bin_lo, const int* bin_hi, int nbins, const double* vec, int64_t n,
                        void* out_ee, void* out_bb, void* out_eb, int64_t ld, int64_t stride, double* state,
                        int lresume)
{
    if (!c) return CFMI_EINVAL;
    if (nbins <= 0) return CFMI_OK;
    PairArgs a = {};
    a.vr = vec; a.vc = vec; a.nr = n; a.nc = n; a.L = lmax; a.lrun = lmax;
    a.blo = bin_lo; a.bhi = bin_hi; a.nbins = nbins; a.sum_mode = 0;
    a.out[0] = out_ee; a.out[1] = out_bb; a.out[2] = out_eb; a.ld = ld; a.stride = stride; a.np = n;
    a.state = state; a.lresume = lresume;
    return launch_pairs(c, K22, a, w_ee, w_bb, w_eb, dtype);
}

int cfmi_deriv_stack_02(cfmi_ctx* c, int dtype, const double* w_te, const double* w_tb, int lmax,
                        const int* bin_lo, const int* bin_hi, int nbins, const double* vec_t, int64_t n_t,
                        const double* vec_p, int64_t n_p, void* out_te, void* out_tb, int64_t ld, int64_t stride,
                        void* mirror_te, void* mirror_tb, int64_t ld_mirror, int64_t stride_mirror, double* state,
                        int lresume)
{
    if (!c) return CFMI_EINVAL;
    if (nbins <= 0) return CFMI_OK;
    PairArgs a = {};
    a.vr = vec_p; a.vc = vec_t; a.nr = n_p; a.nc = n_t; a.L = lmax; a.lrun = lmax;
    a.blo = bin_lo; a.bhi = bin_hi; a.nbins = nbins; a.sum_mode = 0;
    a.out[0] = out_te; a.out[1] = out_tb; a.mir[0] = mirror_te; a.mir[1] = mirror_tb;
    a.ld = ld; a.stride = stride; a.ldm = ld_mirror; a.stride_m = stride_mirror; a.np = n_p;
    a.state = state; a.lresume = lresume;
    return launch_pairs(c, K02, a, w_te, w_tb, nullptr, dtype);
}

}  // extern "C"
