// One-wave Cholesky factor and inverse of a 64 x 64 diagonal block, shared by
// the blocked Cholesky (dense.hip) and the batched likelihood (batched.hip).
//
// Factor: lane r holds row r of the block in registers.  Step q reads the
// pivot from lane q (a scalar), forms 1 / L_qq once, scales column q by it,
// publishes the column through LDS and every lane updates its row
// (right-looking).  A single wave needs no barrier between its LDS store and
// load: the LDS unit processes one wave's operations in order.
//
// Inverse: lane c owns column c of L^-1 and solves L u = e_c right-looking,
// u_k = s_k / L_kk, then s_i -= L_ik u_k for i > k, so the dependent chain
// is one multiply and one FMA per k while the other updates are independent;
// L's columns are LDS broadcasts.
#pragma once

#include <hip/hip_runtime.h>

template <typename T>
__device__ __forceinline__ T lane_value(T v, int lane)
{
    if constexpr (sizeof(T) == 8) {
        const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
        const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
        return __hiloint2double(hi, lo);
    } else {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
    }
}

// A column-major 64 x 64 block (src, 4096 contiguous elements) into LDS as
// dst[i][j] = block[i][j], by NT threads: each thread issues its 4096 / NT
// loads together before its LDS stores (a load-store loop waits for every
// load in turn).
template <int NT, typename T, int LD>
__device__ __forceinline__ void block64_to_lds(T (*dst)[LD], const T* __restrict__ src, int tid)
{
    constexpr int PER = 4096 / NT;
    static_assert(PER * NT == 4096 && LD >= 64, "block64_to_lds: thread count must divide 4096");
    T v[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) v[e] = src[tid + e * NT];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int idx = tid + e * NT;
        dst[idx & 63][idx >> 6] = v[e];
    }
}

// row: lane r's row of the kb x kb block, zero above the diagonal and beyond
// kb (consumed: entry q is dead after step q, so the registers free up as
// the steps go).  On return lt[k][i] = L[i][k] (zero for k >= kb and above
// the diagonal) and rdiag[k] = 1 / L_kk (zero for k >= kb).  Returns 0, or
// the 1-based index of the first non-positive pivot (lt and rdiag then hold
// no factor).
//
// The steps have no branches (a branch per step costs register copies at
// every join, and keeps a step's updates from overlapping the next pivot's
// square root and reciprocal): a non-positive pivot is recorded and replaced
// by 1.  Past kb the rows are zero, so those steps publish the identity's
// columns, cleared at the end, and record their zero pivots past kb, which
// are not failures.  (A test of q against kb inside the steps would be
// computed for all 64 steps up front and kept in scalar registers.)
template <typename T>
__device__ __forceinline__ int wave_chol64(T (&row)[64], int kb, T (*lt)[65], T* rdiag)
{
    const int r = threadIdx.x & 63;
    int bad = 0;
#pragma clang loop unroll(full)
    for (int q = 0; q < 64; ++q) {
        const T d = lane_value(row[q], q);
        const bool pos = d > T(0);
        bad = (bad == 0 && !pos) ? q + 1 : bad;
        const T piv = sqrt(pos ? d : T(1));
        const T rp = T(1) / piv;
        const T l = r > q ? row[q] * rp : (r == q ? piv : T(0));
        lt[q][r] = l;
        rdiag[q] = rp;   // the same value from every lane
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // the column in groups of 16 reads, each group consumed (waited for)
        // as a whole before its updates: left alone, the compiler waits for
        // every pair of reads in turn
#pragma clang loop unroll(full)
        for (int g = 0; g < 4; ++g) {
            if (16 * g + 15 <= q) continue;
            T v[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = lt[q][16 * g + j];
#pragma unroll
            for (int j = 0; j < 16; ++j) asm volatile("" ::"v"(v[j]));
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (16 * g + j > q) row[16 * g + j] -= l * v[j];
        }
    }
    if (r >= kb) {
        lt[r][r] = T(0);
        rdiag[r] = T(0);
    }
    return bad > kb ? 0 : bad;
}

// From wave_chol64's lt and rdiag: U = L^-1, left in LDS as lt[c][i] = U[i][c]
// (so that lane i can then read row i of U, lt[.][i], for coalesced stores).
// u_k is final at step k and goes to lt[c][k] at once: the values that
// step k uses from row k of lt lie beyond column k, and those that step c
// uses from row c beyond column c, so no step uses an entry written this
// way (lane k's own write to lt[k][k] is read but not used).  (Kept in
// registers until the end instead, the 64 columns measured slower.)
template <typename T>
__device__ __forceinline__ void wave_inv64_to_lds(T (*lt)[65], const T* rdiag)
{
    __syncthreads();
    const int c = threadIdx.x & 63;
    T u[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) u[i] = (i == c) ? T(1) : T(0);
#pragma clang loop unroll(full)
    for (int k = 0; k < 64; ++k) {
        u[k] *= rdiag[k];
        lt[c][k] = u[k];
        // L's column k in groups of 16 reads, each consumed as a whole
#pragma clang loop unroll(full)
        for (int g = 0; g < 4; ++g) {
            if (16 * g + 15 <= k) continue;
            T v[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = lt[k][16 * g + j];
#pragma unroll
            for (int j = 0; j < 16; ++j) asm volatile("" ::"v"(v[j]));
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (16 * g + j > k) u[16 * g + j] -= v[j] * u[k];
        }
    }
    __syncthreads();
}

// U = L^-1 for one wave on the FP64 matrix cores, from wave_chol64's lt and
// rdiag, into ut as ut[c][i] = U[i][c] (the layout of wave_inv64_to_lds):
// the four 16 x 16 diagonal blocks by substitution (lane c takes column c %
// 16 of block c / 16, all four blocks at once), then U_BA = -U_BB L_BA U_AA
// for the two pairs of 16-blocks and once for the two 32-blocks, each
// product on v_mfma_f64_16x16x4 (lane (m, k) holds A[m][k] and B[k][m], and
// C[k + 4 r][m] after) with the operands read from LDS.  tb: 32 x 33
// staging for the intermediate products.  Entries past kb come out zero.
typedef double chol_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void wave_inv64_mfma(double (*lt)[65], const double* rdiag, double (*ut)[65],
                                                double (*tb)[33])
{
    __syncthreads();
    const int c = threadIdx.x & 63;
    const int lm = c & 15, lk = c >> 4;
    // zero U, then the 16 x 16 diagonal blocks by substitution
#pragma unroll
    for (int i = 0; i < 64; ++i) ut[c][i] = 0.0;
    {
        const int b = c >> 4, j = c & 15, o = 16 * b;
        double u[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) u[i] = (i == j) ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            u[k] *= rdiag[o + k];
#pragma unroll
            for (int i = k + 1; i < 16; ++i) u[i] -= lt[o + k][o + i] * u[k];
        }
        lds_fence();
#pragma unroll
        for (int i = 0; i < 16; ++i) ut[o + j][o + i] = u[i];
    }
    lds_fence();
    // pairs of 16-blocks (A, B) = (0, 1) and (2, 3): T = L_BA U_AA, U_BA = -U_BB T
#pragma unroll
    for (int pr = 0; pr < 2; ++pr) {
        const int a = 32 * pr, bb = a + 16;
        chol_d4 t = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int k = 4 * s + lk;
            t = __builtin_amdgcn_mfma_f64_16x16x4f64(lt[a + k][bb + lm], ut[a + lm][a + k], t, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) tb[lk + 4 * r][lm] = t[r];
        lds_fence();
        chol_d4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int k = 4 * s + lk;
            y = __builtin_amdgcn_mfma_f64_16x16x4f64(ut[bb + k][bb + lm], tb[k][lm], y, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) ut[a + lm][bb + lk + 4 * r] = -y[r];
        lds_fence();
    }
    // the 32-blocks: T = L21 U11 (32 x 32), U21 = -U22 T; U11 and U22 are
    // lower triangular, so their zero 16 x 16 corners skip half the K-steps
    chol_d4 t[2][2];
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) {
            t[ti][tj] = chol_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 4 * tj; s < 8; ++s) {
                const int k = 4 * s + lk;
                t[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(lt[k][32 + 16 * ti + lm], ut[16 * tj + lm][k],
                                                                 t[ti][tj], 0, 0, 0);
            }
        }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) tb[16 * ti + lk + 4 * r][16 * tj + lm] = t[ti][tj][r];
    lds_fence();
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) {
            chol_d4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < 4 * (ti + 1); ++s) {
                const int k = 4 * s + lk;
                y = __builtin_amdgcn_mfma_f64_16x16x4f64(ut[32 + k][32 + 16 * ti + lm], tb[k][16 * tj + lm], y, 0, 0,
                                                         0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) ut[16 * tj + lm][32 + 16 * ti + lk + 4 * r] = -y[r];
        }
    __syncthreads();
}
