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
// the 1-based index of the first non-positive pivot (lt and rdiag are then
// incomplete).
template <typename T>
__device__ __forceinline__ int wave_chol64(T (&row)[64], int kb, T (*lt)[65], T* rdiag)
{
    const int r = threadIdx.x & 63;
    int bad = 0;
#pragma unroll
    for (int q = 0; q < 64; ++q) {
        if (q < kb && !bad) {
            const T d = lane_value(row[q], q);
            if (!(d > T(0))) {
                bad = q + 1;
            } else {
                const T piv = sqrt(d);
                const T rp = T(1) / piv;
                const T l = r > q ? row[q] * rp : (r == q ? piv : T(0));
                lt[q][r] = l;
                if (r == 0) rdiag[q] = rp;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
                for (int p = q + 1; p < 64; ++p) row[p] -= l * lt[q][p];
            }
        } else if (q >= kb) {
            lt[q][r] = T(0);
            if (r == 0) rdiag[q] = T(0);
        }
    }
    return bad;
}

// From wave_chol64's lt and rdiag: U = L^-1, left in LDS as lt[c][i] = U[i][c]
// (so that lane i can then read row i of U, lt[.][i], for coalesced stores).
// u_k is final at step k and goes to lt[c][k] at once: step k reads row k
// of lt beyond column k only, and row c is read at step c only beyond
// column c, so no later step reads an entry written this way.
template <typename T>
__device__ __forceinline__ void wave_inv64_to_lds(T (*lt)[65], const T* rdiag)
{
    __syncthreads();
    const int c = threadIdx.x & 63;
    T u[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) u[i] = (i == c) ? T(1) : T(0);
#pragma unroll
    for (int k = 0; k < 64; ++k) {
        u[k] *= rdiag[k];
#pragma unroll
        for (int i = k + 1; i < 64; ++i) u[i] -= lt[k][i] * u[k];
        lt[c][k] = u[k];
    }
    __syncthreads();
}
