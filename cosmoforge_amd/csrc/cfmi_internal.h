// Internal declarations shared by the libcfmi translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/cfmi.h"

struct cfmi_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    // Recurrence coefficient table, l = 0..CFMI_MAX_ELL (see signal.hip).
    double* d_coef = nullptr;
    // Device scratch: LAPACK-style info flag and small staging buffers.
    int* d_info = nullptr;
    double* d_red = nullptr;       // reduction staging (CFMI_RED_SLOTS doubles)
    // Growable workspace for split-K partials and trtri temporaries.
    void* d_work = nullptr;
    size_t work_bytes = 0;
    // The blocked Cholesky's look-ahead: a second stream (default priority,
    // created on first use) with a workspace of its own, the two events that
    // order it against the context's stream, and the int8 slices of the
    // current panel, which both streams read.
    hipStream_t aux_stream = nullptr;
    void* aux_work = nullptr;
    size_t aux_work_bytes = 0;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    void* d_panel = nullptr;
    size_t panel_bytes = 0;
    // Sliced-GEMM launch record (cfmi_s8_stats_*): HIP events around every
    // gemm_s8 kernel launch on the stream while enabled, with each launch's
    // algorithmic flops and slice count.
    bool s8_stats = false;
    struct S8Launch {
        hipEvent_t start, stop;
        double flops;
        int s;
    };
    std::vector<S8Launch> s8_launches;
    std::vector<hipEvent_t> s8_event_pool;
};

#define CFMI_RED_SLOTS 65536

int cfmi_fail(cfmi_ctx* c, int code, const char* fmt, ...);
void* cfmi_workspace(cfmi_ctx* c, size_t bytes);

// C (m x w, FP64, column-major; only the entries on or below the diagonal)
// += alpha P P[0:w]^T, P m x k column-major (w <= m): P's rows as s int8
// slices (in the context workspace) on the int8 matrix cores, exact integer
// sums (gemm_s8.hip).  w = m is the lower triangle of P P^T.
int cfmi_s8_update_lower(cfmi_ctx* c, int s, int64_t m, int64_t w, int64_t k, double alpha, const double* P,
                         int64_t ldp, double* C, int64_t ldc);

#define CFMI_HIP(ctx, expr)                                                              \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return cfmi_fail((ctx), CFMI_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define CFMI_LAUNCH_CHECK(ctx)                                                           \
    do {                                                                                 \
        hipError_t e_ = hipGetLastError();                                               \
        if (e_ != hipSuccess)                                                            \
            return cfmi_fail((ctx), CFMI_EHIP, "kernel launch failed at %s:%d: %s",      \
                             __FILE__, __LINE__, hipGetErrorString(e_));                 \
    } while (0)

// GEMM entry shared by the Cholesky / trtri drivers (dense.hip).
// op(A) is M x K, op(B) is K x N, C is M x N, all column-major.
// tri: bit 0 = op(A) lower (A[m,k]=0 for k>m), bit 1 = op(A) upper,
//      bit 2 = op(B) upper (B[k,n]=0 for k>n), bit 3 = op(B) lower,
//      bit 4 = write only the lower triangle of C (m >= n).
//      bit 20 (CFMI_TRI_NOSPLIT) = never split K (no workspace; for graph capture).
constexpr int CFMI_TRI_NOSPLIT = 1 << 20;
int cfmi_gemm_impl(cfmi_ctx* c, int dtype, int ta, int tb, int64_t M, int64_t N, int64_t K,
                   double alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                   double beta, void* C, int64_t ldc, int tri);

// p[i] when ok, else zero, from an unconditional load at a clamped address
// (p[0] must be readable): a conditional load compiles to a branch that
// waits for its data, so a run of them is serialised, while the
// unconditional loads of a run are all in flight together.
template <typename T>
__device__ __forceinline__ T load_or_zero(const T* p, int64_t i, bool ok)
{
    const T v = p[ok ? i : 0];
    return ok ? v : T(0);
}

// XCD-grouped tile order of the batched GEMMs (gemm_b3_kernel, and the
// batched launches of gemm_kernel).  Workgroup x runs on XCD x % 8
// (round-robin dispatch); its slot x / 8 there picks group 8 (slot / G) +
// x % 8 of gm x gn tiles and tile slot % G in it (rows fastest).  Groups run
// over an odd number gtmp of group rows (a padding row exits at once), then
// group columns, then the batch, so consecutive groups of a trimmed product,
// whose work varies with the row, rotate over the XCDs, and the blocks an XCD
// holds at once share the column panels of B in its L2.  Every (batch, tile)
// pair is reached exactly once; slots beyond the grid of tiles return false.
// A-trimmed (tri & 1) / B-trimmed (tri & 4) products take rows / columns
// in reverse, heaviest first.
// G_ needs gm, gn, gtm, gtmp, gtn, ngroups, tm, tn and tri.
template <typename G_>
__device__ __forceinline__ bool group_tile(const G_& g, int64_t x, int64_t& z, int64_t& ti, int64_t& tj)
{
    const int64_t G = g.gm * g.gn;
    const int64_t slot = x >> 3;
    const int64_t gg = (slot / G) * 8 + (x & 7);
    if (gg >= g.ngroups) return false;
    const int64_t w = slot % G;
    const int64_t gi = gg % g.gtmp, rest = gg / g.gtmp;
    if (gi >= g.gtm) return false;
    z = rest / g.gtn;
    ti = gi * g.gm + w % g.gm;
    tj = (rest % g.gtn) * g.gn + w / g.gm;
    if (ti >= g.tm || tj >= g.tn) return false;
    if (g.tri & 1) ti = g.tm - 1 - ti;
    if (g.tri & 4) tj = g.tn - 1 - tj;
    return true;
}
