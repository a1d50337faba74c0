// Internal declarations shared by the libcfmi translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

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
};

#define CFMI_RED_SLOTS 65536

int cfmi_fail(cfmi_ctx* c, int code, const char* fmt, ...);
void* cfmi_workspace(cfmi_ctx* c, size_t bytes);

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
int cfmi_gemm_impl(cfmi_ctx* c, int dtype, int ta, int tb, int64_t M, int64_t N, int64_t K,
                   double alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                   double beta, void* C, int64_t ldc, int tri);
