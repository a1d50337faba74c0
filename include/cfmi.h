/*
 * cfmi.h -- C ABI of libcfmi, the MI355X (gfx950) pixel-space covariance
 * library behind cosmoforge_amd's Fisher / Spectra / PICSLike classes.
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer owned by the caller (e.g. a
 *     PyTorch-ROCm tensor's data_ptr()), unless its comment says "host".
 *   - Matrices are column-major (Fortran order, as the reference allocates
 *     them) with an explicit leading dimension.  A symmetric n x n matrix is
 *     stored in full.
 *   - Pointing vectors are (n, 3) row-major unit vectors (x, y, z).
 *   - C_l / weight arrays are l-indexed, length lmax + 1.
 *   - Return code: 0 = ok; < 0 = API / HIP error, message in
 *     cfmi_last_error(); > 0 = LAPACK-style info (first non-positive pivot,
 *     1-based) from a Cholesky factorisation.
 *   - All work is enqueued on the context's stream; a call returns when the
 *     work is enqueued, except calls that return a host scalar (they
 *     synchronise the stream).  A context is not thread-safe; separate
 *     contexts are.  No exception or abort crosses this boundary.
 *
 * Reference interfaces replaced (ggalloni/CosmoForge 1.3.0; paths relative
 * to src/cosmoforge.cosmocore/cosmocore/ unless stated):
 *   cfmi_pointings      compute_pointings / _unit_vectors   signal_kernels.py:113-161
 *   cfmi_signal_00      compute_00_contribution             signal_kernels.py:164-242
 *   cfmi_signal_22      compute_22_contribution             signal_kernels.py:245-333
 *   cfmi_signal_02      compute_02_contribution             signal_kernels.py:336-387
 *   cfmi_deriv_stack_*  get_binned_derivative_direct        basis/pixel.py:802-950
 *                       (every bin of one spectrum in one launch)
 *   cfmi_potrf          lapack.dpotrf in matrix_inverse_symm / cholesky_factor
 *                                                           basics/linalg.py:197-214,362-406
 *   cfmi_trtri          lapack.dpotri (via L^-1)            basics/linalg.py:212-215
 *   cfmi_trtri_blocks   lapack.dpotri on the diagonal blocks basics/linalg.py:212-215
 *   cfmi_jacobi_scale   D M D preconditioning               basics/linalg.py:197-207
 *   cfmi_logdet_chol    matrix_slogdet_symm                 basics/linalg.py:430-474
 *   cfmi_gemm           matrix_mult (np.matmul)             basics/linalg.py:18-32
 *   cfmi_gram           matrix_trace over all pairs         basics/linalg.py:66-98,
 *                                                           qube/qube/fisher.py:557-603
 *   cfmi_batched_*      PICSLike._compute_likelihood_point  picslike/picslike/picslike.py:568-664
 *   cfmi_mlp_*          (no reference counterpart: learned likelihood surrogate)
 */
#ifndef CFMI_H
#define CFMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFMI_OK 0
#define CFMI_EINVAL (-1)
#define CFMI_EHIP (-2)
#define CFMI_ENOMEM (-3)

#define CFMI_F64 0
#define CFMI_F32 1
#define CFMI_BF16 2
/* Three BF16 planes per value, x = x0 + x1 + x2 (24 significant bits). */
#define CFMI_BF16X3 3

/* Largest multipole the recurrence kernels accept. */
#define CFMI_MAX_ELL 4096

typedef struct cfmi_ctx cfmi_ctx;

int cfmi_version(void);

/* Create a context on `device`; `stream` is a hipStream_t (NULL: the
 * context creates its own). */
int cfmi_create(int device, void *stream, cfmi_ctx **out);
int cfmi_destroy(cfmi_ctx *ctx);
const char *cfmi_last_error(const cfmi_ctx *ctx);
int cfmi_synchronize(cfmi_ctx *ctx);

/* ---- A1: pixel centres -> unit vectors ------------------------------- */
/* vec[i] = unit vector of HEALPix pixel pix[i] (RING if nest == 0), on the
 * device.  Device sin/cos may differ from the platform libm by an ulp.     */
int cfmi_pointings(cfmi_ctx *ctx, int64_t nside, int nest, const int64_t *pix, int64_t n,
                   double *vec);
/* The same on the host with the platform libm (bit-compatible with the
 * reference's scalar loop); pix and vec are HOST arrays.  Spin-2 entries of
 * pixel pairs on a common meridian depend on the last bit of these vectors,
 * so the reference-compatible path uses this variant.                      */
int cfmi_pointings_host(int64_t nside, int nest, const int64_t *pix, int64_t n, double *vec);

/* ---- A2-A9: signal covariance blocks (FP64) ---------------------------- */
/* Spin-0 x spin-0.  cl: beam-smoothed C_l (device, lmax+1).
 * symmetric != 0: vec_r == vec_c (n_r == n_c), writes the full symmetric
 *   block S[j,i] = sum_l cl_l (2l+1)/(4pi) P_l(v_j . v_i) into out.
 * symmetric == 0: writes out[j + i*ld] for all j < n_r, i < n_c, and, if
 *   mirror != NULL, mirror[i + j*ld_mirror] with the same value.            */
int cfmi_signal_00(cfmi_ctx *ctx, const double *cl, int lmax, const double *vec_r, int64_t n_r,
                   const double *vec_c, int64_t n_c, int symmetric, double *out, int64_t ld,
                   double *mirror, int64_t ld_mirror);

/* Spin-2 auto block: the full (2n x 2n) [Q;U] x [Q;U] block into out. */
int cfmi_signal_22(cfmi_ctx *ctx, const double *cl_ee, const double *cl_bb, const double *cl_eb,
                   int lmax, const double *vec, int64_t n, double *out, int64_t ld);

/* Spin-0 x spin-2: the (2 n_p x n_t) [Q;U] x T block into out and, if
 * mirror != NULL, its transpose (n_t x 2 n_p) into mirror.                 */
int cfmi_signal_02(cfmi_ctx *ctx, const double *cl_te, const double *cl_tb, int lmax,
                   const double *vec_t, int64_t n_t, const double *vec_p, int64_t n_p,
                   double *out, int64_t ld, double *mirror, int64_t ld_mirror);

/* ---- A11: binned derivative stacks (FP64 recurrences) ----------------- */
/* Bin b covers l in [bin_lo[b], bin_hi[b]] (device int32 arrays, bins
 * ascending and disjoint); w is the l-indexed weight w_l * B1(l) B2(l).
 * Output b of a stack starts at out + b * stride, stored as dtype
 * (CFMI_F64; CFMI_F32: the FP64 sums rounded once; CFMI_BF16X3: split
 * into three BF16 planes, plane p of the stack at out + p * nbins * stride,
 * likewise for the mirror).  Geometry and mirror
 * conventions are those of the matching cfmi_signal_* call.  In the 22 and
 * 02 calls any output pointer may be NULL (not produced).
 * state (device, may be NULL) holds cfmi_deriv_state_doubles(n_r, n_c,
 * symmetric) doubles (22: n_r = n_c = n, symmetric; 02: n_r = n_p,
 * n_c = n_t, not symmetric): the per-pair recurrence state after the last
 * bin is saved there, and with lresume > 0 the recurrence starts at
 * l = lresume from the saved state (of a call that ended at lresume - 1
 * with the same geometry), so consecutive bin ranges cost one pass.       */
int64_t cfmi_deriv_state_doubles(int64_t n_r, int64_t n_c, int symmetric);
int cfmi_deriv_stack_00(cfmi_ctx *ctx, int dtype, const double *w, int lmax, const int *bin_lo,
                        const int *bin_hi, int nbins, const double *vec_r, int64_t n_r,
                        const double *vec_c, int64_t n_c, int symmetric, void *out, int64_t ld,
                        int64_t stride, void *mirror, int64_t ld_mirror, int64_t stride_mirror,
                        double *state, int lresume);
int cfmi_deriv_stack_22(cfmi_ctx *ctx, int dtype, const double *w_ee, const double *w_bb,
                        const double *w_eb, int lmax, const int *bin_lo, const int *bin_hi,
                        int nbins, const double *vec, int64_t n, void *out_ee, void *out_bb,
                        void *out_eb, int64_t ld, int64_t stride, double *state, int lresume);
int cfmi_deriv_stack_02(cfmi_ctx *ctx, int dtype, const double *w_te, const double *w_tb, int lmax,
                        const int *bin_lo, const int *bin_hi, int nbins, const double *vec_t,
                        int64_t n_t, const double *vec_p, int64_t n_p, void *out_te,
                        void *out_tb, int64_t ld, int64_t stride, void *mirror_te,
                        void *mirror_tb, int64_t ld_mirror, int64_t stride_mirror, double *state,
                        int lresume);

/* ---- A12-A15: dense linear algebra ------------------------------------- */
/* C = alpha op(A) op(B) + beta C; dtype CFMI_F64 or CFMI_F32; trans* != 0
 * selects the transpose.  flags: bit 0 op(A) lower-triangular, bit 1 op(A)
 * upper, bit 2 op(B) upper, bit 3 op(B) lower (zero parts are skipped, they
 * must hold zeros), bit 4 write only the lower triangle of C, bit 6 (64)
 * the product is symmetric (m == n): only its lower tiles are computed and
 * every batch's upper triangle is mirrored from them.                     */
int cfmi_gemm(cfmi_ctx *ctx, int dtype, int transa, int transb, int64_t m, int64_t n, int64_t k,
              double alpha, const void *A, int64_t lda, const void *B, int64_t ldb, double beta,
              void *C, int64_t ldc, int flags);

/* Same product, written into the packed lower triangle of a pn x pn
 * matrix (column-major packing: element (i, j), i >= j, at
 * j (2 pn - j + 1) / 2 + i - j).  C's element (r, c) is (row0 + r,
 * col0 + c); only entries with row0 + r >= col0 + c are produced, scaled by
 * pdiag on the diagonal and poff elsewhere before beta-accumulation.      */
int cfmi_gemm_packed(cfmi_ctx *ctx, int dtype, int transa, int transb, int64_t m, int64_t n,
                     int64_t k, double alpha, const void *A, int64_t lda, const void *B,
                     int64_t ldb, double beta, void *P, int64_t pn, int64_t row0, int64_t col0,
                     double pdiag, double poff, int flags);

/* Strided batches of the two products above: batch b uses A + b sA,
 * B + b sB and C + b sC (P + b sP); batch <= 65535.                       */
int cfmi_gemm_batched(cfmi_ctx *ctx, int dtype, int transa, int transb, int64_t m, int64_t n,
                      int64_t k, double alpha, const void *A, int64_t lda, int64_t sA, const void *B,
                      int64_t ldb, int64_t sB, double beta, void *C, int64_t ldc, int64_t sC,
                      int64_t batch, int flags);
int cfmi_gemm_packed_batched(cfmi_ctx *ctx, int dtype, int transa, int transb, int64_t m, int64_t n,
                             int64_t k, double alpha, const void *A, int64_t lda, int64_t sA,
                             const void *B, int64_t ldb, int64_t sB, double beta, void *P,
                             int64_t sP, int64_t pn, int64_t row0, int64_t col0, double pdiag,
                             double poff, int64_t batch, int flags);

/* General form of the products above.
 *   dtype / out_dtype: (F64, F64), (F32, F32) or (F32, F64).
 *   flags: bits 0-3 as cfmi_gemm, with the triangular trims shifted by a
 *     diagonal offset: bit 0 skips k > m + off_a, bit 1 k < m + off_a,
 *     bit 2 k > n + off_b, bit 3 k < n + off_b;
 *     CFMI_GEMM_ACC64: FP32 products accumulated by the MFMA in FP32 over
 *     chunks of 256 terms and in FP64 across chunks (always on for
 *     out_dtype F64 with FP32 inputs);
 *     bit 10 (1024): batches take their tiles in the plain order over the
 *     batch grid dimension; by default a batched product (batch > 1) runs
 *     its tiles in groups of 1 x 16 per XCD
 *     over a one-dimensional grid; bits 12-15 / 16-19: v > 0 gives
 *     2^(v - 1) group rows / columns instead.
 *   out_mode 0: dense C (ldc); 1: packed lower triangle of an ldc x ldc
 *     matrix as cfmi_gemm_packed; 2: packed lower rows, row-major: element
 *     (i, j <= i) at C + i (i + 1) / 2 + j - pbase.  In modes 1 and 2,
 *     (r, c) of the product is element (row0 + r, col0 + c), only its
 *     lower-triangle entries are produced, scaled by pdiag / poff.
 *   Batches as cfmi_gemm_batched (sC strides the output).                 */
#define CFMI_GEMM_ACC64 128
int cfmi_gemm_ex(cfmi_ctx *ctx, int dtype, int out_dtype, int transa, int transb, int64_t m,
                 int64_t n, int64_t k, double alpha, const void *A, int64_t lda, int64_t sA,
                 int64_t off_a, const void *B, int64_t ldb, int64_t sB, int64_t off_b,
                 double beta, void *C, int64_t ldc, int64_t sC, int64_t batch, int flags,
                 int out_mode, int64_t row0, int64_t col0, int64_t pbase, double pdiag,
                 double poff);

/* FP32-accurate product on BF16 MFMA (CFMI_BF16X3 operands): C = alpha op
 * with op(A)[m][k] = A[m lda + k] and op(B)[k][n] = B[n ldb + k] (planes pA,
 * pB elements apart; batches sA, sB, sC apart), six piece products per
 * term accumulated in FP32.  flags bit 0: A[m][k] = 0 for k > m + off_a;
 * bit 2: B[n][k] = 0 for k > n + off_b (the operands must hold those
 * zeros; the bits only shorten K); bits 8 / 9 force 128 x 128 / 256 x 256
 * block tiles (default: 256 with at least 512 such tiles for products
 * without a B-trim outside out_mode 2, and for N, K >= 6144); bit 10
 * takes the tiles column by column instead of in groups of 2 x 8 per XCD
 * (bits 12-15 / 16-19: v > 0 gives 2^(v - 1) group rows / columns).
 * out_mode 0: FP32 C[m + n ldc] (+ beta C); 1: three BF16 planes
 * C[m ldc + n] (planes pC apart); 2: FP32 packed lower row band as
 * cfmi_gemm_ex mode 2 (+ beta C).  Replaces the FP32 products of the
 * whitening in fisher.py:506-603 (no reference counterpart of the
 * arithmetic itself).                                                     */
int cfmi_gemm_b3(cfmi_ctx *ctx, int out_mode, int64_t m, int64_t n, int64_t k, double alpha,
                 const void *A, int64_t lda, int64_t sA, int64_t pA, int64_t off_a, const void *B,
                 int64_t ldb, int64_t sB, int64_t pB, int64_t off_b, double beta, void *C, int64_t ldc,
                 int64_t sC, int64_t pC, int64_t batch, int flags, int64_t row0, int64_t col0,
                 int64_t pbase, double pdiag);
/* Column-major FP64/FP32 (rows x cols) -> three BF16 planes, row-major
 * (element (i, j) at dst + i ldd + j), planes ps apart; and back (sum of the
 * pieces, count values, FP64).                                            */
int cfmi_split3(cfmi_ctx *ctx, int dtype, int64_t rows, int64_t cols, const void *src, int64_t lds,
                void *dst, int64_t ldd, int64_t ps);
int cfmi_join3(cfmi_ctx *ctx, int64_t count, const void *src, int64_t ps, double *dst);

/* ---- FP64 products on the int8 matrix cores (sliced operands) ---------- */
/* An operand row x_r (k values) is held as s int8 slices and a scale
 * mu_r = max |x_rk|: x = mu (s_1/127 + s_2/(127*254) + ...), each slice a
 * plane of rows x ldk bytes (k contiguous, zero beyond k), planes ps bytes
 * apart; s = 3, 4, 5 keep the row to 2^-24, 2^-32, 2^-40 of mu.
 * cfmi_slice8 layout 0: row r of the source at src + r lds (k contiguous);
 * layout 1: element (r, k) at src[r + k lds] (rows of a column-major
 * matrix).  ldk a multiple of 64 >= k; dst 16-byte aligned.                */
int cfmi_slice8(cfmi_ctx *ctx, int s, int layout, int64_t rows, int64_t k, const double *src, int64_t lds,
                void *dst, int64_t ldk, int64_t ps, double *mu);
/* The FP64 values the slices represent, dst[r ldd + k] (tests).            */
int cfmi_unslice8(cfmi_ctx *ctx, int s, int64_t rows, int64_t k, const void *src, int64_t ldk, int64_t ps,
                  const double *mu, double *dst, int64_t ldd);
/* dst (n x m) = src^T for a column-major FP64 m x n src (tiled copy).     */
int cfmi_transpose(cfmi_ctx *ctx, int64_t m, int64_t n, const double *src, int64_t lds, double *dst,
                   int64_t ldd);
/* Largest k one launch accumulates exactly in int32 (s k 127^2 < 2^31).    */
int64_t cfmi_s8_kmax(int s);
/* C (m x n, column-major FP64) = alpha A B^T + beta C with A (m rows) and
 * B (n rows) sliced over the same k: the s(s+1)/2 slice products with
 * p + q < s on v_mfma_i32_16x16x64_i8 (exact int32 sums), levels combined
 * in FP64; k beyond cfmi_s8_kmax is split and the parts added in FP64.
 * flags bit 0: symmetric result (m == n), lower tiles computed and mirrored;
 * bit 5: only the entries on or below the diagonal (n <= m: the lower
 * triangle, or a lower trapezoid; the strict upper part of C is left as
 * it was);
 * bit 1: one LDS stage staged through registers instead of two filled by
 * asynchronous global -> LDS copies;
 * bit 2: A[m][k] = 0 for k > m + off_a, bit 3: B[n][k] = 0 for k > n + off_b
 * (the operands hold those zeros; the bits only shorten K per tile);
 * bit 4: C is the packed lower row band of the whitening (element (i, j <= i)
 * at i (i + 1) / 2 + j - pbase for i = row0 + m, j = col0 + n, the diagonal
 * scaled by pdiag; ldc unused).
 * Serves the C^-1 V, V^T C^-1 V and W dC_b W^T products that replace
 * matrix_mult at qube/qube/fisher.py:506-531 and spectra.py:1144-1161 in
 * the mixed precision (no reference counterpart of the arithmetic).       */
int cfmi_gemm_s8(cfmi_ctx *ctx, int s, int64_t m, int64_t n, int64_t k, double alpha, const void *A, int64_t lda,
                 int64_t psA, const double *muA, int64_t off_a, const void *B, int64_t ldb, int64_t psB,
                 const double *muB, int64_t off_b, double beta, double *C, int64_t ldc, int flags, int64_t row0,
                 int64_t col0, int64_t pbase, double pdiag);

/* d[i] = 1/sqrt(|A_ii|) (1 where A_ii == 0), then A <- D A D (FP64). */
int cfmi_jacobi_scale(cfmi_ctx *ctx, int64_t n, double *A, int64_t lda, double *d);

/* Lower Cholesky A = L L^T in place (upper triangle zeroed).  If U != NULL
 * the diagonal 64 x 64 blocks of L^-1 are written into U (rest zeroed), to
 * be completed by cfmi_trtri.  *info (host) receives the LAPACK info.
 * For n > 4096 the trailing updates past the next 2048-column panel run on
 * a second stream that the context owns (look-ahead), joined to the
 * context's stream before the call synchronizes that stream and returns. */
int cfmi_potrf(cfmi_ctx *ctx, int dtype, int64_t n, void *A, int64_t lda, void *U, int64_t ldu,
               int *info);

/* Launch record of the sliced GEMM on this context (no reference
 * counterpart; measurement): while enabled, HIP events on the launch's
 * stream bracket every gemm_s8 kernel launch, including those the library
 * issues itself (cfmi_potrf_s8, some of them on its look-ahead stream).  enable(on) synchronizes and clears the
 * record; read synchronizes and returns the launch count, the summed
 * launch time (ms), the algorithmic flops (2 k per entry computed: 2 m n k,
 * k (m n - n (n - 1) / 2) for a lower or symmetric result, i.e. about m n k
 * for m = n) and the int8 products (flops s (s + 1) / 2).                 */
int cfmi_s8_stats_enable(cfmi_ctx *ctx, int on);
int cfmi_s8_stats_read(cfmi_ctx *ctx, int64_t *launches, double *ms, double *flops, double *products);

/* cfmi_potrf (FP64) with the rank-2048 trailing updates, nearly all of
 * the n^3/3 flops, on the int8 matrix cores: each panel's rows as s int8
 * slices (3 <= s <= 5; each entry to 2^-(8 s) of its row's largest),
 * exact int32 sums combined in FP64.  No reference counterpart: the
 * mixed-precision factorisation of the covariance for the harmonic route. */
int cfmi_potrf_s8(cfmi_ctx *ctx, int s, int64_t n, double *A, int64_t lda, double *U, int64_t ldu,
                  int *info);

/* Complete U = L^-1 from cfmi_potrf's L and diagonal blocks of U.        */
int cfmi_trtri(cfmi_ctx *ctx, int dtype, int64_t n, const void *L, int64_t ldl, void *U,
               int64_t ldu);

/* The inverses of L's diagonal nb x nb blocks (nb = 64 * 2^k; the last
 * block may be short) into the same places of U, from cfmi_potrf's L and
 * 64 x 64 blocks of U: the blocked triangular solves' diagonal products
 * in the harmonic route, the dpotri step (basics/linalg.py:212-215)
 * restricted to the blocks that the solves need.                           */
int cfmi_trtri_blocks(cfmi_ctx *ctx, int dtype, int64_t n, int64_t nb, const void *L, int64_t ldl, void *U,
                      int64_t ldu);

/* log det of D^-1 L L^T D^-1 = 2 sum log L_ii - 2 sum log d_i (d may be
 * NULL); the result is written to *out (host).                             */
int cfmi_logdet_chol(cfmi_ctx *ctx, int64_t n, const double *L, int64_t ldl, const double *d,
                     double *out);

/* A[:, j] *= d[j] (side 0) or A[i, :] *= d[i] (side 1); FP64.           */
int cfmi_scale(cfmi_ctx *ctx, int side, int64_t m, int64_t n, double *A, int64_t lda,
               const double *d);

/* B = alpha A + beta B elementwise on an m x n block; FP64.             */
int cfmi_axpby(cfmi_ctx *ctx, int64_t m, int64_t n, double alpha, const double *A, int64_t lda,
               double beta, double *B, int64_t ldb);

/* Convert between FP64 and FP32 matrices (dir 0: f64 -> f32, 1: f32 -> f64). */
int cfmi_convert(cfmi_ctx *ctx, int dir, int64_t m, int64_t n, const void *src, int64_t lds,
                 void *dst, int64_t ldd);

/* Copy an m x n block (FP64 or FP32 per dtype) between leading dimensions. */
int cfmi_copy(cfmi_ctx *ctx, int dtype, int64_t m, int64_t n, const void *src, int64_t lds, void *dst,
              int64_t ldd);

/* Fill an m x n block with value v (FP64 or FP32 per dtype). */
int cfmi_fill(cfmi_ctx *ctx, int dtype, int64_t m, int64_t n, void *A, int64_t lda, double v);

/* G[i, j] = sum_k X[k, i] Y[k, j] for i < m, j < n over k < K (FP64 or
 * FP32 inputs, FP64 output), with fixed-order split-K partial sums; FP32
 * inputs accumulate as in CFMI_GEMM_ACC64.  X and Y are K x m and K x n
 * column-major (ldx, ldy >= K).                                           */
int cfmi_gram(cfmi_ctx *ctx, int dtype, int64_t m, int64_t n, int64_t K, const void *X,
              int64_t ldx, const void *Y, int64_t ldy, double *G, int64_t ldg);

/* R[a + b ldr] = sum_k A[k + a lda] B[k + b ldb] for a < na, b < nb, k < K
 * (FP64; lda, ldb >= K when na, nb > 1): an HBM-bound pass with fixed-order
 * partial sums.  Replaces the contractions of the QML estimate and noise
 * bias, 1/2 y^T dC_b y and 1/2 Tr[C^-1 N C^-1 dC_b] (spectra.py:1124-1186),
 * as <D_b, M_q> over the flattened derivative blocks.                     */
int cfmi_dot_batched(cfmi_ctx *ctx, int64_t K, int64_t na, const double *A, int64_t lda, int64_t nb,
                     const double *B, int64_t ldb, double *R, int64_t ldr);

/* *out = number of nonzero off-diagonal entries of an n x n FP64 matrix
 * (n <= 65535).  Lets the QML's G = C^-1 N C^-1 use the structure of a
 * diagonal N (the common white-noise case of spectra.py:1124-1186).      */
int cfmi_offdiag_nnz(cfmi_ctx *ctx, int64_t n, const double *A, int64_t lda, int64_t *out);

/* ---- A18 / A21: batched exact likelihood -------------------------------- */
/* A batch of njobs covariance matrices C_j = sum_k coef[j][k] M_{p(j),k},
 * p(j) = job_patch[j] (host array), from per-patch FP64 component matrices
 * (device; patch p's k-th n_p x n_p column-major matrix at
 * comps + comp_off[p] + k n_p^2, comp_off host) and data vectors (device,
 * patch p's at xs + x_off[p], x_off host); n_p <= 2048, ncomp <= 4.
 * Per job: chi2 = x^T C^-1 x and ln|C| of the Jacobi-scaled Cholesky.
 *   mode CFMI_BATCH_MIXED: factor stored in FP32, trailing updates in FP32
 *     on MFMA; diagonal blocks, panels, ln|C| and refinement in FP64;
 *   mode CFMI_BATCH_MIXED_F64_STORAGE: factor stored in FP64, trailing
 *     products in FP32 on MFMA added to it in FP64;
 *   mode CFMI_BATCH_FP64: everything FP64.
 *   nb: panel width 16, 32 or 64; wg_per_cu: persistent workgroups per CU;
 *   refine_iters: refinement steps y += C^-1 (x - C y) (0..8).
 * info[j] > 0 (chi2, logdet NaN) marks a non-positive-definite C_j.
 * cfmi_batch_run enqueues on the context's stream; cfmi_batch_capture
 * records the same work as a graph with fixed argument pointers, which
 * cfmi_batch_replay launches (refill coef in place between replays).     */
#define CFMI_BATCH_MIXED 1
#define CFMI_BATCH_MIXED_F64_STORAGE 2
#define CFMI_BATCH_FP64 3
typedef struct cfmi_batch cfmi_batch;
int cfmi_batch_create(cfmi_ctx *ctx, int mode, int npatch, const int64_t *n, int ncomp,
                      const double *comps, const int64_t *comp_off, const double *xs,
                      const int64_t *x_off, int njobs, const int *job_patch, int nb, int wg_per_cu,
                      int refine_iters, cfmi_batch **out);
int cfmi_batch_run(cfmi_batch *b, const double *coef, double *chi2, double *logdet, int *info);
int cfmi_batch_capture(cfmi_batch *b, const double *coef, double *chi2, double *logdet, int *info);
int cfmi_batch_replay(cfmi_batch *b);
int cfmi_batch_destroy(cfmi_batch *b);

/* ---- harmonic-mode contractions (no reference counterpart in the pixel
 * path; the algebra of src/cosmoforge.qube/qube/fisher.py:416-500 and
 * cosmocore/basis/harmonic_basis.py:19-117) ------------------------------
 * Modes are l-major over l in [lmin, lmax]: m = 0, then the cos and sin
 * modes of m = 1..l; (lmax+1)^2 - lmin^2 modes per group. */

/* Real harmonic modes at the pixel centres vec (npix x 3, row-major unit
 * vectors) into the column-major V (ldv rows, ncols columns; the other
 * entries are left as they are).  spin 0: rows row_a.., columns col_a +
 * mode.  spin 2 (lmin >= 2): E columns col_a + mode and B columns col_b +
 * mode over the Q rows row_a.. and U rows row_b.., signed so that
 * sum_m x_lm y_lm^T is the pixel-space dS/dC_l^XY of the T, E, B blocks. */
int cfmi_harmonic_modes(cfmi_ctx *ctx, int spin, int64_t npix, const double *vec, int lmin, int lmax,
                        double *V, int64_t ldv, int64_t ncols, int64_t row_a, int64_t row_b,
                        int64_t col_a, int64_t col_b);

/* F[p][q] (npar x npar, npar = nspec nbins, parameter p = spectrum p/nbins,
 * bin p%nbins) = 1/2 Tr[C^-1 dC_p C^-1 dC_q] from the symmetric
 * M = V^T C^-1 V (ldm); with M2 = V^T C2^-1 V (non-NULL) the cross-analysis
 * 1/2 Tr[C1^-1 dC_p C2^-1 dC_q].  dC_p = sum_{l in bin} w[s][l] sum_m sum over
 * the spectrum's (u, v) group pairs of u_lm v_lm^T.  Device tables: bin_lo,
 * bin_hi (nbins), pairs (nspec x 4 group ids u0 v0 u1 v1, u1 = -1 for an
 * auto spectrum), goff (first column of each group in M), w (nspec x lw).
 * The caller keeps every index inside M and w.                           */
int cfmi_harmonic_fisher(cfmi_ctx *ctx, const double *M, const double *M2, int64_t ldm, int lmin, int nbins,
                         const int *bin_lo, const int *bin_hi, int nspec, const int *pairs,
                         const int64_t *goff, const double *w, int lw, double *F);

/* out[p + j ldo] = alpha sum_{l in bin(p)} w[s][l] sum_m sum_{(u,v)}
 * sum_{i < ni} X[i + (goff[u] + k) ldx + j sxj] Y[i + (goff[v] + k) ldy + j syj]
 * (k the mode index), for j < ncols: the QML forms u^T y v^T y (ni = 1)
 * and u^T G v (ni = n).                                                 */
int cfmi_harmonic_pairsum(cfmi_ctx *ctx, int64_t ni, const double *X, int64_t ldx, int64_t sxj,
                          const double *Y, int64_t ldy, int64_t syj, int64_t ncols, int lmin, int nbins,
                          const int *bin_lo, const int *bin_hi, int nspec, const int *pairs,
                          const int64_t *goff, const double *w, int lw, double alpha, double *out,
                          int64_t ldo);

/* The noise-bias form of cfmi_harmonic_pairsum over whole columns (ni = n),
 * with optional row weights: out[p] = alpha sum_{l in bin(p)} w[s][l] sum_m
 * sum_{(u,v)} sum_i wrow_i X[i + (goff[u] + k) ldx] Y[i + (goff[v] + k) ldy]
 * (wrow NULL: 1).  Each mode's columns of every group (ngroups <= 3, nmodes
 * modes each) are read once for all group pairs; X == Y with wrow = diag N
 * gives 1/2 Tr[C^-1 N C^-1 dC_p] from R = C^-1 V without forming N R.
 * Replaces spectra.py:1124-1186 (the noise bias, harmonic form).         */
int cfmi_harmonic_coldots(cfmi_ctx *ctx, int64_t ni, const double *X, int64_t ldx, const double *Y,
                          int64_t ldy, const double *wrow, int ngroups, int64_t nmodes, int lmin, int nbins,
                          const int *bin_lo, const int *bin_hi, int nspec, const int *pairs,
                          const int64_t *goff, const double *w, int lw, double alpha, double *out);

/* ---- A21: ensemble MLP surrogate of -2 ln L ---------------------------- */
/* E networks 3 -> 256 -> 256 -> 256 -> 1 (SiLU).  Parameters per member
 * (cfmi_mlp_param_count() values, layout W1[256][3], b1, W2[256][256], b2,
 * W3[256][256], b3, W4[1][256], b4): FP32 master w32 [E][P] and its BF16
 * copy wb16 [E][P] (uint16 bit patterns), the copy used by every product;
 * the 256 x 256 layers run on BF16 MFMA with FP32 accumulation.
 * X: inputs [N][3] FP32; idx: [E][B] rows of X per member (NULL: row b);
 * y: [E][B]; ws: cfmi_mlp_workspace_bytes(E, B) bytes of activations kept
 * by forward for backward.  Backward writes dL/dparams into grad [E][P]
 * from dy = dL/dy [E][B]; weight gradients reduce over batch chunks of
 * `chunk` rows (multiple of 32) through part [E][ceil(B/chunk)][256][256].
 * cfmi_mlp_loss: dy = 2 (y - T[idx]) / B and loss[e] = mean squared error. */
int64_t cfmi_mlp_param_count(void);
int64_t cfmi_mlp_workspace_bytes(int E, int B);
int cfmi_mlp_forward(cfmi_ctx *ctx, int E, int B, const float *X, const int *idx, const float *w32,
                     const void *wb16, float *y, void *ws);
int cfmi_mlp_loss(cfmi_ctx *ctx, int E, int B, const float *y, const float *T, const int *idx, float *dy,
                  float *loss);
int cfmi_mlp_backward(cfmi_ctx *ctx, int E, int B, const float *X, const int *idx, const void *wb16,
                      const float *dy, void *ws, float *grad, float *part, int chunk);
/* cfmi_mlp_backward without the workspace argument: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, training loops that forbid hipMalloc).  The reduction
 * partials live in a static 32 MiB device array of the library (one per
 * GPU, kept until the process ends), so E * ceil(B / chunk) <= 128; larger
 * sizes return CFMI_EINVAL.  Allocates nothing.  Calls that overlap on one
 * GPU share the array and race.  Same results as cfmi_mlp_backward.        */
int cfmi_mlp_backward_noalloc(cfmi_ctx *ctx, int E, int B, const float *X, const int *idx, const void *wb16,
                              const float *dy, void *ws, float *grad, int chunk);
/* Adam (bias-corrected, step >= 1) on total = E * P values; refreshes wb16. */
int cfmi_mlp_adam(cfmi_ctx *ctx, int64_t total, float *w, const float *g, float *m, float *v, double lr,
                  double beta1, double beta2, double eps, int step, void *wb16);
int cfmi_mlp_to_bf16(cfmi_ctx *ctx, int64_t total, const float *w, void *wb16);
/* mean and population standard deviation of y [E][B] over the members. */
int cfmi_mlp_ensemble_stats(cfmi_ctx *ctx, int E, int B, const float *y, float *mean, float *sd);

#ifdef __cplusplus
}
#endif
#endif /* CFMI_H */
