/*
 * FP64 CPU oracle: pixel-pair signal kernels (test infrastructure only).
 *
 * A plain-C restatement of the reference's Numba kernels, kept loop for loop
 * so that the rounding sequence matches the reference's scalar code:
 *
 *   legendre_00 / legendre_22 / legendre_02
 *       src/cosmoforge.cosmocore/cosmocore/basics/legendre.py:10-149
 *   get_rotation_angle / _project_and_norm
 *       src/cosmoforge.cosmocore/cosmocore/basics/geometry.py:14-110
 *   compute_00_contribution / _compute_00_symmetric / _compute_00_general
 *       src/cosmoforge.cosmocore/cosmocore/signal_kernels.py:164-242
 *   compute_22_contribution / _compute_22_symmetric   :245-333
 *   compute_02_contribution / _compute_02_parallel    :336-387
 *
 * Matrices are column-major (Fortran order, as the reference allocates them)
 * with leading dimension ld; "S[j, i]" of the reference is S[j + i*ld].
 * The outer pixel loop is an OpenMP loop, as the reference's prange.
 * Build with -ffp-contract=off: the reference's LLVM code does not fuse.
 */
#include <math.h>
#include <stdlib.h>
#include <stdint.h>

#define PI_D 3.14159265358979311599796346854

static void legendre_00(double x, double *leg, int n_ell)
{
    int lmax = n_ell - 1;
    leg[0] = 1.0;
    if (lmax >= 1) leg[1] = x;
    if (lmax >= 2) leg[2] = 1.5 * x * x - 0.5;
    for (int ell = 3; ell <= lmax; ++ell)
        leg[ell] = ((2 * ell - 1) * x * leg[ell - 1] - (ell - 1) * leg[ell - 2]) / ell;
    for (int ell = 0; ell <= lmax; ++ell)
        leg[ell] *= (2 * ell + 1) / (4 * PI_D);
}

static void legendre_22(double x, double *leg, double *f1, double *f2, int n_ell)
{
    int lmax = n_ell - 1;
    for (int i = 0; i < n_ell; ++i) { leg[i] = 0.0; f1[i] = 0.0; f2[i] = 0.0; }
    if (lmax >= 2) {
        leg[2] = 3.0;
        f1[2] = 6.0 * (1.0 + x * x);
        f2[2] = -12.0 * x;
    }
    for (int ell = 3; ell <= lmax; ++ell) {
        leg[ell] = (x * (2 * ell - 1) * leg[ell - 1] - (ell + 1) * leg[ell - 2]) / (ell - 2);
        f1[ell] = -(2 * ell - 8 + ell * (ell - 1) * (1.0 - x * x)) * leg[ell]
                  + (2 * ell + 4) * x * leg[ell - 1];
        f2[ell] = 4.0 * (-(ell - 1) * x * leg[ell] + (ell + 2) * leg[ell - 1]);
    }
    for (int ell = 2; ell <= lmax; ++ell) {
        double factor2 = 1.0 / ((double)((ell + 2) * (ell + 1) * ell * (ell - 1)));
        double norm = (2 * ell + 1) / (4 * PI_D) * factor2;
        leg[ell] *= norm;
        f1[ell] *= norm;
        f2[ell] *= norm;
    }
}

static void legendre_02(double x, double *leg, int n_ell)
{
    int lmax = n_ell - 1;
    for (int i = 0; i < n_ell; ++i) leg[i] = 0.0;
    if (lmax >= 2) leg[2] = 3.0 * (1.0 - x * x);
    for (int ell = 3; ell <= lmax; ++ell)
        leg[ell] = (x * (2 * ell - 1) * leg[ell - 1] - (ell + 1) * leg[ell - 2]) / (ell - 2);
    for (int ell = 2; ell <= lmax; ++ell) {
        double factor = sqrt(1.0 / ((double)((ell + 2) * (ell + 1) * ell * (ell - 1))));
        leg[ell] *= (2 * ell + 1) / (4 * PI_D) * factor;
    }
}

static const double EPS_BUMP = PI_D / 180.0 / 3600.0 / 100.0;

static void project_and_norm(double vx, double vy, double *px_o, double *py_o)
{
    double px = 0.0 * 0.0 - 1.0 * vy;   /* zz=(0,0,1): px = zzy*vz - zzz*vy */
    double py = 1.0 * vx - 0.0 * 0.0;   /* py = zzz*vx - zzx*vz */
    double nm = sqrt(px * px + py * py);
    if (nm < 1e-8) {
        px += EPS_BUMP;
        nm = sqrt(px * px + py * py);
    }
    *px_o = px / nm;
    *py_o = py / nm;
}

void oracle_rotation_angle(const double *r1, const double *r2, double *a12, double *a21)
{
    double r12x = r1[1] * r2[2] - r1[2] * r2[1];
    double r12y = r1[2] * r2[0] - r1[0] * r2[2];
    double r12z = r1[0] * r2[1] - r1[1] * r2[0];
    double mod = sqrt(r12x * r12x + r12y * r12y + r12z * r12z);
    if (mod < 1e-8) { *a12 = 0.0; *a21 = 0.0; return; }
    r12x /= mod; r12y /= mod; r12z /= mod;
    double r1sx, r1sy, r2sx, r2sy;
    project_and_norm(r1[0], r1[1], &r1sx, &r1sy);
    project_and_norm(r2[0], r2[1], &r2sx, &r2sy);
    double dot1 = r12x * r1sx + r12y * r1sy;
    if (dot1 > 1.0) dot1 = 1.0; else if (dot1 < -1.0) dot1 = -1.0;
    double sign1 = r12z > 0.0 ? 1.0 : -1.0;
    *a12 = 2.0 * acos(dot1) * sign1;
    r12x = -r12x; r12y = -r12y; r12z = -r12z;
    double dot2 = r12x * r2sx + r12y * r2sy;
    if (dot2 > 1.0) dot2 = 1.0; else if (dot2 < -1.0) dot2 = -1.0;
    double sign2 = r12z > 0.0 ? 1.0 : -1.0;
    *a21 = 2.0 * acos(dot2) * sign2;
}

/* Python-visible scalar helpers for closed-form pins. */
void oracle_legendre_00(double x, double *leg, int n_ell) { legendre_00(x, leg, n_ell); }
void oracle_legendre_22(double x, double *leg, double *f1, double *f2, int n_ell)
{
    legendre_22(x, leg, f1, f2, n_ell);
}
void oracle_legendre_02(double x, double *leg, int n_ell) { legendre_02(x, leg, n_ell); }

static inline double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* compute_00_contribution: mode 0 = symmetric (lower triangle + diagonal),
 * mode 1 = general rectangle. vec1 and vec2 are (npix, 3) row-major. */
void oracle_compute_00(const double *cl, int n_ell, double *S, int64_t ld,
                       const double *vec1, const double *vec2, int64_t npix, int mode)
{
    if (mode == 0) {
        double *leg = (double *)malloc(sizeof(double) * n_ell);
        legendre_00(1.0, leg, n_ell);
        double entry = 0.0;
        for (int ell = 0; ell < n_ell; ++ell) entry += cl[ell] * leg[ell];
        for (int64_t i = 0; i < npix; ++i) S[i + i * ld] = entry;
        free(leg);
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t i = 0; i < npix; ++i) {
            double *lg = (double *)malloc(sizeof(double) * n_ell);
            for (int64_t j = i + 1; j < npix; ++j) {
                double x = vec1[3 * j] * vec2[3 * i] + vec1[3 * j + 1] * vec2[3 * i + 1] +
                           vec1[3 * j + 2] * vec2[3 * i + 2];
                legendre_00(x, lg, n_ell);
                double val = 0.0;
                for (int ell = 0; ell < n_ell; ++ell) val += cl[ell] * lg[ell];
                S[j + i * ld] = val;
            }
            free(lg);
        }
    } else {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < npix; ++i) {
            double *lg = (double *)malloc(sizeof(double) * n_ell);
            for (int64_t j = 0; j < npix; ++j) {
                double x = vec1[3 * j] * vec2[3 * i] + vec1[3 * j + 1] * vec2[3 * i + 1] +
                           vec1[3 * j + 2] * vec2[3 * i + 2];
                legendre_00(x, lg, n_ell);
                double val = 0.0;
                for (int ell = 0; ell < n_ell; ++ell) val += cl[ell] * lg[ell];
                S[j + i * ld] = val;
            }
            free(lg);
        }
    }
}

/* compute_22_contribution: S is the (2n x 2n) [Q|U] block. */
void oracle_compute_22(const double *cl11, const double *cl22, const double *cl12, int n_ell,
                       double *S, int64_t ld, const double *vec1, const double *vec2, int64_t npix)
{
    double *leg = (double *)malloc(sizeof(double) * n_ell * 3);
    double *f1 = leg + n_ell, *f2 = leg + 2 * n_ell;
    legendre_22(1.0, leg, f1, f2, n_ell);
    double qq = 0.0, uu = 0.0, qu = 0.0;
    for (int ell = 2; ell < n_ell; ++ell) {
        qq += cl11[ell] * f1[ell] - cl22[ell] * f2[ell];
        uu += cl22[ell] * f1[ell] - cl11[ell] * f2[ell];
        qu += (f1[ell] + f2[ell]) * cl12[ell];
    }
    for (int64_t i = 0; i < npix; ++i) {
        S[i + i * ld] = qq;
        S[(i + npix) + i * ld] = qu;
    }
    for (int64_t i = npix; i < 2 * npix; ++i) S[i + i * ld] = uu;
    free(leg);

#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < npix; ++i) {
        double *lg = (double *)malloc(sizeof(double) * n_ell * 3);
        double *g1 = lg + n_ell, *g2 = lg + 2 * n_ell;
        for (int64_t j = i + 1; j < npix; ++j) {
            double x = vec1[3 * j] * vec2[3 * i] + vec1[3 * j + 1] * vec2[3 * i + 1] +
                       vec1[3 * j + 2] * vec2[3 * i + 2];
            legendre_22(x, lg, g1, g2, n_ell);
            double q = 0.0, u = 0.0, m = 0.0;
            for (int ell = 2; ell < n_ell; ++ell) {
                q += cl11[ell] * g1[ell] - cl22[ell] * g2[ell];
                u += cl22[ell] * g1[ell] - cl11[ell] * g2[ell];
                m += (g1[ell] + g2[ell]) * cl12[ell];
            }
            double a1, a2;
            oracle_rotation_angle(vec1 + 3 * j, vec2 + 3 * i, &a1, &a2);
            double c1 = cos(a1), c2 = cos(a2), s1 = -sin(a1), s2 = -sin(a2);
            double c1c2 = c1 * c2, s1s2 = s1 * s2, c1s2 = c1 * s2, s1c2 = s1 * c2;
            S[j + i * ld] = q * c1c2 + u * s1s2 + m * (c1s2 + s1c2);
            S[(j + npix) + (i + npix) * ld] = q * s1s2 + u * c1c2 - m * (c1s2 + s1c2);
            S[(j + npix) + i * ld] = -q * s1c2 + u * c1s2 + m * (c1c2 - s1s2);
            S[(i + npix) + j * ld] = -q * c1s2 + u * s1c2 + m * (c1c2 - s1s2);
        }
        free(lg);
    }
}

/* compute_02_contribution: S is the (2*n2 x n0) [Q;U] x T block. */
void oracle_compute_02(const double *cl12, const double *cl13, int n_ell, double *S, int64_t ld,
                       const double *vec0, const double *vec2, int64_t n0, int64_t n2)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n0; ++i) {
        double *lg = (double *)malloc(sizeof(double) * n_ell);
        for (int64_t j = 0; j < n2; ++j) {
            double x = vec2[3 * j] * vec0[3 * i] + vec2[3 * j + 1] * vec0[3 * i + 1] +
                       vec2[3 * j + 2] * vec0[3 * i + 2];
            legendre_02(x, lg, n_ell);
            double tq = 0.0, tu = 0.0;
            for (int ell = 2; ell < n_ell; ++ell) {
                tq -= cl12[ell] * lg[ell];
                tu -= cl13[ell] * lg[ell];
            }
            double a1, a2;
            oracle_rotation_angle(vec2 + 3 * j, vec0 + 3 * i, &a1, &a2);
            double c1 = cos(a1), s1 = sin(a1);
            S[j + i * ld] = tq * c1 - tu * s1;
            S[(j + n2) + i * ld] = tq * s1 + tu * c1;
        }
        free(lg);
    }
}

/* Symmetrise: copy the lower triangle to the upper (linalg.py:101-120). */
void oracle_copy_lower_to_upper(double *M, int64_t n, int64_t ld)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < i; ++j) M[j + i * ld] = M[i + j * ld];
}

/* Tr(A B) with 64-row block partials summed in fixed order (linalg.py:66-98).
 * A and B are row-major views as the reference indexes them (A[i,j], B[j,i]);
 * for symmetric inputs the storage order does not matter. */
double oracle_matrix_trace(const double *A, const double *B, int64_t n)
{
    const int64_t bs = 64;
    int64_t nb = (n + bs - 1) / bs;
    double *partial = (double *)calloc(nb, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int64_t ib = 0; ib < nb; ++ib) {
        int64_t i0 = ib * bs, i1 = i0 + bs < n ? i0 + bs : n;
        double acc[64];
        for (int64_t k = 0; k < i1 - i0; ++k) acc[k] = 0.0;
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = i0; i < i1; ++i) acc[i - i0] += A[i * n + j] * B[j * n + i];
        double s = 0.0;
        for (int64_t k = 0; k < i1 - i0; ++k) s += acc[k];
        partial[ib] = s;
    }
    double total = 0.0;
    for (int64_t ib = 0; ib < nb; ++ib) total += partial[ib];
    free(partial);
    return total;
}
