/*
 * jxamd_setup.c -- host setup steps that precede the solve phase.
 *
 * ABF ("alternate block factorisation") decoupling, as applied by the
 * reference driver before the solve (src/test_bsr_cprgmres.c:193-212 ->
 * jxf_BSRMatrixDecouple, jxfpamg-0220/csrc/cpr/jxf_bsr_decoup.c:1273-1348):
 * every block row is multiplied by the inverse of its (first-stored)
 * diagonal block, which is then set to the identity.  The arithmetic follows
 * the reference block by block (closed-form 2x2/3x3/4x4 inverses by
 * cofactors, Gauss-Jordan above 4, the same product orders), so the
 * decoupled system is bit-identical to the reference's.  One reference quirk
 * is kept on purpose because it defines the system that is solved: the
 * right-hand side is multiplied in place (jxf_smat_vec_mul(binv, svec, svec),
 * jxf_bsr_decoup.c:335-343 and :346-357), so components already overwritten
 * feed the later ones and the result is not D^{-1} b.  A second one is kept
 * for the same reason: the closed-form 4x4 inverse (jxf_smat_inv_4x4,
 * jxf_bsr_decoup.c:112-183) is not the inverse.  Its diagonal cofactors are
 * right, but off the diagonal it returns (-1)^(i+j) adj(A)_ji in place of
 * adj(A)_ij, and its determinant (line 127) pairs a12, a13, a14 with minors
 * that themselves contain the first row.  The decoupled 4-unknown blocks are
 * therefore not D^{-1} A_ij, the pressure rows lose the diagonal dominance
 * the AMG relies on, and the reference's CPR-GMRES stalls on such systems.
 *
 * The same holds above 4x4: the reference's Gauss-Jordan (jxf_bsr_decoup.c:
 * 200-266) row-reduces the block in place without an identity beside it, so
 * it returns a column permutation of the identity, not the inverse; the
 * restatement below follows it.
 *
 * jxamd_bsr_decouple_abf_exact is the documented deviation for b >= 4:
 * a true inverse (Gauss-Jordan on [D | I]) for every block size above 3
 * (1x1 to 3x3 keep the reference's exact closed forms), and an out-of-place
 * right-hand-side product, so the decoupled system is D^{-1} A x = D^{-1} b
 * with the solution of A x = b.
 *
 * Pressure extraction (jxf_BSRMatrixGetSubmatrix, jxfpamg-0220/csrc/
 * mat_vec_data/seq_bsr_matrix.c:214-250): A_p keeps A's block pattern with
 * the (p,p) entry of every block.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "jxamd.h"

/* closed-form inverses, jxf_bsr_decoup.c:53-183 */
static void inv2(double *a)
{
    const double a11 = a[0], a12 = a[1], a21 = a[2], a22 = a[3];
    const double det = a11 * a22 - a12 * a21;
    if (fabs(det) < 1e-15) {
        a[0] = 1.0; a[1] = 0.0; a[2] = 0.0; a[3] = 1.0;
    } else {
        double di = 1.0 / det;
        a[0] = a22 * di; a[1] = -a12 * di; a[2] = -a21 * di; a[3] = a11 * di;
    }
}

static void inv3(double *a)
{
    const double a11 = a[0], a12 = a[1], a13 = a[2];
    const double a21 = a[3], a22 = a[4], a23 = a[5];
    const double a31 = a[6], a32 = a[7], a33 = a[8];
    const double M11 = a22 * a33 - a23 * a32, M12 = a21 * a33 - a23 * a31, M13 = a21 * a32 - a22 * a31;
    const double M21 = a12 * a33 - a13 * a32, M22 = a11 * a33 - a13 * a31, M23 = a11 * a32 - a12 * a31;
    const double M31 = a12 * a23 - a13 * a22, M32 = a11 * a23 - a13 * a21, M33 = a11 * a22 - a12 * a21;
    const double det = a11 * M11 - a12 * M12 + a13 * M13;
    if (fabs(det) < 1e-15) {
        for (int i = 0; i < 9; i++) a[i] = 0.0;
        a[0] = a[4] = a[8] = 1.0;
        return;
    }
    double di = 1.0 / det;
    a[0] = M11 * di; a[1] = -M21 * di; a[2] = M31 * di;
    a[3] = -M12 * di; a[4] = M22 * di; a[5] = -M32 * di;
    a[6] = M13 * di; a[7] = -M23 * di; a[8] = M33 * di;
}

static void inv4(double *a)
{
    const double a11 = a[0], a12 = a[1], a13 = a[2], a14 = a[3];
    const double a21 = a[4], a22 = a[5], a23 = a[6], a24 = a[7];
    const double a31 = a[8], a32 = a[9], a33 = a[10], a34 = a[11];
    const double a41 = a[12], a42 = a[13], a43 = a[14], a44 = a[15];
    const double M11 = a22 * a33 * a44 + a23 * a34 * a42 + a24 * a32 * a43 - a22 * a34 * a43 - a23 * a32 * a44 - a24 * a33 * a42;
    const double M12 = a12 * a34 * a43 + a13 * a32 * a44 + a14 * a33 * a42 - a12 * a33 * a44 - a13 * a34 * a42 - a14 * a32 * a43;
    const double M13 = a12 * a23 * a44 + a13 * a24 * a42 + a14 * a22 * a43 - a12 * a24 * a43 - a13 * a22 * a44 - a14 * a23 * a42;
    const double M14 = a12 * a24 * a33 + a13 * a22 * a34 + a14 * a23 * a32 - a12 * a23 * a34 - a13 * a24 * a32 - a14 * a22 * a33;
    const double det = a11 * M11 - a12 * M12 + a13 * M13 - a14 * M14;
    if (fabs(det) < 1e-15) {
        for (int i = 0; i < 16; i++) a[i] = 0.0;
        a[0] = a[5] = a[10] = a[15] = 1.0;
        return;
    }
    const double di = 1.0 / det;
    const double M21 = a21 * a34 * a43 + a23 * a31 * a44 + a24 * a33 * a41 - a21 * a33 * a44 - a23 * a34 * a41 - a24 * a31 * a43;
    const double M22 = a11 * a33 * a44 + a13 * a34 * a41 + a14 * a31 * a43 - a11 * a34 * a43 - a13 * a31 * a44 - a14 * a33 * a41;
    const double M23 = a11 * a24 * a43 + a13 * a21 * a44 + a14 * a23 * a41 - a11 * a23 * a44 - a13 * a24 * a41 - a14 * a21 * a43;
    const double M24 = a11 * a23 * a34 + a13 * a24 * a31 + a14 * a21 * a33 - a11 * a24 * a33 - a13 * a21 * a34 - a14 * a23 * a31;
    const double M31 = a21 * a32 * a44 + a22 * a34 * a41 + a24 * a31 * a42 - a21 * a34 * a42 - a22 * a31 * a44 - a24 * a32 * a41;
    const double M32 = a11 * a34 * a42 + a12 * a31 * a44 + a14 * a32 * a41 - a11 * a32 * a44 - a12 * a34 * a41 - a14 * a31 * a42;
    const double M33 = a11 * a22 * a44 + a12 * a24 * a41 + a14 * a21 * a42 - a11 * a24 * a42 - a12 * a21 * a44 - a14 * a22 * a41;
    const double M34 = a11 * a24 * a32 + a12 * a21 * a34 + a14 * a22 * a31 - a11 * a22 * a34 - a12 * a24 * a31 - a14 * a21 * a32;
    const double M41 = a21 * a33 * a42 + a22 * a31 * a43 + a23 * a32 * a41 - a21 * a32 * a43 - a22 * a33 * a41 - a23 * a31 * a42;
    const double M42 = a11 * a32 * a43 + a12 * a33 * a41 + a13 * a31 * a42 - a11 * a33 * a42 - a12 * a31 * a43 - a13 * a32 * a41;
    const double M43 = a11 * a23 * a42 + a12 * a21 * a43 + a13 * a22 * a41 - a11 * a22 * a43 - a12 * a23 * a41 - a13 * a21 * a42;
    const double M44 = a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a11 * a23 * a32 - a12 * a21 * a33 - a13 * a22 * a31;
    a[0] = M11 * di;  a[1] = -M21 * di; a[2] = M31 * di;   a[3] = -M41 * di;
    a[4] = -M12 * di; a[5] = M22 * di;  a[6] = -M32 * di;  a[7] = M42 * di;
    a[8] = M13 * di;  a[9] = -M23 * di; a[10] = M33 * di;  a[11] = -M43 * di;
    a[12] = -M14 * di; a[13] = M24 * di; a[14] = -M34 * di; a[15] = M44 * di;
}

/* Gauss-Jordan with partial pivoting for n > 4, jxf_bsr_decoup.c:200-266 */
static void inv_gj(double *a, int n)
{
    int pivot[16];
    for (int i = 0; i < n; i++) pivot[i] = i;
    for (int i = 0; i < n; i++) {
        double max_val = fabs(a[i * n + i]);
        int max_idx = i;
        for (int j = i + 1; j < n; j++)
            if (fabs(a[j * n + i]) > max_val) { max_val = fabs(a[j * n + i]); max_idx = j; }
        if (max_val < 1e-15) {
            for (int j = 0; j < n * n; j++) a[j] = 0.0;
            for (int j = 0; j < n; j++) a[j * n + j] = 1.0;
            return;
        }
        if (max_idx != i) {
            for (int j = 0; j < n; j++) { double t = a[i * n + j]; a[i * n + j] = a[max_idx * n + j]; a[max_idx * n + j] = t; }
            int t = pivot[i]; pivot[i] = pivot[max_idx]; pivot[max_idx] = t;
        }
        double tmp = a[i * n + i];
        for (int j = 0; j < n; j++) a[i * n + j] /= tmp;
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            tmp = a[j * n + i];
            for (int k = 0; k < n; k++) a[j * n + k] -= tmp * a[i * n + k];
        }
    }
    double temp[16];
    for (int i = 0; i < n; i++) {
        if (pivot[i] != i) {
            for (int j = 0; j < n; j++) temp[j] = a[j * n + i];
            for (int j = 0; j < n; j++) a[j * n + i] = a[j * n + pivot[i]];
            for (int j = 0; j < n; j++) a[j * n + pivot[i]] = temp[j];
            for (int j = i + 1; j < n; j++)
                if (pivot[j] == i) { pivot[j] = pivot[i]; break; }
        }
    }
}

/* true inverse of an n x n block (n <= 10): Gauss-Jordan on [A | I] with
 * partial pivoting; a singular block gives the identity, as the reference's
 * routines do */
static void inv_exact(double *a, int n)
{
    double m[10][20];
    for (int r = 0; r < n; r++)
        for (int c = 0; c < 2 * n; c++) m[r][c] = c < n ? a[r * n + c] : (c - n == r ? 1.0 : 0.0);
    for (int i = 0; i < n; i++) {
        int p = i;
        for (int r = i + 1; r < n; r++)
            if (fabs(m[r][i]) > fabs(m[p][i])) p = r;
        if (fabs(m[p][i]) < 1e-15) {
            for (int r = 0; r < n; r++)
                for (int c = 0; c < n; c++) a[r * n + c] = r == c ? 1.0 : 0.0;
            return;
        }
        if (p != i)
            for (int c = 0; c < 2 * n; c++) { double t = m[i][c]; m[i][c] = m[p][c]; m[p][c] = t; }
        const double piv = m[i][i];
        for (int c = 0; c < 2 * n; c++) m[i][c] /= piv;
        for (int r = 0; r < n; r++) {
            if (r == i) continue;
            const double f = m[r][i];
            if (f != 0.0)
                for (int c = 0; c < 2 * n; c++) m[r][c] -= f * m[i][c];
        }
    }
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++) a[r * n + c] = m[r][n + c];
}

static int decouple(int nb, int b, const int *ia, double *a, double *rhs, int exact);

int jxamd_bsr_decouple_abf(int nb, int b, const int *ia, const int *ja, double *a, double *rhs)
{
    (void)ja;
    return decouple(nb, b, ia, a, rhs, 0);
}

int jxamd_bsr_decouple_abf_exact(int nb, int b, const int *ia, const int *ja, double *a, double *rhs)
{
    (void)ja;
    return decouple(nb, b, ia, a, rhs, 1);
}

static int decouple(int nb, int b, const int *ia, double *a, double *rhs, int exact)
{
    if (nb < 0 || b < 1 || b > 10 || !ia || !a) return JXAMD_ERROR_ARG;
    const int bb = b * b;
    double binv[100], smat[100];
    for (int i = 0; i < nb; i++) {
        if (ia[i + 1] <= ia[i]) return JXAMD_ERROR_ARG;
        double *row = a + (size_t)ia[i] * bb;
        memcpy(binv, row, sizeof(double) * bb);
        if (b == 1) binv[0] = fabs(binv[0]) > 1e-15 ? 1.0 / binv[0] : 1.0;
        else if (b == 2) inv2(binv);
        else if (b == 3) inv3(binv);
        else if (exact) inv_exact(binv, b); /* b >= 4 */
        else if (b == 4) inv4(binv);
        else inv_gj(binv, b);
        for (int k = ia[i]; k < ia[i + 1]; k++) {
            double *blk = a + (size_t)k * bb;
            if (k == ia[i]) {
                for (int r = 0; r < b; r++)
                    for (int c = 0; c < b; c++) blk[r * b + c] = (r == c) ? 1.0 : 0.0;
                continue;
            }
            if (b == 2) { /* jxf_smat_mul_2x2 */
                smat[0] = binv[0] * blk[0] + binv[1] * blk[2];
                smat[1] = binv[0] * blk[1] + binv[1] * blk[3];
                smat[2] = binv[2] * blk[0] + binv[3] * blk[2];
                smat[3] = binv[2] * blk[1] + binv[3] * blk[3];
            } else if (b == 4) { /* jxf_smat_mul_4x4: one expression per entry */
                for (int r = 0; r < 4; r++)
                    for (int c = 0; c < 4; c++)
                        smat[r * 4 + c] = binv[r * 4 + 0] * blk[0 * 4 + c] + binv[r * 4 + 1] * blk[1 * 4 + c] +
                                          binv[r * 4 + 2] * blk[2 * 4 + c] + binv[r * 4 + 3] * blk[3 * 4 + c];
            } else { /* jxf_smat_mul: accumulate from 0 */
                for (int r = 0; r < b; r++)
                    for (int c = 0; c < b; c++) {
                        double acc = 0.0;
                        for (int q = 0; q < b; q++) acc += binv[r * b + q] * blk[q * b + c];
                        smat[r * b + c] = acc;
                    }
            }
            memcpy(blk, smat, sizeof(double) * bb);
        }
        if (rhs && exact) {
            double *v = rhs + (size_t)i * b, w[10];
            for (int r = 0; r < b; r++) {
                w[r] = 0.0;
                for (int q = 0; q < b; q++) w[r] += binv[r * b + q] * v[q];
            }
            memcpy(v, w, sizeof(double) * b);
        } else if (rhs) {
            /* in-place product with aliased input and output, as in the reference */
            double *v = rhs + (size_t)i * b;
            if (b == 2) {
                v[0] = binv[0] * v[0] + binv[1] * v[1];
                v[1] = binv[2] * v[0] + binv[3] * v[1];
            } else if (b == 4) {
                v[0] = binv[0] * v[0] + binv[1] * v[1] + binv[2] * v[2] + binv[3] * v[3];
                v[1] = binv[4] * v[0] + binv[5] * v[1] + binv[6] * v[2] + binv[7] * v[3];
                v[2] = binv[8] * v[0] + binv[9] * v[1] + binv[10] * v[2] + binv[11] * v[3];
                v[3] = binv[12] * v[0] + binv[13] * v[1] + binv[14] * v[2] + binv[15] * v[3];
            } else {
                for (int r = 0; r < b; r++) {
                    v[r] = 0.0;
                    for (int q = 0; q < b; q++) v[r] += binv[r * b + q] * v[q];
                }
            }
        }
    }
    return JXAMD_SUCCESS;
}

/* jxf_smat_inv_nc (jxf_bsr_decoup.c:186-266): the closed forms up to 4x4,
 * Gauss-Jordan above */
static void inv_nc(double *a, int n)
{
    if (n == 1) a[0] = fabs(a[0]) > 1e-15 ? 1.0 / a[0] : 1.0;
    else if (n == 2) inv2(a);
    else if (n == 3) inv3(a);
    else if (n == 4) inv4(a);
    else inv_gj(a, n);
}

/* blk <- M blk, jxf_smat_mul (jxf_bsr_decoup.c:286-296): each entry summed
 * from zero in column order */
static void mul_generic(const double *M, double *blk, int b)
{
    double t[100];
    for (int r = 0; r < b; r++)
        for (int c = 0; c < b; c++) {
            double acc = 0.0;
            for (int q = 0; q < b; q++) acc += M[r * b + q] * blk[q * b + c];
            t[r * b + c] = acc;
        }
    memcpy(blk, t, sizeof(double) * b * b);
}

/* v <- M v in place with aliased input and output, jxf_smat_vec_mul
 * (jxf_bsr_decoup.c:335-343) called as (M, svec, svec): each component is
 * cleared before it is accumulated, so it reads its own (now zero) value and
 * the already updated earlier components */
static void vec_inplace(const double *M, double *v, int b)
{
    for (int r = 0; r < b; r++) {
        v[r] = 0.0;
        for (int q = 0; q < b; q++) v[r] += M[r * b + q] * v[q];
    }
}

/* ANL, SEM and QI (jxf_bsr_decoup.c:529-588, 591-657, 659-718): a matrix M
 * built from the row's diagonal block multiplies every block of the row,
 * the diagonal one included, and the rhs (in place) */
static int decouple_rows(int nb, int b, const int *ia, double *a, double *rhs, int type, int is_thermal)
{
    const int bb = b * b;
    double M[100], D[100];
    for (int i = 0; i < nb; i++) {
        if (ia[i + 1] <= ia[i]) return JXAMD_ERROR_ARG;
        memcpy(D, a + (size_t)ia[i] * bb, sizeof(double) * bb);
        if (type == JXAMD_DECOUP_SEM) {
            memcpy(M, D, sizeof(double) * bb);
            inv_nc(M, b);
            if (is_thermal) {
                M[0] = 1.0;
                M[b - 1] = 0.0;
                M[(b - 1) * b] = 0.0;
                M[bb - 1] = 1.0;
                for (int l = 0; l < b - 2; l++) {
                    M[1 + l] = -D[1 + l];
                    M[(b - 1) * b + 1 + l] = -D[(b - 1) * b + 1 + l];
                }
            } else {
                M[0] = 1.0;
                for (int l = 0; l < b - 1; l++) M[1 + l] = -D[1 + l];
            }
        } else {
            for (int r = 0; r < b; r++)
                for (int c = 0; c < b; c++) M[r * b + c] = r == c ? 1.0 : 0.0;
            const int nl = is_thermal ? b - 2 : b - 1;
            for (int l = 0; l < nl; l++) {
                if (type == JXAMD_DECOUP_ANL) {
                    M[1 + l] = -D[1 + l];
                    if (is_thermal) M[(b - 1) * b + 1 + l] = -D[(b - 1) * b + 1 + l];
                } else { /* QI: scaled by the diagonal entry, no guard */
                    M[1 + l] = -D[1 + l] / D[(l + 1) * b + l + 1];
                    if (is_thermal) M[(b - 1) * b + 1 + l] = -D[(b - 1) * b + 1 + l] / D[(l + 1) * b + l + 1];
                }
            }
        }
        for (int k = ia[i]; k < ia[i + 1]; k++) mul_generic(M, a + (size_t)k * bb, b);
        if (rhs) vec_inplace(M, rhs + (size_t)i * b, b);
    }
    return JXAMD_SUCCESS;
}

/* True-IMPES and its second form (jxf_bsr_decoup.c:721-808, 911-997,
 * 999-1099, 1102-1262): column sums Ac_c = sum_i A_ic (rows in ascending
 * order, from zero), then row i's pressure equation is recombined with
 * coefficients from Ac_i = Ac_ii.  The arithmetic and its quirks follow the
 * reference per block size: 2x2 uses -a12/a22 (first form, zero when Ac_ii
 * or a22 is near singular) or -a12/a22 with only the determinant guarded
 * (second form); 3x3 keeps the first two entries of Ac_ii^{-1}'s first row
 * for the matrix and all three for the rhs (first form), or adds c0 x row 1
 * + c1 x row 2 to row 0 (second form); other sizes replace row 0 by
 * (first row of Ac_ii^{-1}) x block (first form) or add the rest of that row
 * times rows 1.. (second form, which leaves the rhs unchanged). */
static int decouple_timpes(int nb, int b, const int *ia, const int *ja, double *a, double *rhs, int second)
{
    const int bb = b * b;
    double *Ac = (double *)calloc((size_t)nb * bb + 1, sizeof(double));
    double *cf = (double *)calloc((size_t)nb * b + 1, sizeof(double));
    if (!Ac || !cf) { free(Ac); free(cf); return JXAMD_ERROR_MEMORY; }
    for (int i = 0; i < nb; i++)
        for (int k = ia[i]; k < ia[i + 1]; k++) {
            if (ja[k] < 0 || ja[k] >= nb) { free(Ac); free(cf); return JXAMD_ERROR_ARG; }
            const double *blk = a + (size_t)k * bb;
            double *ac = Ac + (size_t)ja[k] * bb;
            for (int e = 0; e < bb; e++) ac[e] += blk[e];
        }
    for (int i = 0; i < nb; i++) {
        double binv[100];
        const double *ac = Ac + (size_t)i * bb;
        double *c = cf + (size_t)i * b;
        if (b == 2) {
            const double a11 = ac[0], a12 = ac[1], a21 = ac[2], a22 = ac[3];
            const double det = a11 * a22 - a12 * a21;
            if (fabs(det) < 1e-15) c[0] = 0.0;
            else if (second) c[0] = -a12 / a22;
            else c[0] = fabs(a22) > 1e-15 ? -a12 / a22 : 0.0;
        } else {
            memcpy(binv, ac, sizeof(double) * bb);
            if (b == 3) inv3(binv);
            else inv_nc(binv, b);
            if (!second) for (int k = 0; k < b; k++) c[k] = binv[k];
            else if (b == 3) { c[0] = binv[0]; c[1] = binv[1]; }
            else for (int k = 0; k < b - 1; k++) c[k] = binv[k + 1];
        }
    }
    for (int i = 0; i < nb; i++) {
        const double *c = cf + (size_t)i * b;
        for (int k = ia[i]; k < ia[i + 1]; k++) {
            double *blk = a + (size_t)k * bb;
            if (b == 2) {
                blk[0] = blk[0] + c[0] * blk[2];
                blk[1] = blk[1] + c[0] * blk[3];
            } else if (b == 3 && !second) {
                const double n0 = c[0] * blk[0] + c[1] * blk[3];
                const double n1 = c[0] * blk[1] + c[1] * blk[4];
                const double n2 = c[0] * blk[2] + c[1] * blk[5];
                blk[0] = n0; blk[1] = n1; blk[2] = n2;
            } else if (b == 3) {
                blk[0] = blk[0] + c[0] * blk[3] + c[1] * blk[6];
                blk[1] = blk[1] + c[0] * blk[4] + c[1] * blk[7];
                blk[2] = blk[2] + c[0] * blk[5] + c[1] * blk[8];
            } else if (!second) {
                double n[10];
                for (int col = 0; col < b; col++) {
                    n[col] = 0.0;
                    for (int q = 0; q < b; q++) n[col] += c[q] * blk[q * b + col];
                }
                for (int col = 0; col < b; col++) blk[col] = n[col];
            } else {
                for (int col = 0; col < b; col++) {
                    double sum = blk[col];
                    for (int q = 0; q < b - 1; q++) sum += c[q] * blk[(q + 1) * b + col];
                    blk[col] = sum;
                }
            }
        }
        if (!rhs) continue;
        double *r = rhs + (size_t)i * b;
        if (b == 2) r[0] = r[0] + c[0] * r[1];
        else if (b == 3 && !second) r[0] = c[0] * r[0] + c[1] * r[1] + c[2] * r[2];
        else if (b == 3) r[0] = r[0] + c[0] * r[1] + c[1] * r[2];
        else if (!second) {
            double n0 = 0.0;
            for (int q = 0; q < b; q++) n0 += c[q] * r[q];
            r[0] = n0;
        }
    }
    free(Ac);
    free(cf);
    return JXAMD_SUCCESS;
}

int jxamd_bsr_decouple(int nb, int b, const int *ia, const int *ja, double *a, double *rhs, int type, int is_thermal)
{
    if (nb < 0 || b < 1 || b > 10 || !ia || !ja || !a) return JXAMD_ERROR_ARG;
    switch (type) {
    case JXAMD_DECOUP_NONE: return JXAMD_SUCCESS;
    case JXAMD_DECOUP_ABF: return decouple(nb, b, ia, a, rhs, 0);
    case JXAMD_DECOUP_ANL:
    case JXAMD_DECOUP_SEM:
    case JXAMD_DECOUP_QI: return decouple_rows(nb, b, ia, a, rhs, type, is_thermal ? 1 : 0);
    case JXAMD_DECOUP_TIMPES: return decouple_timpes(nb, b, ia, ja, a, rhs, 0);
    case JXAMD_DECOUP_TIMPES2: return decouple_timpes(nb, b, ia, ja, a, rhs, 1);
    default: return JXAMD_ERROR_ARG;
    }
}

int jxamd_bsr_extract_pressure(int nb, int b, const int *ia, const int *ja, const double *a, int pidx, int *p_ia,
                               int *p_ja, double *p_a)
{
    if (nb < 0 || b < 1 || pidx < 0 || pidx >= b) return JXAMD_ERROR_ARG;
    const int bb = b * b, pos = pidx * b + pidx;
    for (int i = 0; i <= nb; i++) p_ia[i] = ia[i];
    for (int k = 0; k < ia[nb]; k++) {
        p_ja[k] = ja[k];
        p_a[k] = a[(size_t)k * bb + pos];
    }
    return JXAMD_SUCCESS;
}
