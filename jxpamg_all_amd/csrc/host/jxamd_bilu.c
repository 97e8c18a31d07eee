/*
 * jxamd_bilu.c -- exact block ILU(0) factorisation on the host (IKJ order, no
 * fill outside A's block pattern).  The factors feed the GPU's level-scheduled
 * block triangular solves (stage-2 solver type 5).
 *
 * The reference's BILU (jxfpamg-0220/csrc/ilu/par_bilu_setup.c:17 with type 3,
 * jxf_ILUSetupFGPBILU_v3 :4996, and its solve par_bilu_solve.c:360-366) is an
 * iterative fine-grained approximation of these factors with Jacobi
 * triangular solves; this is the exact BILU(0) those iterations approach.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "jxamd.h"

/* b x b inverse by Gauss-Jordan with partial pivoting; 1 if singular */
static int inv_block(const double *M, double *out, int b)
{
    double a[JXAMD_MAX_BLOCK * JXAMD_MAX_BLOCK], x[JXAMD_MAX_BLOCK * JXAMD_MAX_BLOCK];
    memcpy(a, M, sizeof(double) * b * b);
    for (int i = 0; i < b * b; i++) x[i] = 0.0;
    for (int i = 0; i < b; i++) x[i * b + i] = 1.0;
    for (int c = 0; c < b; c++) {
        int pr = c;
        for (int r = c + 1; r < b; r++)
            if (fabs(a[r * b + c]) > fabs(a[pr * b + c])) pr = r;
        if (a[pr * b + c] == 0.0) return 1;
        if (pr != c)
            for (int j = 0; j < b; j++) {
                double t = a[c * b + j];
                a[c * b + j] = a[pr * b + j];
                a[pr * b + j] = t;
                t = x[c * b + j];
                x[c * b + j] = x[pr * b + j];
                x[pr * b + j] = t;
            }
        const double d = 1.0 / a[c * b + c];
        for (int j = 0; j < b; j++) {
            a[c * b + j] *= d;
            x[c * b + j] *= d;
        }
        for (int r = 0; r < b; r++) {
            if (r == c || a[r * b + c] == 0.0) continue;
            const double f = a[r * b + c];
            for (int j = 0; j < b; j++) {
                a[r * b + j] -= f * a[c * b + j];
                x[r * b + j] -= f * x[c * b + j];
            }
        }
    }
    memcpy(out, x, sizeof(double) * b * b);
    return 0;
}

int jxamd_bilu0_factor(int nb, int b, const int *ia, const int *ja, const double *a, double *lu, double *dinv)
{
    if (nb < 1 || b < 1 || b > JXAMD_MAX_BLOCK || !ia || !ja || !a || !lu || !dinv) return JXAMD_ERROR_ARG;
    const int bb = b * b;
    for (int i = 0; i < nb; i++) {
        if (ia[i + 1] <= ia[i] || ja[ia[i]] != i) return JXAMD_ERROR_ARG; /* diagonal block first */
        for (int k = ia[i] + 2; k < ia[i + 1]; k++)
            if (ja[k] <= ja[k - 1]) return JXAMD_ERROR_ARG; /* then ascending block columns */
    }
    int *slot = (int *)malloc(sizeof(int) * nb);
    if (!slot) return JXAMD_ERROR_MEMORY;
    for (int j = 0; j < nb; j++) slot[j] = -1;
    memcpy(lu, a, sizeof(double) * (size_t)ia[nb] * bb);
    int rc = 0;
    for (int i = 0; i < nb && !rc; i++) {
        for (int k = ia[i]; k < ia[i + 1]; k++) slot[ja[k]] = k;
        /* lower blocks in ascending column order: each is final when reached */
        for (int kk = ia[i] + 1; kk < ia[i + 1] && ja[kk] < i; kk++) {
            const int k = ja[kk];
            double *L = lu + (size_t)kk * bb;
            const double *Dk = dinv + (size_t)k * bb;
            double t[JXAMD_MAX_BLOCK * JXAMD_MAX_BLOCK];
            for (int r = 0; r < b; r++)
                for (int c = 0; c < b; c++) {
                    double s = 0.0;
                    for (int q = 0; q < b; q++) s += L[r * b + q] * Dk[q * b + c];
                    t[r * b + c] = s;
                }
            memcpy(L, t, sizeof(double) * bb);
            /* row i -= L_ik * (row k of U), restricted to row i's pattern */
            for (int jj = ia[k] + 1; jj < ia[k + 1]; jj++) {
                const int j = ja[jj];
                if (j <= k || slot[j] < 0) continue;
                double *Aij = lu + (size_t)slot[j] * bb;
                const double *Ukj = lu + (size_t)jj * bb;
                for (int r = 0; r < b; r++)
                    for (int c = 0; c < b; c++) {
                        double s = 0.0;
                        for (int q = 0; q < b; q++) s += L[r * b + q] * Ukj[q * b + c];
                        Aij[r * b + c] -= s;
                    }
            }
        }
        if (inv_block(lu + (size_t)ia[i] * bb, dinv + (size_t)i * bb, b)) rc = JXAMD_ERROR_GENERIC;
        for (int k = ia[i]; k < ia[i + 1]; k++) slot[ja[k]] = -1;
    }
    free(slot);
    return rc;
}
