/*
 * cpu_ref.c -- TEST INFRASTRUCTURE: the oracle (see cpu_ref.h).
 *
 * FP64, one thread, reference operation order.  Paths below are relative to
 * the reference repository; J = jxfpamg-0220.
 */
#include "cpu_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* y = alpha*A*x + beta*y for BSR, J/csrc/operation/mvops/seq_bsr_matvec.c:21-134:
 * y *= beta/alpha (zeroed when that ratio is 0), y += A*x accumulated block by
 * block in column order, then y *= alpha. */
void cref_bsr_matvec(const cref_bsr *A, double alpha, const double *x, double beta, double *y)
{
    const int nb = A->nb, b = A->b, bb = b * b;
    const long n = (long)nb * b;
    if (alpha == 0.0) {
        for (long i = 0; i < n; i++) y[i] *= beta;
        return;
    }
    double temp = beta / alpha;
    if (temp != 1.0) {
        if (temp == 0.0)
            for (long i = 0; i < n; i++) y[i] = 0.0;
        else
            for (long i = 0; i < n; i++) y[i] *= temp;
    }
    for (int i = 0; i < nb; i++) {
        for (int jj = A->ia[i]; jj < A->ia[i + 1]; jj++) {
            const double *blk = A->a + (long)jj * bb;
            const double *xc = x + (long)A->ja[jj] * b;
            for (int r = 0; r < b; r++) {
                double t = y[(long)i * b + r];
                for (int s = 0; s < b; s++) t += blk[r * b + s] * xc[s];
                y[(long)i * b + r] = t;
            }
        }
    }
    if (alpha != 1.0)
        for (long i = 0; i < n; i++) y[i] *= alpha;
}

/* y = alpha*A*x + beta*y for CSR, J/csrc/operation/mvops/seq_csr_matvec.c:20-200
 * (all rows, one vector; the rownnz branch visits the same rows in the
 * same order). */
void cref_csr_matvec(const cref_csr *A, double alpha, const double *x, double beta, double *y)
{
    const int n = A->n;
    if (alpha == 0.0) {
        for (int i = 0; i < n; i++) y[i] *= beta;
        return;
    }
    double temp = beta / alpha;
    if (temp != 1.0) {
        if (temp == 0.0)
            for (int i = 0; i < n; i++) y[i] = 0.0;
        else
            for (int i = 0; i < n; i++) y[i] *= temp;
    }
    for (int i = 0; i < n; i++) {
        double t = y[i];
        for (int jj = A->ia[i]; jj < A->ia[i + 1]; jj++) t += A->a[jj] * x[A->ja[jj]];
        y[i] = t;
    }
    if (alpha != 1.0)
        for (int i = 0; i < n; i++) y[i] *= alpha;
}

/* y = alpha*A^T*x + beta*y, J/csrc/operation/mvops/seq_csr_matvec.c:211-330,
 * one-thread branch: scatter row by row. */
void cref_csr_matvecT(const cref_csr *A, double alpha, const double *x, double beta, double *y)
{
    const int n = A->n, m = A->ncols;
    if (alpha == 0.0) {
        for (int i = 0; i < m; i++) y[i] *= beta;
        return;
    }
    double temp = beta / alpha;
    if (temp != 1.0) {
        if (temp == 0.0)
            for (int i = 0; i < m; i++) y[i] = 0.0;
        else
            for (int i = 0; i < m; i++) y[i] *= temp;
    }
    for (int i = 0; i < n; i++)
        for (int jj = A->ia[i]; jj < A->ia[i + 1]; jj++) y[A->ja[jj]] += A->a[jj] * x[i];
    if (alpha != 1.0)
        for (int i = 0; i < m; i++) y[i] *= alpha;
}

/* J/csrc/operation/vecops/seq_vecop.c:239-260 */
double cref_dot(int n, const double *x, const double *y)
{
    double r = 0.0;
    for (int i = 0; i < n; i++) r += y[i] * x[i];
    return r;
}

/* Hybrid SSOR, relax_weight = omega = 1, one thread:
 * J/csrc/operation/relaxations/par_relax_6.c:224-273.  Forward sweep then
 * backward sweep; a row with a zero diagonal (stored first) is skipped. */
void cref_relax6(const cref_csr *A, const double *f, double *u)
{
    const int n = A->n;
    for (int i = 0; i < n; i++) {
        double d = A->a[A->ia[i]];
        if (d != 0.0) {
            double res = f[i];
            for (int jj = A->ia[i] + 1; jj < A->ia[i + 1]; jj++) res -= A->a[jj] * u[A->ja[jj]];
            u[i] = res / d;
        }
    }
    for (int i = n - 1; i > -1; i--) {
        double d = A->a[A->ia[i]];
        if (d != 0.0) {
            double res = f[i];
            for (int jj = A->ia[i] + 1; jj < A->ia[i + 1]; jj++) res -= A->a[jj] * u[A->ja[jj]];
            u[i] = res / d;
        }
    }
}

/* Hybrid SSOR with num_threads > 1, J/csrc/operation/relaxations/
 * par_relax_6.c:132-223: rows split into num_threads contiguous chunks
 * (the first n % T chunks one row longer); inside a chunk a forward then a
 * backward GS sweep, neighbours outside the chunk read from the copy of u
 * taken at entry.  Chunks are independent, so running them one after the
 * other reproduces the threaded result. */
void cref_relax6_threads(const cref_csr *A, const double *f, double *u, int num_threads)
{
    const int n = A->n;
    if (num_threads <= 1) {
        cref_relax6(A, f, u);
        return;
    }
    double *tmp = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) tmp[i] = u[i];
    for (int t = 0; t < num_threads; t++) {
        int size = n / num_threads, rest = n - size * num_threads, ns, ne;
        if (t < rest) {
            ns = t * size + t;
            ne = (t + 1) * size + t + 1;
        } else {
            ns = t * size + rest;
            ne = (t + 1) * size + rest;
        }
        for (int i = ns; i < ne; i++) {
            double d = A->a[A->ia[i]];
            if (d != 0.0) {
                double res = f[i];
                for (int jj = A->ia[i] + 1; jj < A->ia[i + 1]; jj++) {
                    int ii = A->ja[jj];
                    res -= A->a[jj] * ((ii >= ns && ii < ne) ? u[ii] : tmp[ii]);
                }
                u[i] = res / d;
            }
        }
        for (int i = ne - 1; i > ns - 1; i--) {
            double d = A->a[A->ia[i]];
            if (d != 0.0) {
                double res = f[i];
                for (int jj = A->ia[i] + 1; jj < A->ia[i + 1]; jj++) {
                    int ii = A->ja[jj];
                    res -= A->a[jj] * ((ii >= ns && ii < ne) ? u[ii] : tmp[ii]);
                }
                u[i] = res / d;
            }
        }
    }
    free(tmp);
}

/* Polynomial smoother (no reference counterpart): degree Richardson steps
 * with the Jacobi preconditioner, u <- u + w_k D^{-1} (f - A u); the error
 * propagates by prod_k (I - w_k D^{-1} A). */
void cref_poly_smooth(const cref_csr *A, const double *f, double *u, int degree, const double *omega)
{
    const int n = A->n;
    double *r = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int k = 0; k < degree; k++) {
        for (int i = 0; i < n; i++) {
            double acc = f[i];
            for (int jj = A->ia[i]; jj < A->ia[i + 1]; jj++) acc -= A->a[jj] * u[A->ja[jj]];
            r[i] = acc;
        }
        for (int i = 0; i < n; i++) {
            const double d = A->a[A->ia[i]];
            if (d != 0.0) u[i] += omega[k] * r[i] / d;
        }
    }
    free(r);
}

/* Gaussian elimination without pivoting, J/csrc/operation/matops/par_gs.c:20-74 */
int cref_gselim(double *A, double *x, int size)
{
    if (size == 1) {
        if (A[0] != 0.0) {
            x[0] = x[0] / A[0];
            return 0;
        }
        return 1;
    }
    for (int k = 0; k < size - 1; k++) {
        if (A[(long)k * size + k] != 0.0) {
            for (int j = k + 1; j < size; j++) {
                if (A[(long)j * size + k] != 0.0) {
                    double factor = A[(long)j * size + k] / A[(long)k * size + k];
                    for (int m = k + 1; m < size; m++) A[(long)j * size + m] -= factor * A[(long)k * size + m];
                    x[j] -= factor * x[k];
                }
            }
        }
    }
    for (int k = size - 1; k > 0; --k) {
        x[k] /= A[(long)k * size + k];
        for (int j = 0; j < k; j++)
            if (A[(long)j * size + k] != 0.0) x[j] -= x[k] * A[(long)j * size + k];
    }
    x[0] /= A[0];
    return 0;
}

/* Coarsest-level direct solve, J/csrc/operation/relaxations/par_relax_9.c:20-142:
 * densify the whole matrix and eliminate from scratch. */
int cref_relax9(const cref_csr *A, const double *f, double *u)
{
    const int n = A->n;
    if (n == 0) return 0;
    double *M = (double *)calloc((size_t)n * n, sizeof(double));
    double *bv = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) {
        for (int jj = A->ia[i]; jj < A->ia[i + 1]; jj++) M[(long)i * n + A->ja[jj]] = A->a[jj];
        bv[i] = f[i];
    }
    int err = cref_gselim(M, bv, n);
    for (int i = 0; i < n; i++) u[i] = bv[i];
    free(M);
    free(bv);
    return err;
}

/* ov = alpha*mat*v + beta*ov with a compensated (Kahan) row sum,
 * J/csrc/mat_vec_data/seq_bsr_matrix.c:841-883 */
void cref_bsr_block_matvec_stable(double alpha, const double *mat, const double *v, double beta, double *ov, int b)
{
    if (alpha == 0.0) {
        for (int j = 0; j < b; j++) ov[j] *= beta;
        return;
    }
    double tmp[8];
    for (int j = 0; j < b; j++) tmp[j] = ov[j] * beta;
    for (int i = 0; i < b; i++) {
        double sum = 0.0, c = 0.0;
        for (int j = 0; j < b; j++) {
            double y = mat[i * b + j] * v[j] - c;
            double t = sum + y;
            c = (t - sum) - y;
            sum = t;
        }
        tmp[i] += alpha * sum;
        ov[i] = tmp[i];
    }
}

/* ov = mat^{-1} v: b == 2 closed form with an ill-conditioning fallback,
 * otherwise partial-pivot LU, J/csrc/mat_vec_data/seq_bsr_matrix.c:948-1097.
 * The LU's "singular" branch stops the elimination early and is then
 * overwritten by the substitutions, exactly as there. */
void cref_bsr_block_invmatvec_stable(const double *mat, const double *v, double *ov, int b)
{
    if (b == 2) {
        double a11 = mat[0], a12 = mat[1], a21 = mat[2], a22 = mat[3];
        double det = a11 * a22 - a12 * a21;
        if (fabs(det) < 1e-12 * (fabs(a11 * a22) + fabs(a12 * a21))) {
            ov[0] = v[0] / (a11 + 1e-12);
            ov[1] = v[1] / (a22 + 1e-12);
            return;
        }
        double inv_det = 1.0 / det;
        double b1 = v[0], b2 = v[1];
        ov[0] = (a22 * b1 - a12 * b2) * inv_det;
        ov[1] = (-a21 * b1 + a11 * b2) * inv_det;
        return;
    }
    double m[64], vd[8], od[8];
    for (int i = 0; i < b * b; i++) m[i] = mat[i];
    for (int i = 0; i < b; i++) {
        vd[i] = v[i];
        od[i] = 0.0;
    }
    for (int i = 0; i < b; i++) {
        int max_row = i;
        double max_val = fabs(m[i * b + i]);
        for (int k = i + 1; k < b; k++) {
            if (fabs(m[k * b + i]) > max_val) {
                max_val = fabs(m[k * b + i]);
                max_row = k;
            }
        }
        if (max_val < 1e-15) {
            for (int j = 0; j < b; j++) {
                double diag = m[j * b + j];
                od[j] = (fabs(diag) > 1e-12) ? vd[j] / diag : 0.0;
            }
            break;
        }
        if (max_row != i) {
            for (int j = 0; j < b; j++) {
                double t = m[i * b + j];
                m[i * b + j] = m[max_row * b + j];
                m[max_row * b + j] = t;
            }
            double t = vd[i];
            vd[i] = vd[max_row];
            vd[max_row] = t;
        }
        double pivot = m[i * b + i];
        for (int k = i + 1; k < b; k++) {
            double factor = m[k * b + i] / pivot;
            m[k * b + i] = factor;
            for (int j = i + 1; j < b; j++) m[k * b + j] -= factor * m[i * b + j];
        }
    }
    for (int i = 0; i < b; i++) {
        od[i] = vd[i];
        for (int j = 0; j < i; j++) od[i] -= m[i * b + j] * od[j];
    }
    for (int i = b - 1; i >= 0; i--) {
        for (int j = i + 1; j < b; j++) od[i] -= m[i * b + j] * od[j];
        double diag = m[i * b + i];
        if (fabs(diag) > 1e-15)
            od[i] /= diag;
        else
            od[i] = 0.0;
    }
    for (int i = 0; i < b; i++) ov[i] = od[i];
}

/* One block hybrid SOR sweep on one rank, J/csrc/amg/par_relax_bsr.c:193-573
 * (num_threads fixed to 1 at :286).  vtemp receives the incoming x (:317). */
void cref_bsr_hgs(const cref_bsr *A, double *x, const double *rhs, double *vtemp, double w, double omega, int fwd)
{
    const int nb = A->nb, b = A->b, bb = b * b;
    const long n = (long)nb * b;
    double res[8], res0[8], res2[8], tmp[8], out[8];
    for (long i = 0; i < n; i++) vtemp[i] = x[i];
    int begin = 0, end = nb, step = 1;
    if (fwd == -1) {
        begin = nb - 1;
        end = -1;
        step = -1;
    }
    const double eps = 2.220446049250313e-16; /* JXF_REAL_EPSILON for double */
    if (fabs(w - 1) < eps && fabs(omega - 1) < eps) {
        for (int i = begin; i != end; i += step) {
            for (int k = 0; k < b; k++) res[k] = rhs[(long)i * b + k];
            for (int jj = A->ia[i] + 1; jj < A->ia[i + 1]; jj++) {
                int ii = A->ja[jj];
                cref_bsr_block_matvec_stable(-1.0, A->a + (long)jj * bb, x + (long)ii * b, 1.0, res, b);
            }
            cref_bsr_block_invmatvec_stable(A->a + (long)A->ia[i] * bb, res, out, b);
            for (int k = 0; k < b; k++) x[(long)i * b + k] = out[k];
        }
        return;
    }
    const double prod = 1.0 - w * omega;
    const double one_minus_omega = 1.0 - omega;
    for (int i = begin; i != end; i += step) {
        for (int k = 0; k < b; k++) {
            res[k] = rhs[(long)i * b + k];
            res0[k] = 0.0;
            res2[k] = 0.0;
        }
        for (int jj = A->ia[i] + 1; jj < A->ia[i + 1]; jj++) {
            int ii = A->ja[jj];
            cref_bsr_block_matvec_stable(-1.0, A->a + (long)jj * bb, x + (long)ii * b, 1.0, res0, b);
            cref_bsr_block_matvec_stable(1.0, A->a + (long)jj * bb, vtemp + (long)ii * b, 1.0, res2, b);
        }
        for (int k = 0; k < b; k++) tmp[k] = omega * res[k] + res0[k] + one_minus_omega * res2[k];
        cref_bsr_block_invmatvec_stable(A->a + (long)A->ia[i] * bb, tmp, out, b);
        for (int k = 0; k < b; k++) {
            x[(long)i * b + k] *= prod;
            x[(long)i * b + k] += w * out[k];
        }
    }
}

/* The same forward sweep as run on nparts MPI ranks, each owning a
 * contiguous range of block rows (first nb % nparts ranks one row more,
 * the row_part split of src/jxpamg_interface.c:177-199): Gauss-Seidel on the
 * rank's rows, neighbours on other ranks at their entry values in the
 * omega-weighted right-hand side, J/csrc/amg/par_relax_bsr.c:296-330
 * (Vext_data) and :523-557 (A_offd terms into res). */
void cref_bsr_hgs_parts(const cref_bsr *A, double *x, const double *rhs, double *vtemp, double w, double omega,
                        int nparts)
{
    const int nb = A->nb, b = A->b, bb = b * b;
    const long n = (long)nb * b;
    if (nparts <= 1) {
        cref_bsr_hgs(A, x, rhs, vtemp, w, omega, 1);
        return;
    }
    double res[8], res0[8], res2[8], tmp[8], out[8];
    for (long i = 0; i < n; i++) vtemp[i] = x[i];
    const double prod = 1.0 - w * omega, one_minus_omega = 1.0 - omega;
    const int size = nb / nparts, rest = nb - size * nparts;
    for (int t = 0; t < nparts; t++) {
        const int ns = t < rest ? t * (size + 1) : t * size + rest;
        const int ne = ns + size + (t < rest ? 1 : 0);
        for (int i = ns; i < ne; i++) {
            for (int k = 0; k < b; k++) {
                res[k] = rhs[(long)i * b + k];
                res0[k] = 0.0;
                res2[k] = 0.0;
            }
            for (int jj = A->ia[i] + 1; jj < A->ia[i + 1]; jj++) {
                const int ii = A->ja[jj];
                if (ii >= ns && ii < ne) {
                    cref_bsr_block_matvec_stable(-1.0, A->a + (long)jj * bb, x + (long)ii * b, 1.0, res0, b);
                    cref_bsr_block_matvec_stable(1.0, A->a + (long)jj * bb, vtemp + (long)ii * b, 1.0, res2, b);
                } else {
                    cref_bsr_block_matvec_stable(-1.0, A->a + (long)jj * bb, vtemp + (long)ii * b, 1.0, res, b);
                }
            }
            for (int k = 0; k < b; k++) tmp[k] = omega * res[k] + res0[k] + one_minus_omega * res2[k];
            cref_bsr_block_invmatvec_stable(A->a + (long)A->ia[i] * bb, tmp, out, b);
            for (int k = 0; k < b; k++) {
                x[(long)i * b + k] *= prod;
                x[(long)i * b + k] += w * out[k];
            }
        }
    }
}

/* One V-cycle, J/csrc/amg/par_cycle.c:20-969 with the CPR settings
 * (J/csrc/cpr/jxf_cpr.c:319-348): cycle_type 1, one sweep per visit, relax
 * type 6 down/up and 9 on the coarsest level, relax_order 1 so every visit
 * runs two hSGS calls (J/csrc/amg/par_relax_if.c:36-60 with relax_points
 * forced to 0 at J/csrc/amg/par_relax.c:48), restriction by P^T
 * (R_array == P_array). */
static int cref_pamg_cycle(const cref_hier *H, double **F, double **U, double *vtemp)
{
    const int L = H->num_levels;
    int lev_counter[CREF_MAX_LEVELS];
    lev_counter[0] = 1;
    for (int k = 1; k < L; k++) lev_counter[k] = 1;
    int level = 0, cycle_param = 1, err = 0;
    for (;;) {
        if (L > 1) {
            if (cycle_param < 3 && H->poly_degree > 0) {
                cref_poly_smooth(&H->A[level], F[level], U[level], H->poly_degree, H->poly_omega[level]);
            } else if (cycle_param < 3) {
                cref_relax6_threads(&H->A[level], F[level], U[level], H->relax_threads[level]);
                cref_relax6_threads(&H->A[level], F[level], U[level], H->relax_threads[level]);
            } else {
                err = cref_relax9(&H->A[level], F[level], U[level]);
                if (err) return err;
            }
        } else {
            /* no coarsening: one user-type (6) sweep, par_cycle.c:236-255 */
            if (H->poly_degree > 0)
                cref_poly_smooth(&H->A[level], F[level], U[level], H->poly_degree, H->poly_omega[level]);
            else
                cref_relax6_threads(&H->A[level], F[level], U[level], H->relax_threads[level]);
        }
        --lev_counter[level];
        if (lev_counter[level] >= 0 && level != L - 1) {
            const int fine = level, coarse = level + 1;
            for (int i = 0; i < H->A[coarse].n; i++) U[coarse][i] = 0.0;
            memcpy(vtemp, F[fine], sizeof(double) * H->A[fine].n);
            cref_csr_matvec(&H->A[fine], -1.0, U[fine], 1.0, vtemp);
            cref_csr_matvecT(&H->P[fine], 1.0, vtemp, 0.0, F[coarse]);
            ++level;
            if (lev_counter[level] < 1) lev_counter[level] = 1;
            cycle_param = (level == L - 1) ? 3 : 1;
        } else if (level != 0) {
            cref_csr_matvec(&H->P[level - 1], 1.0, U[level], 1.0, U[level - 1]);
            --level;
            cycle_param = 2;
        } else {
            break;
        }
    }
    return err;
}

/* jxf_PAMGSolve with max_iter 1, tol 1e-7, conv_criteria 0, rhsnrm_threshold 0
 * (J/csrc/amg/par_amg_solve.c:206-350, defaults J/csrc/amg/par_amg.c:240-246):
 * the cycle runs unless the initial relative residual is already below tol. */
int cref_pamg_solve(const cref_hier *H, const double *f, double *u)
{
    const int L = H->num_levels;
    const int n0 = H->A[0].n;
    int nmax = 0;
    for (int l = 0; l < L; l++) nmax = H->A[l].n > nmax ? H->A[l].n : nmax;
    double *vtemp = (double *)malloc(sizeof(double) * (nmax > 0 ? nmax : 1));
    memcpy(vtemp, f, sizeof(double) * n0);
    cref_csr_matvec(&H->A[0], -1.0, u, 1.0, vtemp);
    double resid = sqrt(cref_dot(n0, vtemp, vtemp));
    double rhs_norm = sqrt(cref_dot(n0, f, f));
    double relative = rhs_norm ? resid / rhs_norm : resid;
    int err = 0;
    if (relative >= 1.0e-7) {
        double *F[CREF_MAX_LEVELS], *U[CREF_MAX_LEVELS];
        F[0] = (double *)f;
        U[0] = u;
        for (int l = 1; l < L; l++) {
            F[l] = (double *)calloc((size_t)H->A[l].n + 1, sizeof(double));
            U[l] = (double *)calloc((size_t)H->A[l].n + 1, sizeof(double));
        }
        err = cref_pamg_cycle(H, F, U, vtemp);
        for (int l = 1; l < L; l++) {
            free(F[l]);
            free(U[l]);
        }
    }
    free(vtemp);
    return err;
}

/* x = M^{-1} r, J/csrc/cpr/jxf_cpr.c:424-610: restrict (:74-100), x_p := 0,
 * stage1_maxit AMG solves, prolongate type 0 (overwrite, :127-133), then
 * stage2_maxit forward block HGS sweeps on A x = r.  The debug residuals
 * printed at :518-524 and :599-605 change nothing and are not repeated. */
/* ---------------------------------------------------------------------
 * BILU(0) -- no reference restatement: the reference's BILU
 * (J/csrc/ilu/par_bilu_setup.c:17, jxf_ILUSetupFGPBILU_v3 :4996) is an
 * iterative fine-grained variant and is not on the test_bsr_cprgmres path.
 * This is the textbook exact block ILU(0) (IKJ order, no fill outside A's
 * pattern), pinned by its defining properties in the tests.
 * --------------------------------------------------------------------- */

/* In-place inverse of a b x b block, Gauss-Jordan with partial pivoting;
 * returns 1 if singular. */
static int block_inverse(double *M, int b)
{
    double inv[16], a[16];
    for (int i = 0; i < b * b; i++) a[i] = M[i];
    for (int i = 0; i < b; i++)
        for (int j = 0; j < b; j++) inv[i * b + j] = (i == j) ? 1.0 : 0.0;
    for (int c = 0; c < b; c++) {
        int pr = c;
        for (int r = c + 1; r < b; r++)
            if (fabs(a[r * b + c]) > fabs(a[pr * b + c])) pr = r;
        if (a[pr * b + c] == 0.0) return 1;
        if (pr != c)
            for (int j = 0; j < b; j++) {
                double t = a[c * b + j]; a[c * b + j] = a[pr * b + j]; a[pr * b + j] = t;
                t = inv[c * b + j]; inv[c * b + j] = inv[pr * b + j]; inv[pr * b + j] = t;
            }
        double d = 1.0 / a[c * b + c];
        for (int j = 0; j < b; j++) { a[c * b + j] *= d; inv[c * b + j] *= d; }
        for (int r = 0; r < b; r++) {
            if (r == c) continue;
            double f = a[r * b + c];
            if (f == 0.0) continue;
            for (int j = 0; j < b; j++) { a[r * b + j] -= f * a[c * b + j]; inv[r * b + j] -= f * inv[c * b + j]; }
        }
    }
    for (int i = 0; i < b * b; i++) M[i] = inv[i];
    return 0;
}

/* lu (same pattern as A): strictly lower blocks hold L (unit block lower),
 * the diagonal and upper blocks hold U; dinv[i] = U_ii^{-1}.  Columns of a
 * row must be ascending after the diagonal block. */
int cref_bilu0_factor(const cref_bsr *A, double *lu, double *dinv)
{
    const int nb = A->nb, b = A->b, bb = b * b;
    int *pos = (int *)malloc(sizeof(int) * (nb > 0 ? nb : 1));
    if (!pos) return 2;
    for (int j = 0; j < nb; j++) pos[j] = -1;
    for (long k = 0; k < (long)A->ia[nb] * bb; k++) lu[k] = A->a[k];
    int err = 0;
    for (int i = 0; i < nb && !err; i++) {
        for (int k = A->ia[i]; k < A->ia[i + 1]; k++) pos[A->ja[k]] = k;
        for (int kk = A->ia[i] + 1; kk < A->ia[i + 1]; kk++) {
            const int k = A->ja[kk];
            if (k > i) break;
            /* L_ik = A_ik U_kk^{-1} */
            double t[16];
            const double *Dk = dinv + (long)k * bb;
            double *Lik = lu + (long)kk * bb;
            for (int r = 0; r < b; r++)
                for (int c = 0; c < b; c++) {
                    double acc = 0.0;
                    for (int q = 0; q < b; q++) acc += Lik[r * b + q] * Dk[q * b + c];
                    t[r * b + c] = acc;
                }
            for (int e = 0; e < bb; e++) Lik[e] = t[e];
            /* A_ij -= L_ik U_kj for j > k in row k that are in row i */
            for (int jj = A->ia[k] + 1; jj < A->ia[k + 1]; jj++) {
                const int j = A->ja[jj];
                if (j <= k || pos[j] < 0) continue;
                double *Aij = lu + (long)pos[j] * bb;
                const double *Ukj = lu + (long)jj * bb;
                for (int r = 0; r < b; r++)
                    for (int c = 0; c < b; c++) {
                        double acc = 0.0;
                        for (int q = 0; q < b; q++) acc += Lik[r * b + q] * Ukj[q * b + c];
                        Aij[r * b + c] -= acc;
                    }
            }
        }
        double *Di = dinv + (long)i * bb;
        for (int e = 0; e < bb; e++) Di[e] = lu[(long)A->ia[i] * bb + e];
        if (block_inverse(Di, b)) err = 1;
        for (int k = A->ia[i]; k < A->ia[i + 1]; k++) pos[A->ja[k]] = -1;
    }
    free(pos);
    return err;
}

/* z = (LU)^{-1} r: y_i = r_i - sum_{k<i} L_ik y_k, then
 * z_i = U_ii^{-1} (y_i - sum_{j>i} U_ij z_j) */
void cref_bilu0_solve(const cref_bsr *A, const double *lu, const double *dinv, const double *r, double *z)
{
    const int nb = A->nb, b = A->b, bb = b * b;
    for (int i = 0; i < nb; i++) {
        double acc[4];
        for (int q = 0; q < b; q++) acc[q] = r[(long)i * b + q];
        for (int kk = A->ia[i] + 1; kk < A->ia[i + 1]; kk++) {
            const int k = A->ja[kk];
            if (k > i) continue;
            const double *L = lu + (long)kk * bb;
            for (int q = 0; q < b; q++)
                for (int c = 0; c < b; c++) acc[q] -= L[q * b + c] * z[(long)k * b + c];
        }
        for (int q = 0; q < b; q++) z[(long)i * b + q] = acc[q];
    }
    for (int i = nb - 1; i >= 0; i--) {
        double acc[4];
        for (int q = 0; q < b; q++) acc[q] = z[(long)i * b + q];
        for (int jj = A->ia[i] + 1; jj < A->ia[i + 1]; jj++) {
            const int j = A->ja[jj];
            if (j < i) continue;
            const double *U = lu + (long)jj * bb;
            for (int q = 0; q < b; q++)
                for (int c = 0; c < b; c++) acc[q] -= U[q * b + c] * z[(long)j * b + c];
        }
        const double *D = dinv + (long)i * bb;
        for (int q = 0; q < b; q++) {
            double o = 0.0;
            for (int c = 0; c < b; c++) o += D[q * b + c] * acc[c];
            z[(long)i * b + q] = o;
        }
    }
}

int cref_cpr_precond(const cref_bsr *A, const cref_hier *H, const cref_params *p, const double *r, double *x)
{
    const int nb = A->nb, b = A->b, pi = p->pressure_index;
    double *rp = (double *)malloc(sizeof(double) * nb);
    double *xp = (double *)calloc((size_t)nb, sizeof(double));
    double *work = (double *)malloc(sizeof(double) * (size_t)nb * b);
    for (int i = 0; i < nb; i++) rp[i] = r[(long)i * b + pi];
    int err = 0;
    for (int it = 0; it < p->stage1_maxit && !err; it++) err = cref_pamg_solve(H, rp, xp);
    for (int i = 0; i < nb; i++)
        for (int j = 0; j < b; j++) x[(long)i * b + j] = (j == pi) ? xp[i] : 0.0;
    if (p->stage2_solver_type == 5) {
        /* BILU(0) corrections from the prolongated pressure correction */
        const long n = (long)nb * b;
        double *res = (double *)malloc(sizeof(double) * n), *z = (double *)malloc(sizeof(double) * n);
        for (int it = 0; it < p->stage2_maxit; it++) {
            memcpy(res, r, sizeof(double) * n);
            cref_bsr_matvec(A, -1.0, x, 1.0, res);
            cref_bilu0_solve(A, p->bilu_lu, p->bilu_dinv, res, z);
            for (long q = 0; q < n; q++) x[q] += z[q];
        }
        free(res);
        free(z);
    } else {
        for (int it = 0; it < p->stage2_maxit; it++)
            cref_bsr_hgs_parts(A, x, r, work, p->stage2_weight, p->stage2_omega, p->stage2_parts);
    }
    free(rp);
    free(xp);
    free(work);
    return err;
}

/* Right-preconditioned restarted GMRES(k) with modified Gram-Schmidt,
 * J/csrc/krylov/pgmres.c:950-1713 with rel_change 0, cf_tol 0, min_iter 0,
 * convfac_threshold 1e4 (defaults :440-467).  res_history (length
 * max_iter+1) receives the norms[] log (:1154, :1307). */
int cref_gmres_solve(const cref_bsr *A, const cref_hier *H, const cref_params *p, const double *bvec, double *x,
                     double *res_history, cref_stats *st)
{
    const int k_dim = p->k_dim, max_iter = p->max_iter;
    const long n = (long)A->nb * A->b;
    const double r_tol = p->tol, a_tol = p->a_tol, epsmac = 1.e-16;
    const double convfac_threshold = p->convfac_threshold > 0.0 ? p->convfac_threshold : 1.0e4;
    double **P = (double **)malloc(sizeof(double *) * (k_dim + 1));
    for (int i = 0; i <= k_dim; i++) P[i] = (double *)malloc(sizeof(double) * n);
    double *R = (double *)malloc(sizeof(double) * n);
    double *W = (double *)malloc(sizeof(double) * n);
    double *rs = (double *)calloc(k_dim + 1, sizeof(double));
    double *c = (double *)calloc(k_dim, sizeof(double));
    double *s = (double *)calloc(k_dim, sizeof(double));
    double **hh = (double **)malloc(sizeof(double *) * (k_dim + 1));
    for (int i = 0; i <= k_dim; i++) hh[i] = (double *)calloc(k_dim, sizeof(double));
    int err = 0, converged = 0, nprec = 0, nfalse = 0;
    int i = 0, j, k, iter = 0;
    double t, gamma, r_norm, b_norm, den_norm, epsilon, r_norm_0;
    double normup, normdown, conv_factor = 0.0;

    memcpy(P[0], bvec, sizeof(double) * n);
    cref_bsr_matvec(A, -1.0, x, 1.0, P[0]);
    b_norm = sqrt(cref_dot(n, bvec, bvec));
    if (b_norm != 0.0 && (b_norm / b_norm) != (b_norm / b_norm)) { err = 1; goto done; }
    r_norm = sqrt(cref_dot(n, P[0], P[0]));
    r_norm_0 = r_norm;
    if (r_norm != 0.0 && (r_norm / r_norm) != (r_norm / r_norm)) { err = 1; goto done; }
    normdown = r_norm;
    if (res_history) res_history[0] = r_norm;
    den_norm = b_norm > 0.0 ? b_norm : r_norm;
    epsilon = a_tol > r_tol * den_norm ? a_tol : r_tol * den_norm;

    while (iter < max_iter) {
        rs[0] = r_norm;
        if (r_norm == 0.0) goto done;
        if (r_norm <= epsilon) {
            memcpy(R, bvec, sizeof(double) * n);
            cref_bsr_matvec(A, -1.0, x, 1.0, R);
            r_norm = sqrt(cref_dot(n, R, R));
            if (r_norm <= epsilon) {
                converged = 1;
                break;
            }
            nfalse++;
            if (p->is_check_restarted == 0) break;
        }
        t = 1.0 / r_norm;
        for (long q = 0; q < n; q++) P[0][q] *= t;
        i = 0;
        while (i < k_dim && iter < max_iter) {
            i++;
            iter++;
            memset(R, 0, sizeof(double) * n);
            err = cref_cpr_precond(A, H, p, P[i - 1], R);
            nprec++;
            if (err) goto done;
            cref_bsr_matvec(A, 1.0, R, 0.0, P[i]);
            for (j = 0; j < i; j++) {
                hh[j][i - 1] = cref_dot(n, P[j], P[i]);
                double a = -hh[j][i - 1];
                for (long q = 0; q < n; q++) P[i][q] += a * P[j][q];
            }
            t = sqrt(cref_dot(n, P[i], P[i]));
            hh[i][i - 1] = t;
            if (t != 0.0) {
                t = 1.0 / t;
                for (long q = 0; q < n; q++) P[i][q] *= t;
            }
            for (j = 1; j < i; j++) {
                t = hh[j - 1][i - 1];
                hh[j - 1][i - 1] = s[j - 1] * hh[j][i - 1] + c[j - 1] * t;
                hh[j][i - 1] = -s[j - 1] * t + c[j - 1] * hh[j][i - 1];
            }
            t = hh[i][i - 1] * hh[i][i - 1];
            t += hh[i - 1][i - 1] * hh[i - 1][i - 1];
            gamma = sqrt(t);
            if (gamma == 0.0) gamma = epsmac;
            c[i - 1] = hh[i - 1][i - 1] / gamma;
            s[i - 1] = hh[i][i - 1] / gamma;
            rs[i] = -hh[i][i - 1] * rs[i - 1];
            rs[i] /= gamma;
            rs[i - 1] = c[i - 1] * rs[i - 1];
            hh[i - 1][i - 1] = s[i - 1] * hh[i][i - 1] + c[i - 1] * hh[i - 1][i - 1];
            r_norm = fabs(rs[i]);
            normup = r_norm;
            conv_factor = normup / normdown;
            normdown = normup;
            if (res_history) res_history[iter] = r_norm;
            if (iter > 1 && conv_factor > convfac_threshold) break;
            if (r_norm <= epsilon) break;
        }
        /* solution update: back-solve, w = sum y_j p_j, x += M^{-1} w (:1477-1503) */
        rs[i - 1] = rs[i - 1] / hh[i - 1][i - 1];
        for (k = i - 2; k >= 0; k--) {
            t = 0.0;
            for (j = k + 1; j < i; j++) t -= hh[k][j] * rs[j];
            t += rs[k];
            rs[k] = t / hh[k][k];
        }
        memcpy(W, P[i - 1], sizeof(double) * n);
        for (long q = 0; q < n; q++) W[q] *= rs[i - 1];
        for (j = i - 2; j >= 0; j--)
            for (long q = 0; q < n; q++) W[q] += rs[j] * P[j][q];
        memset(R, 0, sizeof(double) * n);
        err = cref_cpr_precond(A, H, p, W, R);
        nprec++;
        if (err) goto done;
        for (long q = 0; q < n; q++) x[q] += 1.0 * R[q];
        if (r_norm <= epsilon) {
            memcpy(R, bvec, sizeof(double) * n);
            cref_bsr_matvec(A, -1.0, x, 1.0, R);
            r_norm = sqrt(cref_dot(n, R, R));
            if (r_norm <= epsilon) {
                converged = 1;
                break;
            }
            nfalse++;
            if (p->is_check_restarted == 1) {
                memcpy(P[0], R, sizeof(double) * n);
                i = 0;
            } else {
                break;
            }
        }
        if (iter > 1 && conv_factor > convfac_threshold) break;
        /* restart residual from the Givens recurrence (:1617-1637) */
        for (j = i; j > 0; j--) {
            rs[j - 1] = -s[j - 1] * rs[j];
            rs[j] = c[j - 1] * rs[j];
        }
        if (i) {
            double a = rs[i] - 1.0;
            for (long q = 0; q < n; q++) P[i][q] += a * P[i][q];
        }
        for (j = i - 1; j > 0; j--)
            for (long q = 0; q < n; q++) P[i][q] += rs[j] * P[j][q];
        if (i) {
            double a = rs[0] - 1.0;
            for (long q = 0; q < n; q++) P[0][q] += a * P[0][q];
            for (long q = 0; q < n; q++) P[0][q] += 1.0 * P[i][q];
        }
    }
done:
    if (st) {
        st->iterations = iter;
        st->converged = converged;
        st->num_precond = nprec;
        st->b_norm = b_norm;
        st->final_rel_res = b_norm > 0.0 ? r_norm / b_norm : r_norm / r_norm_0;
        st->num_false_conv = nfalse;
    }
    for (int q = 0; q <= k_dim; q++) {
        free(P[q]);
        free(hh[q]);
    }
    free(P);
    free(hh);
    free(R);
    free(W);
    free(rs);
    free(c);
    free(s);
    return err;
}
