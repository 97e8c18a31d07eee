/*
 * jxamd_amg_setup.c -- host construction of the pressure-AMG hierarchy.
 *
 * The solve phase imports the hierarchy the reference builds in
 * JXF_CPRSetup (jxfpamg-0220/csrc/cpr/jxf_cpr.c:319-377 -> JXF_PAMGSetup,
 * jxfpamg-0220/csrc/amg/par_amg_setup.c:50).  This file builds the same
 * hierarchy on one rank, for the settings CPR uses: classical strength
 * (theta 0.25, max_row_sum 0.9), HMIS coarsening (coarsen_type 10),
 * extended+i interpolation (interp_type 6, no truncation), Galerkin
 * A_{l+1} = P^T A_l P, coarsening until at most coarse_threshold points.
 * Every step keeps the reference's traversal and accumulation order, so the
 * operators come out bit-identical (checked against the reference's own
 * hierarchies by the tests):
 *   strength        jxf_PAMGCreateS            amg/par_strength.c:20
 *   HMIS            jxf_PAMGCoarsenHMIS        amg/par_coarsen_hmis.c:21-37
 *     = Ruge-Stueben first pass (Z_PT variant) amg/par_coarsen_ruge.c, amg/amg_linklist.c
 *     + PMIS with CF_init 1                    amg/par_coarsen_pmis.c, amg/par_indepset.c,
 *                                              utilities/random.c (Park-Miller, seed 2747)
 *   ext+i interp    jxf_PAMGBuildExtPIInterp   amg/par_interp_6.c:19
 *   RAP             jxf_PAMGBuildCoarseOperatorKT  operation/matops/par_rap.c:185
 *   level loop      amg/par_amg_setup.c:525-1320
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "jxamd.h"

#define C_PT 1
#define F_PT -1
#define Z_PT -2
#define SF_PT -3
#define UNDECIDED 0

typedef struct {
    int n, ncols;
    int *ia, *ja;
    double *a;
} mat;

static void mat_free(mat *m)
{
    free(m->ia); free(m->ja); free(m->a);
    memset(m, 0, sizeof(*m));
}

/* ------------------------------------------------------------------------ */
/* strength of connection (amg/par_strength.c:20-360, one rank, 1 function) */
/* ------------------------------------------------------------------------ */
static int create_strength(const mat *A, double theta, double max_row_sum, int **S_i, int **S_j)
{
    const int n = A->n;
    const int nnz = A->ia[n];
    int *si = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int *sj = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    if (!si || !sj) { free(si); free(sj); return JXAMD_ERROR_MEMORY; }
    for (int k = 0; k < nnz; k++) sj[k] = A->ja[k];
    for (int i = 0; i < n; i++) {
        const int b = A->ia[i], e = A->ia[i + 1];
        const double diag = A->a[b];
        double row_scale = 0.0, row_sum = diag;
        if (diag < 0) {
            for (int k = b + 1; k < e; k++) {
                row_scale = row_scale > A->a[k] ? row_scale : A->a[k];
                row_sum += A->a[k];
            }
        } else {
            for (int k = b + 1; k < e; k++) {
                row_scale = row_scale < A->a[k] ? row_scale : A->a[k];
                row_sum += A->a[k];
            }
        }
        row_sum = fabs(row_sum / diag);
        sj[b] = -1;
        if (row_sum > max_row_sum && max_row_sum < 1.0) {
            for (int k = b + 1; k < e; k++) sj[k] = -1;
        } else if (diag < 0) {
            for (int k = b + 1; k < e; k++)
                if (A->a[k] <= theta * row_scale) sj[k] = -1;
        } else {
            for (int k = b + 1; k < e; k++)
                if (A->a[k] >= theta * row_scale) sj[k] = -1;
        }
    }
    int js = 0;
    for (int i = 0; i < n; i++) {
        si[i] = js;
        for (int k = A->ia[i]; k < A->ia[i + 1]; k++)
            if (sj[k] > -1) sj[js++] = sj[k];
    }
    si[n] = js;
    *S_i = si;
    *S_j = sj;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* list of lists keyed by measure, with FIFO order inside a measure         */
/* (same observable behaviour as amg/amg_linklist.c)                        */
/* ------------------------------------------------------------------------ */
typedef struct {
    int maxm;      /* capacity in measures */
    int *head, *tail;
    int *next, *prev; /* per point */
    int top;       /* largest measure that may be non-empty */
} lol;

static int lol_init(lol *L, int n, int maxm)
{
    L->maxm = maxm;
    L->head = (int *)malloc(sizeof(int) * ((size_t)maxm + 1));
    L->tail = (int *)malloc(sizeof(int) * ((size_t)maxm + 1));
    L->next = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    L->prev = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    if (!L->head || !L->tail || !L->next || !L->prev) return JXAMD_ERROR_MEMORY;
    for (int m = 0; m <= maxm; m++) L->head[m] = L->tail[m] = -1;
    L->top = 0;
    return 0;
}

static void lol_free(lol *L)
{
    free(L->head); free(L->tail); free(L->next); free(L->prev);
}

static int lol_grow(lol *L, int m)
{
    if (m <= L->maxm) return 0;
    int nm = L->maxm * 2 > m ? L->maxm * 2 : m;
    int *h = (int *)realloc(L->head, sizeof(int) * ((size_t)nm + 1));
    if (!h) return JXAMD_ERROR_MEMORY;
    L->head = h;
    int *t = (int *)realloc(L->tail, sizeof(int) * ((size_t)nm + 1));
    if (!t) return JXAMD_ERROR_MEMORY;
    L->tail = t;
    for (int q = L->maxm + 1; q <= nm; q++) L->head[q] = L->tail[q] = -1;
    L->maxm = nm;
    return 0;
}

static int lol_enter(lol *L, int m, int idx)
{
    if (lol_grow(L, m)) return JXAMD_ERROR_MEMORY;
    L->next[idx] = -1;
    L->prev[idx] = L->tail[m];
    if (L->tail[m] >= 0) L->next[L->tail[m]] = idx;
    else L->head[m] = idx;
    L->tail[m] = idx;
    if (m > L->top) L->top = m;
    return 0;
}

static void lol_remove(lol *L, int m, int idx)
{
    if (m < 0 || m > L->maxm || L->head[m] < 0) return; /* "No such list" */
    int p = L->prev[idx], q = L->next[idx];
    if (p >= 0) L->next[p] = q;
    else L->head[m] = q;
    if (q >= 0) L->prev[q] = p;
    else L->tail[m] = p;
}

static int lol_first(lol *L)
{
    while (L->top > 0 && L->head[L->top] < 0) L->top--;
    return L->head[L->top];
}

/* Ruge-Stueben first pass with coarsen_type 10 (f_pnt = Z_PT), one rank,
 * measure_type 0: amg/par_coarsen_ruge.c (first coarsening phase) */
static int ruge_first_pass(int n, const int *S_i, const int *S_j, int *CF)
{
    const int nnz = S_i[n];
    int *ST_i = (int *)calloc((size_t)n + 1, sizeof(int));
    int *ST_j = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    int *measure = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    lol L;
    memset(&L, 0, sizeof(L));
    int e = 0;
    if (!ST_i || !ST_j || !measure) { e = JXAMD_ERROR_MEMORY; goto out; }
    for (int k = 0; k < nnz; k++) ST_i[S_j[k] + 1]++;
    for (int i = 0; i < n; i++) ST_i[i + 1] += ST_i[i];
    for (int i = 0; i < n; i++)
        for (int k = S_i[i]; k < S_i[i + 1]; k++) ST_j[ST_i[S_j[k]]++] = i;
    for (int i = n; i > 0; i--) ST_i[i] = ST_i[i - 1];
    ST_i[0] = 0;
    int maxm = 8;
    for (int i = 0; i < n; i++) {
        measure[i] = ST_i[i + 1] - ST_i[i];
        if (measure[i] + 2 > maxm) maxm = measure[i] + 2;
    }
    if ((e = lol_init(&L, n, maxm))) goto out;
    int num_left = 0;
    for (int j = 0; j < n; j++) {
        if (S_i[j + 1] - S_i[j] == 0) {
            CF[j] = SF_PT;
            measure[j] = 0;
        } else {
            CF[j] = UNDECIDED;
            num_left++;
        }
    }
    for (int j = 0; j < n; j++) {
        int m = measure[j];
        if (CF[j] == SF_PT) continue;
        if (m > 0) {
            if ((e = lol_enter(&L, m, j))) goto out;
        } else {
            CF[j] = Z_PT;
            for (int k = S_i[j]; k < S_i[j + 1]; k++) {
                int nb = S_j[k];
                if (CF[nb] == SF_PT) continue;
                if (nb < j) {
                    int nm = measure[nb];
                    if (nm > 0) lol_remove(&L, nm, nb);
                    nm = ++measure[nb];
                    if ((e = lol_enter(&L, nm, nb))) goto out;
                } else {
                    ++measure[nb];
                }
            }
            --num_left;
        }
    }
    while (num_left > 0) {
        int index = lol_first(&L);
        if (index < 0) break; /* lists exhausted */
        CF[index] = C_PT;
        int m = measure[index];
        measure[index] = 0;
        --num_left;
        lol_remove(&L, m, index);
        for (int k = ST_i[index]; k < ST_i[index + 1]; k++) {
            int nb = ST_j[k];
            if (CF[nb] != UNDECIDED) continue;
            CF[nb] = F_PT;
            lol_remove(&L, measure[nb], nb);
            --num_left;
            for (int q = S_i[nb]; q < S_i[nb + 1]; q++) {
                int n2 = S_j[q];
                if (CF[n2] == UNDECIDED) {
                    lol_remove(&L, measure[n2], n2);
                    int nm = ++measure[n2];
                    if ((e = lol_enter(&L, nm, n2))) goto out;
                }
            }
        }
        for (int k = S_i[index]; k < S_i[index + 1]; k++) {
            int nb = S_j[k];
            if (CF[nb] != UNDECIDED) continue;
            int mm = measure[nb];
            lol_remove(&L, mm, nb);
            measure[nb] = --mm;
            if (mm > 0) {
                if ((e = lol_enter(&L, mm, nb))) goto out;
            } else {
                CF[nb] = F_PT;
                --num_left;
                for (int q = S_i[nb]; q < S_i[nb + 1]; q++) {
                    int n2 = S_j[q];
                    if (CF[n2] == UNDECIDED) {
                        lol_remove(&L, measure[n2], n2);
                        int nm = ++measure[n2];
                        if ((e = lol_enter(&L, nm, n2))) goto out;
                    }
                }
            }
        }
    }
out:
    lol_free(&L);
    free(ST_i); free(ST_j); free(measure);
    return e;
}

/* Park-Miller minimal standard generator, utilities/random.c:25-57 */
static int g_seed = 13579;
static double rand_next(void)
{
    const int a = 16807, m = 2147483647, q = 127773, r = 2836;
    int high = g_seed / q, low = g_seed % q;
    int test = a * low - r * high;
    g_seed = test > 0 ? test : test + m;
    return (double)g_seed / m;
}

/* PMIS with CF_init 1 on one rank (amg/par_coarsen_pmis.c, amg/par_indepset.c) */
static int pmis(int n, const int *S_i, int *S_j, int *CF)
{
    double *measure = (double *)calloc((size_t)n + 1, sizeof(double));
    int *graph = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    if (!measure || !graph) { free(measure); free(graph); return JXAMD_ERROR_MEMORY; }
    for (int k = 0; k < S_i[n]; k++) measure[S_j[k]] += 1.0;
    g_seed = 2747; /* jxf_PAMGIndepSetInit: 2747 + my_id */
    for (int i = 0; i < n; i++) measure[i] += rand_next();
    int cnt = 0;
    for (int i = 0; i < n; i++) {
        if (CF[i] == -1) CF[i] = 0;
        if (CF[i] == Z_PT) {
            if (measure[i] >= 1.0 || (S_i[i + 1] - S_i[i]) > 0) {
                CF[i] = 0;
                graph[cnt++] = i;
            } else {
                CF[i] = F_PT;
            }
        } else if (CF[i] == SF_PT) {
            measure[i] = 0;
        } else {
            graph[cnt++] = i;
        }
    }
    int graph_size = cnt, iter = 0;
    while (graph_size > 0) {
        if (iter) {
            /* jxf_PAMGIndepSet */
            for (int ig = 0; ig < graph_size; ig++) {
                int i = graph[ig];
                if (measure[i] > 1) CF[i] = 1;
            }
            for (int ig = 0; ig < graph_size; ig++) {
                int i = graph[ig];
                if (measure[i] > 1) {
                    for (int k = S_i[i]; k < S_i[i + 1]; k++) {
                        int j = S_j[k];
                        if (j < 0) j = -j - 1;
                        if (measure[j] > 1) {
                            if (measure[i] > measure[j]) CF[j] = 0;
                            else if (measure[j] > measure[i]) CF[i] = 0;
                        }
                    }
                }
            }
        }
        iter++;
        for (int ig = 0; ig < graph_size; ig++) {
            int i = graph[ig];
            if (measure[i] < 1.0) {
                CF[i] = F_PT;
                for (int k = S_i[i]; k < S_i[i + 1]; k++)
                    if (S_j[k] > -1) S_j[k] = -S_j[k] - 1;
            }
            if (CF[i] > 0) {
                CF[i] = C_PT;
                for (int k = S_i[i]; k < S_i[i + 1]; k++)
                    if (S_j[k] > -1) S_j[k] = -S_j[k] - 1;
            } else {
                for (int k = S_i[i]; k < S_i[i + 1]; k++) {
                    int j = S_j[k];
                    if (j < 0) j = -j - 1;
                    if (CF[j] > 0) CF[i] = F_PT;
                }
                if (CF[i] == F_PT)
                    for (int k = S_i[i]; k < S_i[i + 1]; k++)
                        if (S_j[k] > -1) S_j[k] = -S_j[k] - 1;
            }
        }
        for (int ig = 0; ig < graph_size; ig++) {
            int i = graph[ig];
            if (CF[i] == 0) {
                for (int k = S_i[i]; k < S_i[i + 1]; k++) {
                    int j = S_j[k];
                    if (j < 0) j = -j - 1;
                    if (CF[j] != 0 && S_j[k] > -1) S_j[k] = -S_j[k] - 1;
                }
            }
        }
        for (int ig = 0; ig < graph_size; ig++) {
            int i = graph[ig];
            if (CF[i] != 0) {
                measure[i] = 0;
                graph_size--;
                graph[ig] = graph[graph_size];
                graph[graph_size] = i;
                ig--;
            }
        }
    }
    for (int k = 0; k < S_i[n]; k++)
        if (S_j[k] < 0) S_j[k] = -S_j[k] - 1;
    /* special F points keep SF_PT (-3) here: the reference's conversion to
     * F_PT is commented out (amg/par_coarsen_pmis.c:688-695), so the
     * interpolation skips them (amg/par_interp_6.c, CF_marker != -3) */
    free(measure);
    free(graph);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* extended+i interpolation, one rank (amg/par_interp_6.c:19-1140)          */
/* ------------------------------------------------------------------------ */
static int build_ext_pi(const mat *A, const int *CF, const int *S_i, const int *S_j, mat *P)
{
    const int n = A->n;
    int *P_i = (int *)calloc((size_t)n + 1, sizeof(int));
    int *f2c = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    int *mark = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    if (!P_i || !f2c || !mark) { free(P_i); free(f2c); free(mark); return JXAMD_ERROR_MEMORY; }
    for (int i = 0; i < n; i++) { f2c[i] = -1; mark[i] = -1; }
    /* structure */
    int cnt = 0, nc = 0;
    for (int i = 0; i < n; i++) {
        P_i[i] = cnt;
        if (CF[i] >= 0) {
            cnt++;
            f2c[i] = nc++;
        } else if (CF[i] != -3) {
            for (int jj = S_i[i]; jj < S_i[i + 1]; jj++) {
                int i1 = S_j[jj];
                if (CF[i1] >= 0) {
                    if (mark[i1] < P_i[i]) mark[i1] = cnt++;
                } else if (CF[i1] != -3) {
                    for (int kk = S_i[i1]; kk < S_i[i1 + 1]; kk++) {
                        int k1 = S_j[kk];
                        if (CF[k1] >= 0 && mark[k1] < P_i[i]) mark[k1] = cnt++;
                    }
                }
            }
        }
    }
    P_i[n] = cnt;
    int *P_j = (int *)malloc(sizeof(int) * (cnt > 0 ? cnt : 1));
    double *P_a = (double *)calloc((size_t)(cnt > 0 ? cnt : 1), sizeof(double));
    if (!P_j || !P_a) { free(P_i); free(f2c); free(mark); free(P_j); free(P_a); return JXAMD_ERROR_MEMORY; }
    for (int i = 0; i < n; i++) mark[i] = -1;
    int jc = 0, strong_f_marker = -2;
    for (int i = 0; i < n; i++) {
        const int jj_begin = jc;
        if (CF[i] >= 0) {
            P_j[jc] = f2c[i];
            P_a[jc] = 1.0;
            jc++;
        } else if (CF[i] != -3) {
            strong_f_marker--;
            for (int jj = S_i[i]; jj < S_i[i + 1]; jj++) {
                int i1 = S_j[jj];
                if (CF[i1] >= 0) {
                    if (mark[i1] < jj_begin) {
                        mark[i1] = jc;
                        P_j[jc] = f2c[i1];
                        P_a[jc] = 0.0;
                        jc++;
                    }
                } else if (CF[i1] != -3) {
                    mark[i1] = strong_f_marker;
                    for (int kk = S_i[i1]; kk < S_i[i1 + 1]; kk++) {
                        int k1 = S_j[kk];
                        if (CF[k1] >= 0 && mark[k1] < jj_begin) {
                            mark[k1] = jc;
                            P_j[jc] = f2c[k1];
                            P_a[jc] = 0.0;
                            jc++;
                        }
                    }
                }
            }
            const int jj_end = jc;
            double diagonal = A->a[A->ia[i]];
            for (int jj = A->ia[i] + 1; jj < A->ia[i + 1]; jj++) {
                int i1 = A->ja[jj];
                if (mark[i1] >= jj_begin) {
                    P_a[mark[i1]] += A->a[jj];
                } else if (mark[i1] == strong_f_marker) {
                    double sum = 0.0;
                    int sgn = A->a[A->ia[i1]] < 0 ? -1 : 1;
                    for (int jj1 = A->ia[i1] + 1; jj1 < A->ia[i1 + 1]; jj1++) {
                        int i2 = A->ja[jj1];
                        if ((mark[i2] >= jj_begin || i2 == i) && (sgn * A->a[jj1]) < 0) sum += A->a[jj1];
                    }
                    if (sum != 0) {
                        double distribute = A->a[jj] / sum;
                        for (int jj1 = A->ia[i1] + 1; jj1 < A->ia[i1 + 1]; jj1++) {
                            int i2 = A->ja[jj1];
                            if (mark[i2] >= jj_begin && (sgn * A->a[jj1]) < 0)
                                P_a[mark[i2]] += distribute * A->a[jj1];
                            if (i2 == i && (sgn * A->a[jj1]) < 0) diagonal += distribute * A->a[jj1];
                        }
                    } else {
                        diagonal += A->a[jj];
                    }
                } else if (CF[i1] != -3) {
                    diagonal += A->a[jj];
                }
            }
            for (int jj = jj_begin; jj < jj_end; jj++) P_a[jj] /= -diagonal;
        }
        strong_f_marker--;
    }
    free(f2c);
    free(mark);
    P->n = n;
    P->ncols = nc;
    P->ia = P_i;
    P->ja = P_j;
    P->a = P_a;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Galerkin coarse operator P^T A P (operation/matops/par_rap.c:185, one    */
/* rank, one thread): row ic starts with its diagonal, then columns in order */
/* of first touch over R = P^T row ic -> A row i1 -> P row i2.              */
/* ------------------------------------------------------------------------ */
static int build_rap(const mat *A, const mat *P, mat *C)
{
    const int n = A->n, nc = P->ncols;
    /* R = P^T, rows ascending in the fine index */
    int nnzP = P->ia[n];
    int *R_i = (int *)calloc((size_t)nc + 1, sizeof(int));
    int *R_j = (int *)malloc(sizeof(int) * (nnzP > 0 ? nnzP : 1));
    double *R_a = (double *)malloc(sizeof(double) * (nnzP > 0 ? nnzP : 1));
    int *P_marker = (int *)malloc(sizeof(int) * (nc > 0 ? nc : 1));
    int *A_marker = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    if (!R_i || !R_j || !R_a || !P_marker || !A_marker) {
        free(R_i); free(R_j); free(R_a); free(P_marker); free(A_marker);
        return JXAMD_ERROR_MEMORY;
    }
    for (int k = 0; k < nnzP; k++) R_i[P->ja[k] + 1]++;
    for (int c = 0; c < nc; c++) R_i[c + 1] += R_i[c];
    {
        int *pos = (int *)malloc(sizeof(int) * ((size_t)nc + 1));
        if (!pos) { free(R_i); free(R_j); free(R_a); free(P_marker); free(A_marker); return JXAMD_ERROR_MEMORY; }
        memcpy(pos, R_i, sizeof(int) * ((size_t)nc + 1));
        for (int i = 0; i < n; i++)
            for (int k = P->ia[i]; k < P->ia[i + 1]; k++) {
                int q = pos[P->ja[k]]++;
                R_j[q] = i;
                R_a[q] = P->a[k];
            }
        free(pos);
    }
    /* symbolic pass */
    for (int c = 0; c < nc; c++) P_marker[c] = -1;
    for (int i = 0; i < n; i++) A_marker[i] = -1;
    long cnt = 0;
    for (int ic = 0; ic < nc; ic++) {
        long row_begin = cnt;
        P_marker[ic] = (int)cnt++;
        for (int jj1 = R_i[ic]; jj1 < R_i[ic + 1]; jj1++) {
            int i1 = R_j[jj1];
            for (int jj2 = A->ia[i1]; jj2 < A->ia[i1 + 1]; jj2++) {
                int i2 = A->ja[jj2];
                if (A_marker[i2] != ic) {
                    A_marker[i2] = ic;
                    for (int jj3 = P->ia[i2]; jj3 < P->ia[i2 + 1]; jj3++) {
                        int i3 = P->ja[jj3];
                        if (P_marker[i3] < row_begin) P_marker[i3] = (int)cnt++;
                    }
                }
            }
        }
        if (cnt > 0x7fffffffL) { free(R_i); free(R_j); free(R_a); free(P_marker); free(A_marker); return JXAMD_ERROR_ARG; }
    }
    int *C_i = (int *)malloc(sizeof(int) * ((size_t)nc + 1));
    int *C_j = (int *)malloc(sizeof(int) * (cnt > 0 ? cnt : 1));
    double *C_a = (double *)malloc(sizeof(double) * (cnt > 0 ? cnt : 1));
    if (!C_i || !C_j || !C_a) {
        free(R_i); free(R_j); free(R_a); free(P_marker); free(A_marker); free(C_i); free(C_j); free(C_a);
        return JXAMD_ERROR_MEMORY;
    }
    /* numeric pass */
    for (int c = 0; c < nc; c++) P_marker[c] = -1;
    for (int i = 0; i < n; i++) A_marker[i] = -1;
    int jc = 0;
    for (int ic = 0; ic < nc; ic++) {
        const int row_begin = jc;
        C_i[ic] = jc;
        P_marker[ic] = jc;
        C_a[jc] = 0.0;
        C_j[jc] = ic;
        jc++;
        for (int jj1 = R_i[ic]; jj1 < R_i[ic + 1]; jj1++) {
            const int i1 = R_j[jj1];
            const double r_entry = R_a[jj1];
            for (int jj2 = A->ia[i1]; jj2 < A->ia[i1 + 1]; jj2++) {
                const int i2 = A->ja[jj2];
                const double r_a = r_entry * A->a[jj2];
                if (A_marker[i2] != ic) {
                    A_marker[i2] = ic;
                    for (int jj3 = P->ia[i2]; jj3 < P->ia[i2 + 1]; jj3++) {
                        const int i3 = P->ja[jj3];
                        const double rap = r_a * P->a[jj3];
                        if (P_marker[i3] < row_begin) {
                            P_marker[i3] = jc;
                            C_a[jc] = rap;
                            C_j[jc] = i3;
                            jc++;
                        } else {
                            C_a[P_marker[i3]] += rap;
                        }
                    }
                } else {
                    for (int jj3 = P->ia[i2]; jj3 < P->ia[i2 + 1]; jj3++)
                        C_a[P_marker[P->ja[jj3]]] += r_a * P->a[jj3];
                }
            }
        }
    }
    C_i[nc] = jc;
    free(R_i); free(R_j); free(R_a); free(P_marker); free(A_marker);
    C->n = nc;
    C->ncols = nc;
    C->ia = C_i;
    C->ja = C_j;
    C->a = C_a;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* level loop (amg/par_amg_setup.c:525-1320 with coarse_threshold_2 = -1)   */
/* ------------------------------------------------------------------------ */
void jxamd_amg_params_default(jxamd_amg_params *p)
{
    p->max_levels = 25;
    p->coarse_threshold = 1000;
    p->strong_threshold = 0.25;
    p->max_row_sum = 0.9;
}

int jxamd_amg_setup(const jxamd_csr *A0, const jxamd_amg_params *prm, jxamd_hierarchy *H)
{
    if (!A0 || !prm || !H || A0->n < 1 || A0->ncols != A0->n) return JXAMD_ERROR_ARG;
    if (prm->max_levels < 1 || prm->max_levels > JXAMD_MAX_LEVELS) return JXAMD_ERROR_ARG;
    memset(H, 0, sizeof(*H));
    mat lev[JXAMD_MAX_LEVELS], P[JXAMD_MAX_LEVELS];
    memset(lev, 0, sizeof(lev));
    memset(P, 0, sizeof(P));
    const int n0 = A0->n, nnz0 = A0->ia[n0];
    for (int i = 0; i < n0; i++)
        if (A0->ia[i + 1] <= A0->ia[i] || A0->ja[A0->ia[i]] != i) return JXAMD_ERROR_ARG; /* diagonal first */
    lev[0].n = lev[0].ncols = n0;
    lev[0].ia = (int *)malloc(sizeof(int) * ((size_t)n0 + 1));
    lev[0].ja = (int *)malloc(sizeof(int) * (nnz0 > 0 ? nnz0 : 1));
    lev[0].a = (double *)malloc(sizeof(double) * (nnz0 > 0 ? nnz0 : 1));
    if (!lev[0].ia || !lev[0].ja || !lev[0].a) { mat_free(&lev[0]); return JXAMD_ERROR_MEMORY; }
    memcpy(lev[0].ia, A0->ia, sizeof(int) * ((size_t)n0 + 1));
    memcpy(lev[0].ja, A0->ja, sizeof(int) * nnz0);
    memcpy(lev[0].a, A0->a, sizeof(double) * nnz0);

    int level = 0, e = 0;
    for (;;) {
        const mat *A = &lev[level];
        const int n = A->n;
        if (prm->max_levels == 1) break;
        int *S_i = NULL, *S_j = NULL;
        int *CF = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
        if (!CF) { e = JXAMD_ERROR_MEMORY; break; }
        e = create_strength(A, prm->strong_threshold, prm->max_row_sum, &S_i, &S_j);
        if (!e) e = ruge_first_pass(n, S_i, S_j, CF);
        if (!e) e = pmis(n, S_i, S_j, CF);
        int coarse_size = 0;
        for (int i = 0; i < n && !e; i++)
            if (CF[i] == 1) coarse_size++;
        if (e || coarse_size == 0 || coarse_size == n) {
            free(CF); free(S_i); free(S_j);
            break;
        }
        e = build_ext_pi(A, CF, S_i, S_j, &P[level]);
        free(CF); free(S_i); free(S_j);
        if (e) break;
        e = build_rap(A, &P[level], &lev[level + 1]);
        if (e) break;
        level++;
        if (level == prm->max_levels - 1 || coarse_size <= prm->coarse_threshold) break;
    }
    if (e) {
        for (int l = 0; l < JXAMD_MAX_LEVELS; l++) { mat_free(&lev[l]); mat_free(&P[l]); }
        return e;
    }
    H->num_levels = level + 1;
    for (int l = 0; l <= level; l++) {
        H->A[l].n = lev[l].n;
        H->A[l].ncols = lev[l].ncols;
        H->A[l].nnz = lev[l].ia[lev[l].n];
        H->A[l].ia = lev[l].ia;
        H->A[l].ja = lev[l].ja;
        H->A[l].a = lev[l].a;
        if (l < level) {
            H->P[l].n = P[l].n;
            H->P[l].ncols = P[l].ncols;
            H->P[l].nnz = P[l].ia[P[l].n];
            H->P[l].ia = P[l].ia;
            H->P[l].ja = P[l].ja;
            H->P[l].a = P[l].a;
        }
    }
    return 0;
}

void jxamd_hierarchy_free(jxamd_hierarchy *H)
{
    if (!H) return;
    for (int l = 0; l < JXAMD_MAX_LEVELS; l++) {
        free((void *)H->A[l].ia); free((void *)H->A[l].ja); free((void *)H->A[l].a);
        free((void *)H->P[l].ia); free((void *)H->P[l].ja); free((void *)H->P[l].a);
    }
    memset(H, 0, sizeof(*H));
}
