/*
 * jxamd_interface.c -- the CSR-array solve entry with the arguments of
 * JXFPAMG_Solver_Interface (src/jxpamg_interface.c:118-257) for
 * solver_id 220, CPR-GMRES (src/jxpamg_interface.c:1967-2043), on one GPU.
 *
 * What it keeps from the reference entry:
 *   - a 1-based ia/ja is decremented in place (jxpamg_interface.c:143-150);
 *   - every row's diagonal entry is swapped to the front of the row in place,
 *     jxf_CSRMatrixReorder (operation/matops/seq_csr_matop.c:189-231);
 *   - the solve starts from x = 0 and copies the solution into ser_x
 *     (jxpamg_interface.c:205-208, 233-238); *iter = GMRES iterations;
 *   - case 220 settings: k_dim 30, max_iter 1000, is_check_restarted 1
 *     (jxpamg_interface.c:582-586), CPR with pressure_index 0, one AMG
 *     V-cycle and one forward block hybrid-SOR sweep (w 0.7, omega 0.9,
 *     jxf_cpr.c:547-549), no decoupling;
 *   - argv: -kdim and -maxiter as in the reference parser; other reference
 *     options are skipped with their value.
 * What it adds: the reference passes its ParCSR matrix to the BSR-only CPR
 * (which does not compile as shipped); here num_functions = b regroups the
 * node-major CSR (b consecutive rows per cell, the row_part logic of
 * jxpamg_interface.c:177-199) into BSR with the diagonal block first and the
 * other blocks by ascending block column, and the pressure-AMG hierarchy is
 * built on the host (jxamd_amg_setup).  Extra argv options (no reference
 * counterpart): -rchunk c / -rthreads T select the chunked relax-6 smoother.
 */
#include <stdlib.h>
#include <string.h>

#include "jxamd.h"

static int reorder_diag_first(int n, const int *ia, int *ja, double *a)
{
    for (int i = 0; i < n; i++) {
        int k0 = ia[i], k1 = ia[i + 1], found = 0;
        for (int k = k0; k < k1; k++) {
            if (ja[k] == i) {
                int tj = ja[k0];
                ja[k0] = ja[k];
                ja[k] = tj;
                double ta = a[k0];
                a[k0] = a[k];
                a[k] = ta;
                found = 1;
                break;
            }
        }
        if (!found) return JXAMD_ERROR_ARG; /* the reference returns -2: diagonal missing */
    }
    return 0;
}

/* node-major CSR (n = nb*b rows) -> BSR, diagonal block first, then ascending
 * block columns; entries missing from a block are zero */
static int csr_to_bsr(int n, int b, const int *ia, const int *ja, const double *a, int **oia, int **oja,
                      double **oa, int *onnzb)
{
    const int nb = n / b, bb = b * b;
    int *bia = (int *)malloc(sizeof(int) * ((size_t)nb + 1));
    int *mark = (int *)malloc(sizeof(int) * (nb > 0 ? nb : 1));
    int *cols = (int *)malloc(sizeof(int) * (nb > 0 ? nb : 1));
    if (!bia || !mark || !cols) { free(bia); free(mark); free(cols); return JXAMD_ERROR_MEMORY; }
    for (int J = 0; J < nb; J++) mark[J] = -1;
    long total = 0;
    bia[0] = 0;
    for (int I = 0; I < nb; I++) {
        int cnt = 0;
        for (int r = 0; r < b; r++)
            for (int k = ia[I * b + r]; k < ia[I * b + r + 1]; k++) {
                int J = ja[k] / b;
                if (mark[J] != I) { mark[J] = I; cnt++; }
            }
        if (mark[I] != I) cnt++; /* diagonal block always stored */
        total += cnt;
        if (total > 0x7fffffffL) { free(bia); free(mark); free(cols); return JXAMD_ERROR_ARG; }
        bia[I + 1] = (int)total;
    }
    int *bja = (int *)malloc(sizeof(int) * (size_t)(total > 0 ? total : 1));
    double *ba = (double *)calloc((size_t)(total > 0 ? total : 1) * bb, sizeof(double));
    int *slot = (int *)malloc(sizeof(int) * (nb > 0 ? nb : 1));
    if (!bja || !ba || !slot) {
        free(bia); free(mark); free(cols); free(bja); free(ba); free(slot);
        return JXAMD_ERROR_MEMORY;
    }
    for (int J = 0; J < nb; J++) mark[J] = -1;
    for (int I = 0; I < nb; I++) {
        int cnt = 0;
        mark[I] = I;
        for (int r = 0; r < b; r++)
            for (int k = ia[I * b + r]; k < ia[I * b + r + 1]; k++) {
                int J = ja[k] / b;
                if (mark[J] != I) { mark[J] = I; cols[cnt++] = J; }
            }
        /* ascending block columns after the diagonal (insertion sort: rows are short) */
        for (int x = 1; x < cnt; x++) {
            int v = cols[x], y = x - 1;
            while (y >= 0 && cols[y] > v) { cols[y + 1] = cols[y]; y--; }
            cols[y + 1] = v;
        }
        const int base = bia[I];
        bja[base] = I;
        slot[I] = base;
        for (int x = 0; x < cnt; x++) {
            bja[base + 1 + x] = cols[x];
            slot[cols[x]] = base + 1 + x;
        }
        for (int r = 0; r < b; r++)
            for (int k = ia[I * b + r]; k < ia[I * b + r + 1]; k++) {
                int J = ja[k] / b, c = ja[k] % b;
                ba[(size_t)slot[J] * bb + r * b + c] += a[k];
            }
    }
    free(mark);
    free(cols);
    free(slot);
    *oia = bia;
    *oja = bja;
    *oa = ba;
    *onnzb = (int)total;
    return 0;
}

int jxamd_solver_interface(int argc, char *argv[], int row, int *ia, int *ja, double *a, double *ser_x,
                           double *ser_b, double tol, int solver_id, int num_functions, int *iter)
{
    if (!ia || !ja || !a || !ser_x || !ser_b || !iter || row < 1) return JXAMD_ERROR_ARG;
    if (solver_id != 220) return JXAMD_ERROR_ARG;
    const int b = num_functions;
    if (b < 2 || b > JXAMD_MAX_BLOCK || row % b != 0) return JXAMD_ERROR_ARG;

    jxamd_params prm;
    jxamd_params_default(&prm);
    prm.k_dim = 30;
    prm.max_iter = 1000;
    prm.is_check_restarted = 1;
    prm.tol = tol;
    prm.pressure_index = 0;
    prm.stage1_maxit = 1;
    prm.stage2_maxit = 1;
    prm.threshold = 1e-15;
    for (int i = 1; i < argc; i++) {
        if (!argv[i]) continue;
        if (!strcmp(argv[i], "-kdim") && i + 1 < argc) prm.k_dim = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-maxiter") && i + 1 < argc) prm.max_iter = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-rchunk") && i + 1 < argc) prm.relax_chunk = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-rthreads") && i + 1 < argc) prm.relax_threads = atoi(argv[++i]);
    }

    if (ia[0] == 1) {
        const int nnz = ia[row] - 1;
        for (int i = 0; i <= row; i++) ia[i]--;
        for (int k = 0; k < nnz; k++) ja[k]--;
    }
    if (ia[0] != 0) return JXAMD_ERROR_ARG;
    for (int i = 0; i < row; i++)
        if (ia[i + 1] < ia[i]) return JXAMD_ERROR_ARG;
    for (int k = 0; k < ia[row]; k++)
        if (ja[k] < 0 || ja[k] >= row) return JXAMD_ERROR_ARG;
    int e = reorder_diag_first(row, ia, ja, a);
    if (e) return e;

    int *bia = NULL, *bja = NULL, *pia = NULL, *pja = NULL, nnzb = 0;
    double *ba = NULL, *pa = NULL;
    jxamd_hierarchy H;
    memset(&H, 0, sizeof(H));
    jxamd_solver *s = NULL;
    const int nb = row / b;
    e = csr_to_bsr(row, b, ia, ja, a, &bia, &bja, &ba, &nnzb);
    if (!e) {
        pia = (int *)malloc(sizeof(int) * ((size_t)nb + 1));
        pja = (int *)malloc(sizeof(int) * (size_t)nnzb);
        pa = (double *)malloc(sizeof(double) * (size_t)nnzb);
        if (!pia || !pja || !pa) e = JXAMD_ERROR_MEMORY;
    }
    if (!e) e = jxamd_bsr_extract_pressure(nb, b, bia, bja, ba, prm.pressure_index, pia, pja, pa);
    if (!e) {
        jxamd_csr A0 = {nb, nb, nnzb, pia, pja, pa};
        jxamd_amg_params ap;
        jxamd_amg_params_default(&ap);
        e = jxamd_amg_setup(&A0, &ap, &H);
    }
    if (!e) {
        jxamd_bsr A = {nb, nb, nnzb, b, bia, bja, ba};
        e = jxamd_solver_create(&s, &A, &H, &prm);
    }
    if (!e) {
        jxamd_stats st;
        for (int i = 0; i < row; i++) ser_x[i] = 0.0;
        e = jxamd_solver_solve(s, ser_b, ser_x, NULL, &st);
        if (!e) *iter = st.iterations;
    }
    if (s) jxamd_solver_destroy(s);
    jxamd_hierarchy_free(&H);
    free(bia); free(bja); free(ba); free(pia); free(pja); free(pa);
    return e;
}
