/*
 * drive.c -- TEST INFRASTRUCTURE: builds solvers for a generated system under
 * every layout / precision / stage-2 option with the host-memory device stub,
 * so AddressSanitizer checks the host layout code.  Exit status 0 = all
 * creations and destructions succeeded.
 *   drive nx ny nz b
 */
#include <stdio.h>
#include <stdlib.h>

#include "jxamd.h"

int main(int argc, char **argv)
{
    if (argc < 5) return 2;
    jxamd_gen_params g = {atoi(argv[1]), atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), JXAMD_FIELD_LAYERED, 1.0, 3.0,
                          2, 100.0, 0.1, 2, 1e3, 7};
    int nb, nnzb;
    if (jxamd_gen_sizes(&g, &nb, &nnzb)) return 3;
    const int b = g.block_size, bb = b * b;
    int *ia = malloc(sizeof(int) * (nb + 1)), *ja = malloc(sizeof(int) * nnzb);
    double *a = malloc(sizeof(double) * (size_t)nnzb * bb), *rhs = malloc(sizeof(double) * nb * b);
    double *x = calloc((size_t)nb * b, sizeof(double));
    if (jxamd_gen_reservoir(&g, ia, ja, a, rhs)) return 4;
    if (jxamd_bsr_decouple_abf(nb, b, ia, ja, a, rhs)) return 5;
    int *pia = malloc(sizeof(int) * (nb + 1)), *pja = malloc(sizeof(int) * nnzb);
    double *pa = malloc(sizeof(double) * nnzb);
    if (jxamd_bsr_extract_pressure(nb, b, ia, ja, a, 0, pia, pja, pa)) return 6;
    jxamd_csr A0 = {nb, nb, nnzb, pia, pja, pa};
    jxamd_amg_params ap;
    jxamd_amg_params_default(&ap);
    jxamd_hierarchy H;
    if (jxamd_amg_setup(&A0, &ap, &H)) return 7;
    jxamd_bsr A = {nb, nb, nnzb, b, ia, ja, a};
    const int chunks[] = {0, 1, 3, 16, 40, 0, 0};
    int fails = 0, runs = 0;
    for (int ci = 0; ci < 7; ci++)
        for (int prec = 0; prec <= 2; prec++)
            for (int s2 = 0; s2 < 4; s2++) {
                jxamd_params p;
                jxamd_params_default(&p);
                p.relax_chunk = chunks[ci];
                p.precond_precision = prec;
                p.stage2_solver_type = s2 < 3 ? 2 : 5;
                p.stage2_maxit = s2 < 3 ? s2 + 1 : 1;
                if (ci == 4) { p.relax_chunk = 0; p.relax_threads = 4; }
                if (ci == 5) p.poly_degree = 4;
                if (ci == 6) p.poly_degree = 1;
                if (ci == 5 && s2 == 1) p.poly_degree = 4;
                if (ci == 1 || ci == 3) p.stage2_chunk = ci == 1 ? 1 : 5;
                p.solution_update = (prec + s2) % 2; /* both end-of-cycle updates */
                p.stage2_precision = (ci + s2) % 3;  /* FP64 / FP32 / FP16 stage-2 blocks or BILU factors */
                p.arnoldi_kernel = (ci + prec) % 3;  /* cooperative, per projection, ordinary one-launch */
                if (ci == 2) p.coarse_rows = 64;     /* V-cycle cut at the first level of <= 64 rows */
                jxamd_solver *s = NULL;
                int rc = jxamd_solver_create(&s, &A, &H, &p);
                jxamd_stats st;
                if (!rc && p.poly_degree > 0) {
                    double w[8], lam;
                    rc = jxamd_solver_get_poly(s, 0, p.poly_degree, w, &lam);
                    if (!rc) rc = jxamd_solver_set_poly(s, 0, p.poly_degree, w);
                    if (!rc && p.poly_degree == 4) {
                        float feat[JXAMD_MLP_NFEAT], params[JXAMD_MLP_NPARAM] = {0};
                        int used;
                        double fl, fd;
                        rc = jxamd_solver_poly_features(s, 0, feat);
                        if (!rc) rc = jxamd_solver_apply_mlp(s, params, &used, &fl, &fd);
                    }
                }
                if (!rc) rc = jxamd_solver_solve(s, rhs, x, NULL, &st);
                if (s) jxamd_solver_destroy(s);
                runs++;
                if (rc) {
                    fails++;
                    fprintf(stderr, "chunk %d prec %d stage2 %d: rc %d\n", chunks[ci], prec, s2, rc);
                }
            }
    /* pinned words that never arrive (the norms, then an Arnoldi column):
     * each solve ends with JXAMD_ERROR_GENERIC after its bounded wait, and
     * the next solve on the same solver succeeds */
    {
        extern int jxamd_stub_faults;
        const int faults[2] = {1, 2 | 4};
        for (int f = 0; f < 2; f++) {
            jxamd_params p;
            jxamd_params_default(&p);
            p.max_iter = 3;
            jxamd_solver *s = NULL;
            jxamd_stats st;
            int rc = jxamd_solver_create(&s, &A, &H, &p);
            if (!rc) {
                jxamd_stub_faults = faults[f];
                rc = jxamd_solver_solve(s, rhs, x, NULL, &st);
                jxamd_stub_faults = 0;
                runs++;
                if (rc != JXAMD_ERROR_GENERIC) {
                    fails++;
                    fprintf(stderr, "missing pinned word (faults %d): rc %d, expected %d\n", faults[f], rc,
                            JXAMD_ERROR_GENERIC);
                }
                rc = jxamd_solver_solve(s, rhs, x, NULL, &st);
            }
            runs++;
            if (rc) {
                fails++;
                fprintf(stderr, "solve after a missing pinned word (faults %d): rc %d\n", faults[f], rc);
            }
            if (s) jxamd_solver_destroy(s);
        }
    }
    jxamd_hierarchy_free(&H);
    free(ia); free(ja); free(a); free(rhs); free(x); free(pia); free(pja); free(pa);
    printf("levels %d runs %d fails %d\n", H.num_levels, runs, fails);
    return fails ? 1 : 0;
}
