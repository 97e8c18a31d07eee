/*
 * ref_driver.c -- TEST INFRASTRUCTURE (oracle side).  Links the reference
 * library built from /root/reference sources in FP64 (oracle/Makefile) and
 * replays the call sequence of src/test_bsr_cprgmres.c:108-459:
 *   read BSR + RHS -> ABF decouple (rank 0) -> BSRMatrixToParBSRMatrix ->
 *   CPR create/set (pressure_index 0, stage1_maxit 1, stage2_maxit S,
 *   stage1 AMG, stage2 BGS) -> CPRSetup -> ParBSR GMRES(k_dim 30,
 *   is_check_restarted 1, logging 1, print_level 1) -> initial residual ->
 *   GMRESSolve -> stats.
 * Differences from that driver: rtol, max_iter and stage2_maxit are
 * arguments (the reference edits them into the source, test-hybrid.sh:53-61),
 * and intermediate objects are dumped as golden vectors.  Nothing in the
 * product links or runs this program.
 *
 * usage: ref_driver A.bin b.txt out.dump rtol max_iter stage2_maxit decouple dump_level [k_dim [is_thermal]]
 *   decouple: the JXF_DecoupType value (0 none, 1 ABF as the reference
 *   driver, 2 ANL, 3 SEM, 4 QI, 5 True-IMPES, 6 True-IMPES 2).
 *   dump_level 0: scalars + residual history only; 1: + hierarchy and
 *   single-operator vectors; 2: + decoupled system; 3: the decoupled system
 *   only (no solve).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mpi.h"

#include "jxf_mv.h"
#include "jxf_bsr_mv.h"
#include "jxf_parbsr_mv.h"
#include "jxf_cpr.h"
#include "jxf_bsr_decoup.h"
#include "jxf_pamg.h"
#include "jxf_krylov.h"

static FILE *g_out;

/* record: int32 name_len, name, int32 kind (0 int32, 1 float64), int64 count, payload */
static void dump_raw(const char *name, int kind, long count, const void *data)
{
    int len = (int)strlen(name);
    fwrite(&len, sizeof(int), 1, g_out);
    fwrite(name, 1, (size_t)len, g_out);
    fwrite(&kind, sizeof(int), 1, g_out);
    long long c = count;
    fwrite(&c, sizeof(long long), 1, g_out);
    if (count > 0) fwrite(data, kind == 0 ? sizeof(int) : sizeof(double), (size_t)count, g_out);
}
static void dump_i(const char *name, long n, const int *v) { dump_raw(name, 0, n, v); }
static void dump_d(const char *name, long n, const double *v) { dump_raw(name, 1, n, v); }
static void dump_scalar_d(const char *name, double v) { dump_raw(name, 1, 1, &v); }
static void dump_scalar_i(const char *name, int v) { dump_raw(name, 0, 1, &v); }

static void dump_csr(const char *pfx, jxf_CSRMatrix *A)
{
    char nm[128];
    int n = jxf_CSRMatrixNumRows(A), m = jxf_CSRMatrixNumCols(A);
    int nnz = jxf_CSRMatrixI(A)[n];
    int shape[3] = {n, m, nnz};
    snprintf(nm, sizeof nm, "%s_shape", pfx); dump_i(nm, 3, shape);
    snprintf(nm, sizeof nm, "%s_ia", pfx); dump_i(nm, n + 1, jxf_CSRMatrixI(A));
    snprintf(nm, sizeof nm, "%s_ja", pfx); dump_i(nm, nnz, jxf_CSRMatrixJ(A));
    snprintf(nm, sizeof nm, "%s_a", pfx); dump_d(nm, nnz, jxf_CSRMatrixData(A));
}

/* deterministic test vector in (-1,1) */
static void fill_vec(double *v, int n, unsigned long long seed)
{
    unsigned long long s = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    for (int i = 0; i < n; i++) {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        v[i] = ((double)(s >> 11) / 9007199254740992.0) * 2.0 - 1.0;
    }
}

static jxf_ParVector *pvec_like_csr(jxf_ParCSRMatrix *A)
{
    jxf_ParVector *v = jxf_ParVectorCreate(jxf_ParCSRMatrixComm(A), jxf_ParCSRMatrixGlobalNumRows(A),
                                           jxf_ParCSRMatrixRowStarts(A));
    jxf_ParVectorInitialize(v);
    jxf_ParVectorSetPartitioningOwner(v, 0);
    return v;
}

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    if (argc < 9) {
        fprintf(stderr, "usage: %s A.bin b.txt out.dump rtol max_iter stage2_maxit decouple dump_level\n", argv[0]);
        MPI_Finalize();
        return 1;
    }
    const char *afile = argv[1], *bfile = argv[2];
    double rtol = atof(argv[4]);
    int max_iter = atoi(argv[5]);
    int stage2_maxit = atoi(argv[6]);
    int decouple = atoi(argv[7]);
    int dump_level = atoi(argv[8]);
    int k_dim = argc > 9 ? atoi(argv[9]) : 30;
    int is_thermal = argc > 10 ? atoi(argv[10]) : 0;
    g_out = fopen(argv[3], "wb");
    if (!g_out) { fprintf(stderr, "cannot open %s\n", argv[3]); return 1; }

    /* 1. read (src/test_bsr_cprgmres.c:127-190) */
    jxf_BSRMatrix *A_bsr = jxf_BSRMatrixRead_Binary(afile);
    int b = jxf_BSRMatrixBlockSize(A_bsr);
    int nb = jxf_BSRMatrixNumRows(A_bsr);
    int N = nb * b;
    jxf_Vector *b_Ser = jxf_SeqVectorCreate(N);
    jxf_SeqVectorInitialize(b_Ser);
    {
        FILE *f = fopen(bfile, "r");
        int n;
        if (!f || fscanf(f, "%d", &n) != 1 || n != N) { fprintf(stderr, "bad rhs\n"); return 1; }
        for (int i = 0; i < N; i++)
            if (fscanf(f, "%lf", &jxf_VectorData(b_Ser)[i]) != 1) { fprintf(stderr, "bad rhs\n"); return 1; }
        fclose(f);
    }
    jxf_Vector *x_Ser = jxf_SeqVectorCreate(N);
    jxf_SeqVectorInitialize(x_Ser);
    for (int i = 0; i < N; i++) jxf_VectorData(x_Ser)[i] = 0.0;

    /* 2. decoupling on the serial system (src/test_bsr_cprgmres.c:193-212 uses ABF) */
    double t0 = MPI_Wtime();
    if (decouple) jxf_BSRMatrixDecouple(A_bsr, b_Ser, (JXF_DecoupType)decouple, is_thermal);
    double t_decouple = MPI_Wtime() - t0;
    if (dump_level >= 2) {
        int nnzb = jxf_BSRMatrixNumNonzeros(A_bsr);
        dump_d("dec_a", (long)nnzb * b * b, jxf_BSRMatrixData(A_bsr));
        dump_d("dec_b", N, jxf_VectorData(b_Ser));
    }
    if (dump_level == 3) {
        dump_scalar_d("t_decouple", t_decouple);
        fclose(g_out);
        MPI_Finalize();
        return 0;
    }

    /* 3. distribute (src/test_bsr_cprgmres.c:248-258) */
    JXF_BigInt *rs = NULL, *cs = NULL;
    jxf_GeneratePartitioning(jxf_BSRMatrixNumRows(A_bsr), 1, &rs);
    jxf_GeneratePartitioning(jxf_BSRMatrixNumCols(A_bsr), 1, &cs);
    jxf_ParBSRMatrix *A = jxf_BSRMatrixToParBSRMatrix(MPI_COMM_WORLD, A_bsr, rs, cs);
    JXF_Int *part;
    jxf_ParBSRMatrixGetRowPartitioning(A, &part);
    jxf_ParVector *par_rhs = jxf_VectorToParVector(jxf_ParBSRMatrixComm(A), b_Ser, part);
    jxf_ParVector *par_sol = jxf_VectorToParVector(jxf_ParBSRMatrixComm(A), x_Ser, part);

    /* 4. CPR (src/test_bsr_cprgmres.c:314-334) */
    jxf_CPRPrecond *cpr = JXF_CPRCreate(MPI_COMM_WORLD);
    JXF_CPRSetParameter(cpr, "pressure_index", 0);
    JXF_CPRSetParameter(cpr, "block_size", b);
    JXF_CPRSetParameter(cpr, "stage1_maxit", 1);
    JXF_CPRSetParameter(cpr, "stage2_maxit", stage2_maxit);
    JXF_CPRSetParameter(cpr, "stage1_solver_type", 1);
    JXF_CPRSetParameter(cpr, "stage2_solver_type", 2);
    JXF_CPRSetParameter(cpr, "print_level", 1);
    JXF_CPRSetRealParameter(cpr, "threshold", 1e-15);
    t0 = MPI_Wtime();
    JXF_CPRSetup(cpr, A);
    double t_setup = MPI_Wtime() - t0;

    jxf_ParAMGData *amg = (jxf_ParAMGData *)cpr->stage1_solver;
    int L = jxf_ParAMGDataNumLevels(amg);
    jxf_ParCSRMatrix **Aarr = jxf_ParAMGDataAArray(amg);
    jxf_ParCSRMatrix **Parr = jxf_ParAMGDataPArray(amg);
    dump_scalar_i("num_levels", L);
    {
        int sizes[64];
        for (int l = 0; l < L && l < 64; l++) sizes[l] = jxf_CSRMatrixNumRows(jxf_ParCSRMatrixDiag(Aarr[l]));
        dump_i("level_sizes", L, sizes);
    }
    if (dump_level >= 1) {
        char nm[64];
        for (int l = 0; l < L; l++) {
            snprintf(nm, sizeof nm, "A%d", l);
            dump_csr(nm, jxf_ParCSRMatrixDiag(Aarr[l]));
            if (jxf_CSRMatrixNumCols(jxf_ParCSRMatrixOffd(Aarr[l])) != 0) fprintf(stderr, "offd nonempty\n");
            if (l < L - 1) {
                snprintf(nm, sizeof nm, "P%d", l);
                dump_csr(nm, jxf_ParCSRMatrixDiag(Parr[l]));
                int *cf = jxf_ParAMGDataCFMarkerArray(amg)[l];
                snprintf(nm, sizeof nm, "CF%d", l);
                if (cf) dump_i(nm, jxf_CSRMatrixNumRows(jxf_ParCSRMatrixDiag(Aarr[l])), cf);
            }
        }

        /* single-operator golden vectors */
        jxf_ParVector *Vt0 = pvec_like_csr(Aarr[0]);
        jxf_ParVector *f0 = pvec_like_csr(Aarr[0]);
        jxf_ParVector *u0 = pvec_like_csr(Aarr[0]);
        int n0 = jxf_ParVectorGlobalSize(f0);
        fill_vec(jxf_VectorData(jxf_ParVectorLocalVector(f0)), n0, 11);
        fill_vec(jxf_VectorData(jxf_ParVectorLocalVector(u0)), n0, 12);
        dump_d("hsgs_f", n0, jxf_VectorData(jxf_ParVectorLocalVector(f0)));
        dump_d("hsgs_u_in", n0, jxf_VectorData(jxf_ParVectorLocalVector(u0)));
        jxf_PAMGRelax6(Aarr[0], f0, NULL, 0, 1.0, 1.0, u0, Vt0);
        dump_d("hsgs_u_out", n0, jxf_VectorData(jxf_ParVectorLocalVector(u0)));

        /* coarsest direct solve */
        jxf_ParVector *fc = pvec_like_csr(Aarr[L - 1]);
        jxf_ParVector *uc = pvec_like_csr(Aarr[L - 1]);
        jxf_ParVector *Vtc = pvec_like_csr(Aarr[L - 1]);
        int nc = jxf_ParVectorGlobalSize(fc);
        fill_vec(jxf_VectorData(jxf_ParVectorLocalVector(fc)), nc, 13);
        jxf_ParVectorSetConstantValues(uc, 0.0);
        dump_d("coarse_f", nc, jxf_VectorData(jxf_ParVectorLocalVector(fc)));
        jxf_PAMGRelax9(Aarr[L - 1], fc, NULL, 0, 1.0, 1.0, uc, Vtc);
        dump_d("coarse_u", nc, jxf_VectorData(jxf_ParVectorLocalVector(uc)));

        /* one V-cycle from zero (the stage-1 solve of src/../jxf_cpr.c:470-476) */
        jxf_ParVector *fv = pvec_like_csr(Aarr[0]);
        jxf_ParVector *uv = pvec_like_csr(Aarr[0]);
        fill_vec(jxf_VectorData(jxf_ParVectorLocalVector(fv)), n0, 14);
        jxf_ParVectorSetConstantValues(uv, 0.0);
        dump_d("vcycle_f", n0, jxf_VectorData(jxf_ParVectorLocalVector(fv)));
        JXF_PAMGSolve(cpr->stage1_solver, (JXF_ParCSRMatrix)Aarr[0], (JXF_Vector)fv, (JXF_ParVector)uv);
        dump_d("vcycle_u", n0, jxf_VectorData(jxf_ParVectorLocalVector(uv)));

        /* one stage-2 block HGS sweep (w 0.7, omega 0.9, forward) */
        jxf_ParVector *bx = jxf_CreateGlobalVector(A);
        jxf_ParVector *bb = jxf_CreateGlobalVector(A);
        jxf_ParVector *bw = jxf_CreateGlobalVector(A);
        fill_vec(jxf_VectorData(jxf_ParVectorLocalVector(bx)), N, 15);
        fill_vec(jxf_VectorData(jxf_ParVectorLocalVector(bb)), N, 16);
        dump_d("bhgs_x_in", N, jxf_VectorData(jxf_ParVectorLocalVector(bx)));
        dump_d("bhgs_b", N, jxf_VectorData(jxf_ParVectorLocalVector(bb)));
        jxf_ParBSRHGSRelax(A, bx, bb, bw, 0.7, 0.9, 1);
        dump_d("bhgs_x_out", N, jxf_VectorData(jxf_ParVectorLocalVector(bx)));

        /* one full CPR application x = M^{-1} r, r = decoupled rhs / ||.|| */
        jxf_ParVector *cr = jxf_CreateGlobalVector(A);
        jxf_ParVector *cx = jxf_CreateGlobalVector(A);
        jxf_ParVectorCopy(par_rhs, cr);
        double nr = jxf_ParVectorNorm2(cr);
        jxf_ParVectorScale(1.0 / nr, cr);
        jxf_ParVectorSetConstantValues(cx, 0.0);
        dump_d("cpr_r", N, jxf_VectorData(jxf_ParVectorLocalVector(cr)));
        JXF_CPRPrecond(cpr, A, cr, cx);
        dump_d("cpr_x", N, jxf_VectorData(jxf_ParVectorLocalVector(cx)));
        /* the apply accumulates stage timers; reset for the solve below */
        cpr->stage1_solve_time = 0.0;
        cpr->stage2_solve_time = 0.0;
    }

    /* 5. GMRES (src/test_bsr_cprgmres.c:346-392) */
    JXF_Solver gmres;
    JXF_ParBSRGMRESCreate(MPI_COMM_WORLD, &gmres);
    JXF_GMRESSetKDim(gmres, k_dim);
    JXF_GMRESSetIsCheckRestarted(gmres, 1);
    JXF_GMRESSetMaxIter(gmres, max_iter);
    JXF_GMRESSetTol(gmres, rtol);
    JXF_GMRESSetLogging(gmres, 1);
    JXF_GMRESSetPrintLevel(gmres, 1);
    JXF_GMRESSetPrecond(gmres, (JXF_PtrToSolverFcn)JXF_CPRPrecond, (JXF_PtrToSolverFcn)JXF_CPRSetup,
                        (JXF_Solver)cpr);
    JXF_GMRESSetup(gmres, (JXF_Matrix)A, (JXF_Vector)par_rhs, (JXF_Vector)par_sol);

    jxf_ParVector *res = jxf_CreateGlobalVector(A);
    jxf_ParVectorCopy(par_rhs, res);
    jxf_ParBSRMatrixMatvec(-1.0, A, par_sol, 1.0, res);
    double r0 = jxf_ParVectorNorm2(res);

    t0 = MPI_Wtime();
    JXF_GMRESSolve(gmres, (JXF_Matrix)A, (JXF_Matrix)A, (JXF_Vector)par_rhs, (JXF_Vector)par_sol);
    double t_solve = MPI_Wtime() - t0;

    jxf_ParVectorCopy(par_rhs, res);
    jxf_ParBSRMatrixMatvec(-1.0, A, par_sol, 1.0, res);
    double rfin = jxf_ParVectorNorm2(res);
    double bnorm = jxf_ParVectorNorm2(par_rhs);
    JXF_Int iters = 0;
    JXF_Real relres = 0.0;
    JXF_GMRESGetNumIterations(gmres, &iters);
    JXF_GMRESGetFinalRelativeResidualNorm(gmres, &relres);
    jxf_GMRESData *gd = (jxf_GMRESData *)gmres;
    JXF_Real stats[4];
    JXF_CPRGetStatistics(cpr, stats);

    dump_scalar_i("iterations", iters);
    dump_scalar_d("final_rel_res", relres);
    dump_scalar_d("true_res", rfin);
    dump_scalar_d("b_norm", bnorm);
    dump_scalar_d("r0_norm", r0);
    dump_scalar_d("t_solve", t_solve);
    dump_scalar_d("t_setup", t_setup);
    dump_scalar_d("t_decouple", t_decouple);
    dump_scalar_d("t_stage1", stats[1]);
    dump_scalar_d("t_stage2", stats[3]);
    dump_scalar_i("converged", gd->converged);
    if (gd->norms) dump_d("res_history", iters + 1, gd->norms);
    if (dump_level >= 1) dump_d("x_final", N, jxf_VectorData(jxf_ParVectorLocalVector(par_sol)));
    fclose(g_out);

    fprintf(stderr, "ORACLE iters=%d rel=%.6e true_rel=%.6e solve=%.3fs setup=%.3fs levels=%d\n", iters, relres,
            rfin / bnorm, t_solve, t_setup, L);
    MPI_Finalize();
    return 0;
}
