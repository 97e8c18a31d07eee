/*
 * ref_bhgs_mpi.c -- TEST INFRASTRUCTURE (oracle side).  One forward stage-2
 * block hybrid SOR sweep of the reference library built from its sources
 * (jxf_ParBSRHGSRelax, jxfpamg-0220/csrc/amg/par_relax_bsr.c:193-573) on as
 * many MPI ranks as mpirun starts, with the weights JXF_CPRPrecond passes
 * (w 0.7, omega 0.9, jxfpamg-0220/csrc/cpr/jxf_cpr.c:547-549).  The matrix is
 * split into contiguous block-row ranges as the reference driver does
 * (src/test_bsr_cprgmres.c:248-258): rank 0 reads A (BSR binary), x0 and b
 * (the driver's text vector format: the length, then one value per line),
 * the sweep runs on the distributed matrix (Gauss-Seidel on each rank's
 * rows, the other ranks' values at entry), and rank 0 writes x in the same
 * text format.  Nothing in the product links or runs this program.
 *
 * usage: mpirun -np P ref_bhgs_mpi A.bin x0.txt b.txt out.txt
 */
#include <stdio.h>
#include <stdlib.h>

#include "mpi.h"

#include "jxf_mv.h"
#include "jxf_bsr_mv.h"
#include "jxf_parbsr_mv.h"
#include "jxf_pamg.h"
#include "jxf_cpr.h"

static jxf_Vector *read_vec(const char *path, int N)
{
    FILE *f = fopen(path, "r");
    int n = -1;
    if (!f || fscanf(f, "%d", &n) != 1 || n != N) {
        if (f) fclose(f);
        return NULL;
    }
    jxf_Vector *v = jxf_SeqVectorCreate(N);
    jxf_SeqVectorInitialize(v);
    for (int i = 0; i < N; i++)
        if (fscanf(f, "%lf", &jxf_VectorData(v)[i]) != 1) {
            fclose(f);
            return NULL;
        }
    fclose(f);
    return v;
}

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    int myid = 0, np = 1;
    MPI_Comm_rank(MPI_COMM_WORLD, &myid);
    MPI_Comm_size(MPI_COMM_WORLD, &np);
    if (argc < 5) {
        if (myid == 0) fprintf(stderr, "usage: %s A.bin x0.txt b.txt out.txt\n", argv[0]);
        MPI_Finalize();
        return 1;
    }
    jxf_BSRMatrix *A_bsr = NULL;
    jxf_Vector *x_Ser = NULL, *b_Ser = NULL;
    int ok = 1;
    JXF_BigInt *rs = NULL, *cs = NULL;
    if (myid == 0) {
        A_bsr = jxf_BSRMatrixRead_Binary(argv[1]);
        if (!A_bsr) ok = 0;
        if (ok) {
            const int N = jxf_BSRMatrixNumRows(A_bsr) * jxf_BSRMatrixBlockSize(A_bsr);
            x_Ser = read_vec(argv[2], N);
            b_Ser = read_vec(argv[3], N);
            if (!x_Ser || !b_Ser) ok = 0;
        }
        if (ok) {
            jxf_GeneratePartitioning(jxf_BSRMatrixNumRows(A_bsr), np, &rs);
            jxf_GeneratePartitioning(jxf_BSRMatrixNumCols(A_bsr), np, &cs);
        }
    }
    MPI_Bcast(&ok, 1, MPI_INT, 0, MPI_COMM_WORLD);
    if (!ok) {
        if (myid == 0) fprintf(stderr, "cannot read the inputs\n");
        MPI_Finalize();
        return 1;
    }
    jxf_ParBSRMatrix *A = jxf_BSRMatrixToParBSRMatrix(MPI_COMM_WORLD, A_bsr, rs, cs);
    JXF_Int *part;
    jxf_ParBSRMatrixGetRowPartitioning(A, &part);
    jxf_ParVector *px = jxf_VectorToParVector(jxf_ParBSRMatrixComm(A), x_Ser, part);
    jxf_ParVector *pb = jxf_VectorToParVector(jxf_ParBSRMatrixComm(A), b_Ser, part);
    jxf_ParVector *pw = jxf_CreateGlobalVector(A);
    jxf_ParBSRHGSRelax(A, px, pb, pw, 0.7, 0.9, 1);
    jxf_Vector *x_all = jxf_ParVectorToVectorAll(px);
    if (myid == 0) {
        FILE *f = fopen(argv[4], "w");
        if (!f) {
            fprintf(stderr, "cannot open %s\n", argv[4]);
            ok = 0;
        } else {
            const int N = jxf_VectorSize(x_all);
            fprintf(f, "%d\n", N);
            for (int i = 0; i < N; i++) fprintf(f, "%.17g\n", jxf_VectorData(x_all)[i]);
            fclose(f);
            fprintf(stderr, "ref_bhgs_mpi: %d ranks, %d values\n", np, N);
        }
    }
    MPI_Finalize();
    return ok ? 0 : 1;
}
