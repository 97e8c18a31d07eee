/*
 * Reference on-disk formats for a BSR system.
 *
 * Binary BSR (jxf_BSRMatrixRead_Binary / jxf_BSRMatrixPrint_Binary,
 * jxfpamg-0220/csrc/mat_vec_data/seq_bsr_matrix.c:1935-2019 and :2038+):
 *   int32 ROW, COL, NNZ, block_size, storage_manner
 *   int32 ROW+1, IA[ROW+1]
 *   int32 NNZ,   JA[NNZ]
 *   int32 NNZ*b*b, VAL[NNZ*b*b]   (value width = sizeof(JXF_Real); 8 here)
 * A 1-based file is detected by IA[0] == 1.  storage_manner 1 means
 * column-major blocks; the reference's reader never reads the data in that
 * branch (seq_bsr_matrix.c:1998-2007), this reader transposes the blocks.
 *
 * RHS text (read_rhs_file, src/test_bsr_cprgmres.c:26-76): the length n,
 * then n values.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "jxamd.h"

int jxamd_bsr_write_binary(const char *path, int nb, int ncols, int nnzb, int b,
                           const int *ia, const int *ja, const double *a)
{
    FILE *fp = fopen(path, "wb");
    if (!fp) return JXAMD_ERROR_GENERIC;
    int hdr[5] = {nb, ncols, nnzb, b, 0};
    int n1 = nb + 1, n2 = nnzb, n3 = nnzb * b * b;
    int ok = fwrite(hdr, sizeof(int), 5, fp) == 5;
    ok = ok && fwrite(&n1, sizeof(int), 1, fp) == 1 && fwrite(ia, sizeof(int), (size_t)n1, fp) == (size_t)n1;
    ok = ok && fwrite(&n2, sizeof(int), 1, fp) == 1 && fwrite(ja, sizeof(int), (size_t)n2, fp) == (size_t)n2;
    ok = ok && fwrite(&n3, sizeof(int), 1, fp) == 1 && fwrite(a, sizeof(double), (size_t)n3, fp) == (size_t)n3;
    fclose(fp);
    return ok ? JXAMD_SUCCESS : JXAMD_ERROR_GENERIC;
}

/* Header only: fills nb, ncols, nnzb, b so the caller can allocate. */
int jxamd_bsr_read_binary_sizes(const char *path, int *nb, int *ncols, int *nnzb, int *b)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return JXAMD_ERROR_GENERIC;
    int hdr[5];
    int ok = fread(hdr, sizeof(int), 5, fp) == 5;
    fclose(fp);
    if (!ok || hdr[0] < 0 || hdr[2] < 0 || hdr[3] < 1) return JXAMD_ERROR_GENERIC;
    *nb = hdr[0];
    *ncols = hdr[1];
    *nnzb = hdr[2];
    *b = hdr[3];
    return JXAMD_SUCCESS;
}

int jxamd_bsr_read_binary(const char *path, int *ia, int *ja, double *a)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return JXAMD_ERROR_GENERIC;
    int hdr[5], n;
    int ok = fread(hdr, sizeof(int), 5, fp) == 5;
    const int nb = hdr[0], nnzb = hdr[2], b = hdr[3], manner = hdr[4];
    ok = ok && fread(&n, sizeof(int), 1, fp) == 1 && n == nb + 1;
    ok = ok && fread(ia, sizeof(int), (size_t)n, fp) == (size_t)n;
    ok = ok && fread(&n, sizeof(int), 1, fp) == 1 && n == nnzb;
    ok = ok && fread(ja, sizeof(int), (size_t)n, fp) == (size_t)n;
    ok = ok && fread(&n, sizeof(int), 1, fp) == 1 && n == nnzb * b * b;
    ok = ok && fread(a, sizeof(double), (size_t)n, fp) == (size_t)n;
    fclose(fp);
    if (!ok) return JXAMD_ERROR_GENERIC;
    if (ia[0] == 1) {
        for (int i = 0; i <= nb; i++) ia[i] -= 1;
        for (int j = 0; j < nnzb; j++) ja[j] -= 1;
    }
    if (manner == 1) {
        const int bb = b * b;
        double tmp[JXAMD_MAX_BLOCK * JXAMD_MAX_BLOCK];
        if (b > JXAMD_MAX_BLOCK) return JXAMD_ERROR_ARG;
        for (long k = 0; k < nnzb; k++) {
            double *blk = a + k * bb;
            for (int r = 0; r < b; r++)
                for (int s = 0; s < b; s++) tmp[r * b + s] = blk[s * b + r];
            memcpy(blk, tmp, sizeof(double) * bb);
        }
    }
    return JXAMD_SUCCESS;
}

int jxamd_rhs_write_text(const char *path, int n, const double *v)
{
    FILE *fp = fopen(path, "w");
    if (!fp) return JXAMD_ERROR_GENERIC;
    fprintf(fp, "%d\n", n);
    for (int i = 0; i < n; i++) fprintf(fp, "%.17g\n", v[i]);
    fclose(fp);
    return JXAMD_SUCCESS;
}

int jxamd_rhs_read_text(const char *path, int n, double *v)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return JXAMD_ERROR_GENERIC;
    int m;
    if (fscanf(fp, "%d", &m) != 1 || m != n) { fclose(fp); return JXAMD_ERROR_ARG; }
    for (int i = 0; i < n; i++) {
        if (fscanf(fp, "%lf", &v[i]) != 1) { fclose(fp); return JXAMD_ERROR_GENERIC; }
    }
    fclose(fp);
    return JXAMD_SUCCESS;
}
