/*
 * jxamd.h -- C-ABI of the MI355X CPR-GMRES solve phase.
 *
 * Plain pointers and sizes only.  Every entry point returns an int error
 * flag with the reference's conventions:
 *   JXF_SUCCESS 0                      jxfpamg-0220/include/jxf_util.h:628
 *   GENERIC 1, MEMORY 2, ARG 4, CONV 256   jxfpamg-0220/include/jxf_util.h:154-159
 * plus JXAMD_ERROR_DEVICE for a failed HIP call (no reference counterpart).
 * As in the reference (pgmres.c:1656-1662 is compiled out), reaching
 * max_iter without convergence is not an error; it is reported in the stats.
 *
 * Matrices use the reference's BSR layout (jxfpamg-0220/include/jxf_bsr_mv.h:36-48):
 * 0-based int32 row pointer ia[nb+1], column index ja[nnzb], values
 * a[nnzb*b*b] with row-major b x b blocks, and the diagonal block stored
 * first in every block row.
 */
#ifndef JXAMD_H
#define JXAMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JXAMD_SUCCESS        0
#define JXAMD_ERROR_GENERIC  1
#define JXAMD_ERROR_MEMORY   2
#define JXAMD_ERROR_ARG      4
#define JXAMD_ERROR_CONV     256
#define JXAMD_ERROR_DEVICE   512

#define JXAMD_MAX_BLOCK  4
#define JXAMD_MAX_LEVELS 25

/* ------------------------------------------------------------------------
 * Synthetic reservoir systems (no reference counterpart: the reference reads
 * SPE10 files, which are absent).  Natural cell order c = i + nx*(j + ny*k).
 * ---------------------------------------------------------------------- */
enum { JXAMD_FIELD_IID = 0, JXAMD_FIELD_SMOOTH = 1, JXAMD_FIELD_LAYERED = 2 };

typedef struct {
    int nx, ny, nz;          /* grid */
    int block_size;          /* unknowns per cell, 2..4, pressure first */
    int field;               /* JXAMD_FIELD_* */
    double sigma_lnk;        /* std of ln k */
    double corr_len;         /* Gaussian correlation length in cells (SMOOTH/LAYERED) */
    int n_smooth_layers;     /* LAYERED: top layers with a smooth field */
    double channel_ratio;    /* LAYERED: channel / background permeability */
    double kz_ratio;         /* kz / kx */
    int n_wells;             /* vertical well columns (<= 16) */
    double well_boost;       /* pressure-diagonal boost per well cell, x mean T */
    uint64_t seed;
} jxamd_gen_params;

/* number of block rows and stored blocks of the 7-point system */
int jxamd_gen_sizes(const jxamd_gen_params *p, int *nb, int *nnzb);
/* fill ia[nb+1], ja[nnzb], a[nnzb*b*b] and (if rhs != NULL) rhs[nb*b] */
int jxamd_gen_reservoir(const jxamd_gen_params *p, int *ia, int *ja, double *a, double *rhs);

#ifdef __cplusplus
}
#endif
#endif /* JXAMD_H */
