/*
 * jxamd_dev.h -- internal C interface between the host solver (C) and the
 * HIP kernels (jxamd_kernels.hip).  All pointers are device pointers unless
 * named otherwise; `stream` is a hipStream_t passed as void*.
 */
#ifndef JXAMD_DEV_H
#define JXAMD_DEV_H

#include <stddef.h>

#include "jxamd.h"

int jxamd_dev_count(int *n);
int jxamd_dev_malloc(void **p, size_t bytes);
int jxamd_dev_free(void *p);
int jxamd_h2d(void *dst, const void *src, size_t bytes, void *stream);
int jxamd_d2h(void *dst, const void *src, size_t bytes, void *stream);
int jxamd_d2d(void *dst, const void *src, size_t bytes, void *stream);
int jxamd_stream_create(void **s);
int jxamd_stream_destroy(void *s);
int jxamd_stream_sync(void *s);
int jxamd_event_create(void **e);
int jxamd_event_destroy(void *e);
int jxamd_event_record(void *e, void *s);
int jxamd_event_elapsed(float *ms, void *e0, void *e1);
/* events without timing (completion only), waiting on one, pinned host memory */
int jxamd_event_create_sync(void **e);
int jxamd_event_sync(void *e);
int jxamd_host_alloc(void **p, size_t bytes); /* mapped, coherent */
int jxamd_host_dev_ptr(void **d, void *h);    /* its device address */
int jxamd_host_free(void *p);
int jxamd_last_error(void);

/* unit: the layout leaves out identity diagonal blocks (y_i starts from x_i) */
int jxamd_k_bsr_spmv(int b, int nb, int unit, const int *sptr, const int *col, const double *val, const double *x,
                     const double *rhs, double *y, int mode, double alpha, double beta, void *stream);
/* storage / vector type codes of the V-cycle (matrix values vt, vectors xt):
 * (F64, F64), (F32, F32), (F16 with a float power-of-two scale per stored
 * row, F32) */
#define JXAMD_T_F64 0
#define JXAMD_T_F32 1
#define JXAMD_T_F16 2
/* or-ed into vt for the SELL-layout level kernels (csr_spmv, poly_step,
 * gs_chunk_lds and the SELL-G row-group forms): col holds int16 offsets
 * column - row packed two slots per int word, and sptr[nsl + 1 + s] is the
 * first word of slice s */
#define JXAMD_T_PK16 0x100
int jxamd_k_csr_spmv(int vt, int xt, int n, const int *sptr, const int *col, const void *val, const float *rsc,
                     const void *x, int xs, const void *rhs, int rs, void *y, int mode, void *stream);
int jxamd_k_poly_step(int vt, int xt, int n, const int *sptr, const int *col, const void *val, const float *rsc,
                      const void *dinv, const void *f, int fs, const void *uo, void *un, double omega, void *stream);
/* the first polynomial step from a zero guess, un = omega D^{-1} f, one thread
 * per row; xt F64 or F32 */
int jxamd_k_poly_first(int xt, int n, const void *dinv, const void *f, int fs, void *un, double omega, void *stream);
/* restriction y = R x (R in SELL-64 when G = 1, SELL-G otherwise) that also
 * writes the next level's first polynomial step from a zero guess,
 * u = omega D^{-1} y, as jxamd_k_poly_first would */
int jxamd_k_restrict_first(int G, int vt, int xt, int n, const int *sptr, const int *col, const void *val,
                           const float *rsc, const void *x, void *y, const void *dinv, void *u, double omega,
                           void *stream);
int jxamd_sweep_waves(int nsl);
int jxamd_k_gs_sweep(int vt, int xt, int nsl, const int *perm, const int *sptr, const int *col, const void *val,
                     const float *rsc, const void *diag, const void *f, int fs, const void *uo, void *un,
                     const int *dep_ptr, const int *dep_list, int *sflag, int epoch, int dir,
                     unsigned long long *ticket, unsigned long long *base, int *err, void *stream);
/* the same sweep for long rows: slices of 64/G rows of one dependency level
 * (perm: position -> row or -1), a group of G lanes (4..32) per row over the
 * plain CSR rows (ia, ja, val; diagonal first, diag and rsc indexed by row) */
int jxamd_k_gs_sweep_grp(int G, int vt, int xt, int nsl, const int *perm, const int *ia, const int *ja,
                         const void *val, const float *rsc, const void *diag, const void *f, int fs, const void *uo,
                         void *un, const int *dep_ptr, const int *dep_list, int *sflag, int epoch, int dir,
                         unsigned long long *ticket, unsigned long long *base, int *err, void *stream);
/* stage 2: x^(-1) into slice order, then K fused sweeps (columns of the
 * layout are positions, ~pos for lower neighbours; xs holds K slice-ordered
 * vectors of xstride doubles; the last sweep also writes xout in row order) */
int jxamd_k_bhgs_gather(int npos, int b, const int *perm, const double *x0, const double *xp, int pidx, double *out,
                        void *stream);
int jxamd_k_bhgs_multi(int b, int nsl, int K, const int *vorder, const int *perm, const int *sptr, const int *col,
                       const double *val, const double *dinv, const double *rhs, const double *xsl0, double *xs,
                       long long xstride, double *xout, const int *dlo_ptr, const int *dlo_list, const int *dall_ptr,
                       const int *dall_list, int *sflag, int epoch, double w, double omega,
                       unsigned long long *ticket, unsigned long long *base, int *err, void *stream);
/* one stage-2 sweep as on nparts ranks (partitions of at most 8 rows);
 * blocks vt = F64 (dinv F64), F32 (dinv F32) or F16 scaled by rsc per block
 * row (dinv F32), stored two slots per 32-bit word from pair pptr[s] of each
 * slice */
int jxamd_k_bhgs_part(int b, int vt, int nb, int nparts, const int *sptr, const int *pptr, const int *col,
                      const void *val, const float *rsc, const void *dinv, const double *rhs, const double *xo,
                      const void *xp, int xp_f32, int pidx, double *xn, double w, double omega, void *stream);
/* BILU(0) block triangular solve (vt: F64 or F32 factor storage, or F16
 * scaled by rsc[position], a power of two per block row, with F32 dinv) */
int jxamd_k_btri(int b, int vt, int nsl, const int *perm, const int *sptr, const int *col, const void *val,
                 const float *rsc, const void *dinv, const double *rhs, double *out, const int *dep_ptr,
                 const int *dep_list, int *sflag, int epoch, unsigned long long *ticket, unsigned long long *base,
                 int *err, void *stream);
/* hybrid SSOR with short chunks (at most JXAMD_CHUNK_LDS_MAX rows): SELL-64
 * of the off-diagonals in natural row order, diagonal in diag[i] */
#define JXAMD_CHUNK_LDS_MAX 16
int jxamd_gs_chunk_lds_rows(int n, int nchunks); /* longest chunk */
int jxamd_k_gs_chunk_lds(int vt, int xt, int n, int nchunks, const int *sptr, const int *col, const void *val,
                         const float *rsc, const void *diag, const void *f, int fs, const void *uo, void *un,
                         void *stream);
int jxamd_k_gs_chunk(int vt, int xt, int n, int nchunks, const int *wstep, const int *sptr, const int *col,
                     const void *val, const float *rsc, const void *diag, const void *f, int fs, const void *uo,
                     void *un, void *stream);
/* row-group variants (G = 4, 8, 16 or 32 lanes per row): SpMV and polynomial
 * step over the SELL-G layout (64/G rows per slice; sptr: slot-row offsets
 * per slice), the chunked smoother over plain CSR */
int jxamd_k_csr_spmv_grp(int G, int vt, int xt, int n, const int *sptr, const int *col, const void *val,
                         const float *rsc, const void *x, int xs, const void *rhs, int rs, void *y, int mode,
                         void *stream);
int jxamd_k_poly_step_grp(int G, int vt, int xt, int n, const int *sptr, const int *col, const void *val,
                          const float *rsc, const void *dinv, const void *f, int fs, const void *uo, void *un,
                          double omega, void *stream);
int jxamd_k_gs_chunk_grp(int G, int vt, int xt, int n, int nchunks, const int *ia, const int *ja, const void *val,
                         const float *rsc, const void *diag, const void *f, int fs, const void *uo, void *un,
                         void *stream);
/* y = M x, M n x n row-major with rows ld apart (ld a multiple of 4 >= n,
 * zeros beyond column n); x readable up to ld entries */
int jxamd_k_dense_gemv(int xt, int n, int ld, const void *M, const void *x, void *y, void *stream);
int jxamd_k_gather_cast(int n, const double *x, int stride, int off, float *y, void *stream);
int jxamd_k_widen(int n, const float *x, double *y, void *stream);
/* y = (float) x[i*stride + off] and un = omega dinv y (a first polynomial
 * step from zero) in one launch */
int jxamd_k_cast_first(int n, const double *x, int stride, int off, float *y, const float *dinv, float *un,
                       double omega, void *stream);
int jxamd_k_fill_f32(int n, float *y, float v, void *stream);
int jxamd_k_gather_stride(int n, const double *x, int stride, int off, double *y, void *stream);
int jxamd_k_prolong(int nb, int b, int pidx, const double *xp, double *x, void *stream);
int jxamd_k_fill(size_t n, double *y, double v, void *stream);
int jxamd_k_copy(size_t n, const double *x, double *y, void *stream);
int jxamd_k_axpy(size_t n, const double *pa, double a, int neg, const double *x, double *y, void *stream);
int jxamd_k_scale(size_t n, const double *pnorm2, double a, double *y, void *stream);
int jxamd_dot_grid(size_t n);
/* one MGS step of the Arnoldi process (k_mgs_coop): plan (grid 0: not
 * available; coop_launch: the device must support cooperative launches) and
 * launch, cooperative or ordinary (a grid within the resident capacity);
 * hval[0..m-1] = projections, hval[m] = squared norm before scaling;
 * part >= (m + 1) * grid doubles; hhost (device address of mapped host
 * memory, or NULL) also receives hval[0..m] and, if hpos >= 0 (> m), then
 * hseq at hhost[hpos] (host polling instead of a completion event); a waiting workgroup gives up
 * after spin_limit polls (err[0] = 1; cnt and flag must then be cleared
 * before the next launch) */
#define JXAMD_SPIN_LIMIT (1 << 24)
int jxamd_mgs_coop_plan(size_t n, int coop_launch, int *grid, int *tm);
int jxamd_k_mgs_coop(int grid, int tm, int coop_launch, size_t n, const double *V, size_t ld, int m, double *w,
                     double *hval, double *part, unsigned int *cnt, int *flag, int base, int *err, double *hhost,
                     int spin_limit, int hpos, double hseq, void *stream);
int jxamd_k_dot_mgs(size_t n, const double *u, double *y, const double *v_prev, const double *h_prev, double *part,
                    unsigned int *count, double *out, double *hout, void *stream);
int jxamd_k_lincomb(size_t n, int m, const double *V, size_t ld, const double *y, double *w, void *stream);
/* the same with m <= JXAMD_COEFS_MAX coefficients from host memory, passed
 * by value in the launch (no copy); add: w += sum_j y_j V_j */
#define JXAMD_COEFS_MAX 32
typedef struct {
    double c[JXAMD_COEFS_MAX];
} jxamd_coefs;
int jxamd_k_lincomb_host(size_t n, int m, const double *V, size_t ld, const double *y, int add, double *w,
                         void *stream);

#endif
