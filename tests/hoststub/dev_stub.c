/*
 * dev_stub.c -- TEST INFRASTRUCTURE: a host-memory stand-in for the device
 * layer (jxamd_dev.h) so the solver's host code (layouts, schedules, level
 * builds, argument checks, the GMRES control loop) runs on a machine without
 * a GPU, under AddressSanitizer.  Kernel launchers do nothing and return
 * success; device memory is host memory.
 */
#include <stdlib.h>
#include <string.h>

#include "jxamd_dev.h"

int jxamd_dev_count(int *n) { *n = 1; return 0; }
int jxamd_dev_malloc(void **p, size_t bytes) { *p = calloc(1, bytes ? bytes : 1); return *p ? 0 : JXAMD_ERROR_MEMORY; }
int jxamd_dev_free(void *p) { free(p); return 0; }
int jxamd_h2d(void *dst, const void *src, size_t bytes, void *stream) { (void)stream; if (bytes) memcpy(dst, src, bytes); return 0; }
int jxamd_d2h(void *dst, const void *src, size_t bytes, void *stream) { (void)stream; if (bytes) memcpy(dst, src, bytes); return 0; }
int jxamd_d2d(void *dst, const void *src, size_t bytes, void *stream) { (void)stream; if (bytes) memmove(dst, src, bytes); return 0; }
static int token;
int jxamd_stream_create(void **s) { *s = &token; return 0; }
int jxamd_stream_destroy(void *s) { (void)s; return 0; }
int jxamd_stream_sync(void *s) { (void)s; return 0; }
int jxamd_event_create(void **e) { *e = &token; return 0; }
int jxamd_event_destroy(void *e) { (void)e; return 0; }
int jxamd_event_record(void *e, void *s) { (void)e; (void)s; return 0; }
int jxamd_event_elapsed(float *ms, void *e0, void *e1) { (void)e0; (void)e1; *ms = 0.f; return 0; }
int jxamd_event_create_sync(void **e) { *e = &token; return 0; }
int jxamd_event_sync(void *e) { (void)e; return 0; }
int jxamd_host_alloc(void **p, size_t bytes) { *p = calloc(1, bytes ? bytes : 1); return *p ? 0 : JXAMD_ERROR_MEMORY; }
int jxamd_host_free(void *p) { free(p); return 0; }
int jxamd_host_dev_ptr(void **d, void *h) { *d = h; return 0; }
int jxamd_last_error(void) { return 0; }

/* failure-path switches set by the driver: bit 0 drops the norms the dot
 * stand-in writes to pinned memory, bit 1 the MGS stand-in's column sequence
 * word (a pinned word that never arrives), bit 2 makes the residual form of
 * the BSR product return its right-hand side (a nonzero residual, so a solve
 * reaches its Arnoldi loop) */
int jxamd_stub_faults = 0;

int jxamd_k_bsr_spmv(int b, int nb, int unit, const int *sptr, const int *col, const double *val, const double *x,
                     const double *rhs, double *y, int mode, double alpha, double beta, void *stream)
{ if ((jxamd_stub_faults & 4) && mode == 1) for (size_t i = 0; i < (size_t)nb * b; i++) y[i] = rhs[i];
  (void)unit; (void)sptr; (void)col; (void)val; (void)x; (void)alpha; (void)beta; (void)stream; return 0; }
int jxamd_k_csr_spmv(int vt, int xt, int n, const int *sptr, const int *col, const void *val, const float *rsc,
                     const void *x, int xs, const void *rhs, int rs, void *y, int mode, void *stream)
{ (void)vt; (void)xt; (void)n; (void)sptr; (void)col; (void)val; (void)rsc; (void)x; (void)xs; (void)rhs; (void)rs; (void)y; (void)mode; (void)stream; return 0; }
int jxamd_k_poly_step(int vt, int xt, int n, const int *sptr, const int *col, const void *val, const float *rsc,
                      const void *dinv, const void *f, int fs, const void *uo, void *un, double omega, void *stream)
{ (void)vt; (void)xt; (void)n; (void)sptr; (void)col; (void)val; (void)rsc; (void)dinv; (void)f; (void)fs; (void)uo; (void)un; (void)omega; (void)stream; return 0; }
int jxamd_k_poly_first(int xt, int n, const void *dinv, const void *f, int fs, void *un, double omega, void *stream)
{ (void)xt; (void)n; (void)dinv; (void)f; (void)fs; (void)un; (void)omega; (void)stream; return 0; }
int jxamd_k_restrict_first(int G, int vt, int xt, int n, const int *sptr, const int *col, const void *val,
                           const float *rsc, const void *x, void *y, const void *dinv, void *u, double omega,
                           void *stream)
{ (void)G; (void)vt; (void)xt; (void)n; (void)sptr; (void)col; (void)val; (void)rsc; (void)x; (void)y; (void)dinv; (void)u; (void)omega; (void)stream; return 0; }
int jxamd_sweep_waves(int nsl) { return nsl < 2048 ? (nsl > 0 ? nsl : 1) : 2048; }
int jxamd_k_gs_sweep(int vt, int xt, int nsl, const int *perm, const int *sptr, const int *col, const void *val,
                     const float *rsc, const void *diag, const void *f, int fs, const void *uo, void *un,
                     const int *dep_ptr, const int *dep_list, int *sflag, int epoch, int dir,
                     unsigned long long *ticket, unsigned long long *base, int *err, void *stream)
{ (void)vt; (void)xt; (void)nsl; (void)perm; (void)sptr; (void)col; (void)val; (void)rsc; (void)diag; (void)f; (void)fs; (void)uo; (void)un; (void)dep_ptr; (void)dep_list; (void)sflag; (void)epoch; (void)dir; (void)ticket; (void)base; (void)err; (void)stream; return 0; }
int jxamd_k_gs_sweep_grp(int G, int vt, int xt, int nsl, const int *perm, const int *ia, const int *ja,
                         const void *val, const float *rsc, const void *diag, const void *f, int fs, const void *uo,
                         void *un, const int *dep_ptr, const int *dep_list, int *sflag, int epoch, int dir,
                         unsigned long long *ticket, unsigned long long *base, int *err, void *stream)
{ (void)G; (void)vt; (void)xt; (void)nsl; (void)perm; (void)ia; (void)ja; (void)val; (void)rsc; (void)diag; (void)f; (void)fs; (void)uo; (void)un; (void)dep_ptr; (void)dep_list; (void)sflag; (void)epoch; (void)dir; (void)ticket; (void)base; (void)err; (void)stream; return 0; }
int jxamd_k_bhgs_gather(int npos, int b, const int *perm, const double *x0, const double *xp, int pidx, double *out,
                        void *stream)
{ (void)npos; (void)b; (void)perm; (void)x0; (void)xp; (void)pidx; (void)out; (void)stream; return 0; }
int jxamd_k_bhgs_multi(int b, int nsl, int K, const int *vorder, const int *perm, const int *sptr, const int *col,
                       const double *val, const double *dinv, const double *rhs, const double *xsl0, double *xs,
                       long long xstride, double *xout, const int *dlo_ptr, const int *dlo_list, const int *dall_ptr,
                       const int *dall_list, int *sflag, int epoch, double w, double omega,
                       unsigned long long *ticket, unsigned long long *base, int *err, void *stream)
{
    (void)b; (void)nsl; (void)perm; (void)sptr; (void)col; (void)val; (void)dinv; (void)rhs; (void)xstride;
    (void)dlo_ptr; (void)dlo_list; (void)sflag; (void)epoch; (void)w; (void)omega; (void)ticket; (void)base;
    (void)err; (void)stream;
    if (!xsl0 || !xs || !xout || (K > 1 && (!vorder || !dall_ptr || !dall_list))) return JXAMD_ERROR_ARG;
    return 0;
}
int jxamd_k_btri(int b, int vt, int nsl, const int *perm, const int *sptr, const int *col, const void *val,
                 const float *rsc, const void *dinv, const double *rhs, double *out, const int *dep_ptr,
                 const int *dep_list, int *sflag, int epoch, unsigned long long *ticket, unsigned long long *base,
                 int *err, void *stream)
{ (void)b; (void)vt; (void)nsl; (void)perm; (void)sptr; (void)col; (void)val; (void)rsc; (void)dinv; (void)rhs; (void)out; (void)dep_ptr; (void)dep_list; (void)sflag; (void)epoch; (void)ticket; (void)base; (void)err; (void)stream; return 0; }
int jxamd_gs_chunk_lds_rows(int n, int nchunks)
{
    if (nchunks < 1 || n < nchunks) return 0;
    return n / nchunks + (n % nchunks ? 1 : 0);
}
int jxamd_k_gs_chunk_lds(int vt, int xt, int n, int nchunks, const int *sptr, const int *col, const void *val,
                         const float *rsc, const void *diag, const void *f, int fs, const void *uo, void *un,
                         void *stream)
{ (void)vt; (void)xt; (void)n; (void)nchunks; (void)sptr; (void)col; (void)val; (void)rsc; (void)diag; (void)f; (void)fs; (void)uo; (void)un; (void)stream; return 0; }
int jxamd_k_gs_chunk(int vt, int xt, int n, int nchunks, const int *wstep, const int *sptr, const int *col,
                     const void *val, const float *rsc, const void *diag, const void *f, int fs, const void *uo,
                     void *un, void *stream)
{ (void)vt; (void)xt; (void)n; (void)nchunks; (void)wstep; (void)sptr; (void)col; (void)val; (void)rsc; (void)diag; (void)f; (void)fs; (void)uo; (void)un; (void)stream; return 0; }
int jxamd_k_dense_gemv(int xt, int n, int ld, const void *M, const void *x, void *y, void *stream)
{ (void)xt; (void)n; (void)ld; (void)M; (void)x; (void)y; (void)stream; return 0; }
int jxamd_k_gather_cast(int n, const double *x, int stride, int off, float *y, void *stream)
{ (void)n; (void)x; (void)stride; (void)off; (void)y; (void)stream; return 0; }
int jxamd_k_widen(int n, const float *x, double *y, void *stream) { (void)n; (void)x; (void)y; (void)stream; return 0; }
int jxamd_k_cast_first(int n, const double *x, int stride, int off, float *y, const float *dinv, float *un,
                       double omega, void *stream)
{ (void)n; (void)x; (void)stride; (void)off; (void)y; (void)dinv; (void)un; (void)omega; (void)stream; return 0; }
int jxamd_k_fill_f32(int n, float *y, float v, void *stream) { (void)n; (void)y; (void)v; (void)stream; return 0; }
int jxamd_k_gather_stride(int n, const double *x, int stride, int off, double *y, void *stream)
{ (void)n; (void)x; (void)stride; (void)off; (void)y; (void)stream; return 0; }
int jxamd_k_prolong(int nb, int b, int pidx, const double *xp, double *x, void *stream)
{ (void)nb; (void)b; (void)pidx; (void)xp; (void)x; (void)stream; return 0; }
int jxamd_k_fill(size_t n, double *y, double v, void *stream) { (void)n; (void)y; (void)v; (void)stream; return 0; }
int jxamd_k_copy(size_t n, const double *x, double *y, void *stream) { (void)n; (void)x; (void)y; (void)stream; return 0; }
int jxamd_k_axpy(size_t n, const double *pa, double a, int neg, const double *x, double *y, void *stream)
{ (void)n; (void)pa; (void)a; (void)neg; (void)x; (void)y; (void)stream; return 0; }
int jxamd_k_scale(size_t n, const double *pnorm2, double a, double *y, void *stream)
{ (void)n; (void)pnorm2; (void)a; (void)y; (void)stream; return 0; }
int jxamd_dot_grid(size_t n) { (void)n; return 1; }
int jxamd_mgs_coop_plan(size_t n, int coop_launch, int *grid, int *tm) { (void)coop_launch; *grid = n > 0 ? 2 : 0; *tm = 4; return 0; }
int jxamd_k_mgs_coop(int grid, int tm, int coop_launch, size_t n, const double *V, size_t ld, int m, double *w,
                     double *hval, double *part, unsigned int *cnt, int *flag, int base, int *err, double *hhost,
                     int spin_limit, int hpos, double hseq, void *stream)
{ if (hhost && hpos >= 0 && !(jxamd_stub_faults & 2)) hhost[hpos] = hseq; /* the column's sequence number at once */
  (void)grid; (void)tm; (void)coop_launch; (void)n; (void)V; (void)ld; (void)m; (void)w; (void)hval; (void)part; (void)cnt; (void)flag; (void)base; (void)err; (void)hhost; (void)spin_limit; (void)stream; return 0; }
int jxamd_k_dot_mgs(size_t n, const double *u, double *y, const double *v_prev, const double *h_prev, double *part,
                    unsigned int *count, double *out, double *hout, void *stream)
{ /* the device stand-in computes the product on the host (device memory is host memory here) */
  double a = 0.0;
  for (size_t i = 0; i < n; i++) {
      if (v_prev) y[i] -= *h_prev * v_prev[i];
      a += u[i] * y[i];
  }
  *out = a;
  if (hout && !(jxamd_stub_faults & 1)) *hout = a;
  (void)part; (void)count; (void)stream; return 0; }
int jxamd_k_lincomb_host(size_t n, int m, const double *V, size_t ld, const double *y, int add, double *w,
                         void *stream)
{ (void)n; (void)V; (void)ld; (void)add; (void)w; (void)stream; if (m < 0 || m > JXAMD_COEFS_MAX || (m && !y)) return JXAMD_ERROR_ARG; return 0; }
int jxamd_k_lincomb(size_t n, int m, const double *V, size_t ld, const double *y, double *w, void *stream)
{ (void)n; (void)m; (void)V; (void)ld; (void)y; (void)w; (void)stream; return 0; }

int jxamd_mlp_forward(const float *d_params, const float *d_x, int n, float *d_y, void *stream)
{
    (void)d_params; (void)d_x; (void)stream;
    for (int i = 0; i < n * JXAMD_MLP_NOUT; i++) d_y[i] = 1.0f;
    return 0;
}
int jxamd_mlp_train(float *d_params, float *d_m, float *d_v, float *d_grad, const float *d_x, const float *d_t, int n,
                    int nout, int steps, int t0, float lr, int update, float *d_loss, void *stream)
{ (void)d_params; (void)d_m; (void)d_v; (void)d_grad; (void)d_x; (void)d_t; (void)n; (void)nout; (void)steps; (void)t0; (void)lr; (void)update; (void)d_loss; (void)stream; return 0; }

int jxamd_k_csr_spmv_grp(int G, int vt, int xt, int n, const int *ia, const int *ja, const void *val,
                         const float *rsc, const void *x, int xs, const void *rhs, int rs, void *y, int mode,
                         void *stream)
{ (void)G; (void)vt; (void)xt; (void)n; (void)ia; (void)ja; (void)val; (void)rsc; (void)x; (void)xs; (void)rhs; (void)rs; (void)y; (void)mode; (void)stream; return 0; }
int jxamd_k_poly_step_grp(int G, int vt, int xt, int n, const int *ia, const int *ja, const void *val,
                          const float *rsc, const void *dinv, const void *f, int fs, const void *uo, void *un,
                          double omega, void *stream)
{ (void)G; (void)vt; (void)xt; (void)n; (void)ia; (void)ja; (void)val; (void)rsc; (void)dinv; (void)f; (void)fs; (void)uo; (void)un; (void)omega; (void)stream; return 0; }
int jxamd_k_gs_chunk_grp(int G, int vt, int xt, int n, int nchunks, const int *ia, const int *ja, const void *val,
                         const float *rsc, const void *diag, const void *f, int fs, const void *uo, void *un,
                         void *stream)
{ (void)G; (void)vt; (void)xt; (void)n; (void)nchunks; (void)ia; (void)ja; (void)val; (void)rsc; (void)diag; (void)f; (void)fs; (void)uo; (void)un; (void)stream; return 0; }

int jxamd_k_bhgs_part(int b, int vt, int nb, int nparts, const int *sptr, const int *pptr, const int *col,
                      const void *val, const float *rsc, const void *dinv, const double *rhs, const double *xo,
                      const void *xp, int xp_f32, int pidx, double *xn, double w, double omega, void *stream)
{ (void)b; (void)vt; (void)nb; (void)nparts; (void)sptr; (void)pptr; (void)col; (void)val; (void)rsc; (void)dinv; (void)rhs; (void)xo; (void)xp; (void)xp_f32; (void)pidx; (void)xn; (void)w; (void)omega; (void)stream; return 0; }
