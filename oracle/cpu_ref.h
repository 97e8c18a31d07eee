/*
 * cpu_ref.h -- TEST INFRASTRUCTURE: the oracle.
 *
 * A single-threaded FP64 CPU restatement of the reference's CPR-GMRES solve
 * phase (jxfpamg-0220, one MPI rank, OMP_NUM_THREADS=1).  Every function
 * names the reference lines it follows and keeps their operation order, so
 * on the same inputs it reproduces the reference bit for bit.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, as the
 * checker; the product never links it.
 */
#ifndef CPU_REF_H
#define CPU_REF_H

#define CREF_MAX_LEVELS 25

typedef struct {
    int n, ncols;
    const int *ia, *ja;
    const double *a;
} cref_csr;

typedef struct {
    int num_levels;
    cref_csr A[CREF_MAX_LEVELS]; /* A[0] = pressure matrix */
    cref_csr P[CREF_MAX_LEVELS]; /* P[l]: n_l x n_{l+1}, l < num_levels-1 */
    int relax_threads[CREF_MAX_LEVELS]; /* relax-6 thread count per level; <= 1: sequential */
    /* poly_degree > 0: every smoothing on level l (pre and post) is instead one
     * polynomial smoother u <- u + w_k D^{-1} (f - A u), k = 0..degree-1,
     * with w_k = poly_omega[l][k] (no reference counterpart) */
    int poly_degree;
    double poly_omega[CREF_MAX_LEVELS][8];
} cref_hier;

typedef struct {
    int nb, b;
    const int *ia, *ja;
    const double *a;
} cref_bsr;

typedef struct {
    int k_dim, max_iter, is_check_restarted;
    double tol, a_tol;
    int pressure_index, stage1_maxit, stage2_maxit;
    double stage2_weight, stage2_omega;
    /* 2: block hybrid SOR sweeps (the reference's BGS); 5: BILU(0)
     * corrections x += (LU)^{-1}(r - A x) with the factors below */
    int stage2_solver_type;
    const double *bilu_lu, *bilu_dinv;
    /* stage2_solver_type 2 with stage2_parts > 1: the sweeps of a run on
     * stage2_parts ranks (contiguous block-row partitions) */
    int stage2_parts;
    /* GMRES stops when one step's residual estimate grows by more than this
     * factor (pgmres.c:1325-1333; JXF_GMRESSetConvFacThreshold, default 1e4) */
    double convfac_threshold;
} cref_params;

typedef struct {
    int iterations, converged, num_precond;
    double final_rel_res, b_norm;
    int num_false_conv; /* estimates at tolerance whose true residual was not (pgmres.c:1583-1608) */
} cref_stats;

void cref_bsr_matvec(const cref_bsr *A, double alpha, const double *x, double beta, double *y);
void cref_csr_matvec(const cref_csr *A, double alpha, const double *x, double beta, double *y);
void cref_csr_matvecT(const cref_csr *A, double alpha, const double *x, double beta, double *y);
double cref_dot(int n, const double *x, const double *y);
void cref_relax6(const cref_csr *A, const double *f, double *u);
void cref_relax6_threads(const cref_csr *A, const double *f, double *u, int num_threads);
void cref_poly_smooth(const cref_csr *A, const double *f, double *u, int degree, const double *omega);
int cref_relax9(const cref_csr *A, const double *f, double *u);
int cref_gselim(double *A, double *x, int n);
void cref_bsr_block_matvec_stable(double alpha, const double *mat, const double *v, double beta, double *ov, int b);
void cref_bsr_block_invmatvec_stable(const double *mat, const double *v, double *ov, int b);
void cref_bsr_hgs(const cref_bsr *A, double *x, const double *rhs, double *vtemp, double w, double omega, int fwd);
void cref_bsr_hgs_parts(const cref_bsr *A, double *x, const double *rhs, double *vtemp, double w, double omega,
                        int nparts);
int cref_pamg_solve(const cref_hier *H, const double *f, double *u);
int cref_bilu0_factor(const cref_bsr *A, double *lu, double *dinv);
void cref_bilu0_solve(const cref_bsr *A, const double *lu, const double *dinv, const double *r, double *z);
int cref_cpr_precond(const cref_bsr *A, const cref_hier *H, const cref_params *p, const double *r, double *x);
int cref_gmres_solve(const cref_bsr *A, const cref_hier *H, const cref_params *p, const double *b, double *x,
                     double *res_history, cref_stats *st);

#endif
