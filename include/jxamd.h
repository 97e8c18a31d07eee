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


/* ------------------------------------------------------------------------
 * Reference file formats and host setup steps.
 * ---------------------------------------------------------------------- */
/* binary BSR: jxf_BSRMatrixPrint_Binary / jxf_BSRMatrixRead_Binary
 * (jxfpamg-0220/csrc/mat_vec_data/seq_bsr_matrix.c:1935-2019, :2038+) */
int jxamd_bsr_write_binary(const char *path, int nb, int ncols, int nnzb, int b, const int *ia, const int *ja,
                           const double *a);
int jxamd_bsr_read_binary_sizes(const char *path, int *nb, int *ncols, int *nnzb, int *b);
int jxamd_bsr_read_binary(const char *path, int *ia, int *ja, double *a);
/* text RHS: read_rhs_file (src/test_bsr_cprgmres.c:26-76) */
int jxamd_rhs_write_text(const char *path, int n, const double *v);
int jxamd_rhs_read_text(const char *path, int n, double *v);
/* ABF decoupling in place, jxf_BSRMatrixDecouple(..., JXF_DECOUP_ABF, 0)
 * (jxfpamg-0220/csrc/cpr/jxf_bsr_decoup.c:1273-1348); rhs may be NULL */
int jxamd_bsr_decouple_abf(int nb, int b, const int *ia, const int *ja, double *a, double *rhs);
/* the same decoupling with true block inverses and an out-of-place rhs
 * product: D^{-1} A, D^{-1} b.  No reference counterpart; it differs from
 * jxamd_bsr_decouple_abf for b >= 4 (the reference's 4x4 closed form,
 * jxf_bsr_decoup.c:112-183, and its Gauss-Jordan above 4x4, :200-266, do not
 * return the inverse) and in the rhs for b > 1 */
int jxamd_bsr_decouple_abf_exact(int nb, int b, const int *ia, const int *ja, double *a, double *rhs);
/* Decoupling types, JXF_DecoupType (jxfpamg-0220/include/jxf_bsr_decoup.h:13-21) */
enum {
    JXAMD_DECOUP_NONE = 0,
    JXAMD_DECOUP_ABF = 1,     /* D_i^{-1} x row i, jxf_bsr_decoup.c:364-527 */
    JXAMD_DECOUP_ANL = 2,     /* analytical, :529-588 */
    JXAMD_DECOUP_SEM = 3,     /* semi-analytical, :591-657 */
    JXAMD_DECOUP_QI = 4,      /* quasi-IMPES, :659-718 */
    JXAMD_DECOUP_TIMPES = 5,  /* True-IMPES (column sums), :721-1099 */
    JXAMD_DECOUP_TIMPES2 = 6  /* True-IMPES, second form, :1102-1262 */
};
/* In-place decoupling of a BSR system, jxf_BSRMatrixDecouple(A, rhs, type,
 * is_thermal) (jxfpamg-0220/csrc/cpr/jxf_bsr_decoup.c:1273-1348), bit for
 * bit on the host, b = 1..10; rhs may be NULL (left alone) */
int jxamd_bsr_decouple(int nb, int b, const int *ia, const int *ja, double *a, double *rhs, int type, int is_thermal);
/* The same on device arrays (b = 2..4), bit for bit with the host routine:
 * one HBM pass over the blocks (True-IMPES: one more pass for the column
 * sums, which needs a structurally symmetric block pattern, else
 * JXAMD_ERROR_ARG with nothing modified).  If d_p_a is not NULL it also
 * receives the pressure matrix values of the decoupled blocks,
 * d_p_a[k] = A_k[pidx][pidx] (the pattern is A's: ia, ja), as
 * jxamd_bsr_extract_pressure_device would, in the same pass.  Synchronous on
 * `stream`; allocates a work buffer of nb*b doubles for True-IMPES. */
int jxamd_bsr_decouple_device(int nb, int b, const int *d_ia, const int *d_ja, double *d_a, double *d_rhs, int type,
                              int is_thermal, int pidx, double *d_p_a, void *stream);
/* pressure matrix on device arrays: d_p_ia = ia, d_p_ja = ja, d_p_a[k] =
 * A_k[pidx][pidx] (jxf_BSRMatrixGetSubmatrix, seq_bsr_matrix.c:214-250, as
 * called by the pressure extraction of par_bsr_matrix.c:327-392); d_p_ia /
 * d_p_ja may be NULL (not written).  Asynchronous on `stream`. */
int jxamd_bsr_extract_pressure_device(int nb, int b, int nnzb, const int *d_ia, const int *d_ja, const double *d_a,
                                      int pidx, int *d_p_ia, int *d_p_ja, double *d_p_a, void *stream);
/* pressure matrix A_p[i,j] = A_ij[p][p] on A's pattern, jxf_BSRMatrixGetSubmatrix
 * (jxfpamg-0220/csrc/mat_vec_data/seq_bsr_matrix.c:214-250); p_ia[nb+1], p_ja/p_a[nnzb] */
int jxamd_bsr_extract_pressure(int nb, int b, const int *ia, const int *ja, const double *a, int pidx, int *p_ia,
                               int *p_ja, double *p_a);

/* ------------------------------------------------------------------------
 * Host-side system description (arrays stay owned by the caller).
 * ---------------------------------------------------------------------- */
typedef struct {
    int nb, ncols, nnzb, b;      /* block rows, block columns, stored blocks, block size */
    const int *ia, *ja;          /* ia[nb+1], ja[nnzb], 0-based, diagonal block first */
    const double *a;             /* a[nnzb*b*b], row-major blocks */
} jxamd_bsr;

typedef struct {
    int n, ncols, nnz;
    const int *ia, *ja;          /* square level matrices: diagonal stored first */
    const double *a;
} jxamd_csr;

/* Pressure-AMG hierarchy imported as built on the host (coarse operators A[l]
 * and interpolation P[l]: n_l x n_{l+1}; restriction is P^T, as
 * R_array == P_array in jxfpamg-0220/csrc/amg/par_amg_setup.c:415-425). */
typedef struct {
    int num_levels;
    jxamd_csr A[JXAMD_MAX_LEVELS];
    jxamd_csr P[JXAMD_MAX_LEVELS];
} jxamd_hierarchy;

/* Host construction of the hierarchy with the settings of JXF_CPRSetup
 * (jxfpamg-0220/csrc/cpr/jxf_cpr.c:319-377): classical strength, HMIS
 * coarsening, extended+i interpolation, Galerkin coarse operators; the
 * operators are bit-identical to the reference's on one rank. */
typedef struct {
    int max_levels;          /* 25 */
    int coarse_threshold;    /* 1000: stop once a level has at most this many points */
    double strong_threshold; /* 0.25 */
    double max_row_sum;      /* 0.9 */
} jxamd_amg_params;
void jxamd_amg_params_default(jxamd_amg_params *p);
/* A0: pressure matrix (square, diagonal first).  Allocates every level
 * (A[0] is a copy); release with jxamd_hierarchy_free. */
int jxamd_amg_setup(const jxamd_csr *A0, const jxamd_amg_params *p, jxamd_hierarchy *H);
void jxamd_hierarchy_free(jxamd_hierarchy *H);

/* Exact block ILU(0) on A's block pattern (diagonal block first, then
 * ascending block columns): lu[nnzb*b*b] holds L (strictly lower blocks, unit
 * diagonal implied) and U (diagonal and upper blocks), dinv[nb*b*b] the
 * inverted diagonal blocks of U.  Stands in for the reference's iterative
 * BILU setup (jxfpamg-0220/csrc/ilu/par_bilu_setup.c:17, type 3). */
int jxamd_bilu0_factor(int nb, int b, const int *ia, const int *ja, const double *a, double *lu, double *dinv);

/* Parameters.  CPR keys mirror JXF_CPRSetParameter / JXF_CPRSetRealParameter
 * (jxfpamg-0220/csrc/cpr/jxf_cpr.c:226-262) plus the stage-2 weights fixed at
 * jxf_cpr.c:547-549; GMRES keys mirror the JXF_GMRESSet* setters
 * (jxfpamg-0220/csrc/krylov/pgmres.c:101-306). */
typedef struct {
    int k_dim, max_iter, is_check_restarted;
    double tol, a_tol;
    int pressure_index, stage1_maxit, stage2_maxit;
    int stage1_solver_type, stage2_solver_type, print_level; /* stage2_solver_type: 2 block hybrid SOR
                                                               * (jxf_cpr.c:573), 5 BILU(0) corrections */
    double threshold, stage2_weight, stage2_omega;
    /* AMG smoother (relax 6) threading, jxfpamg-0220/csrc/operation/
     * relaxations/par_relax_6.c: relax_threads = T > 1 is the reference's
     * T-thread hybrid SSOR (T contiguous chunks, Gauss-Seidel inside a chunk,
     * Jacobi across chunks) on every level, as with OMP_NUM_THREADS = T;
     * relax_chunk = c > 0 uses ceil(n_l / c) chunks on level l instead.
     * Both 0/1: the one-thread sequential sweeps. */
    int relax_threads, relax_chunk;
    /* Stage-1 (pressure AMG V-cycle) precision; no reference counterpart in
     * the FP64 library (the reference's mixed driver, src/test_bsr_cprgmres_hybrid.c:217-439,
     * runs stage 1 in FP32): 0 FP64; 1 FP32 level matrices, vectors and
     * accumulation; 2 FP16 level matrices with a power-of-two scale per row,
     * FP32 vectors and accumulation.  Stage 2 and GMRES stay FP64. */
    int precond_precision;
    /* AMG smoother: 0 = relax 6 as in the reference (two calls per pre- and
     * post-smoothing); d = 1..8 = a degree-d polynomial smoother instead, one
     * application per pre- and post-smoothing, u <- u + w_k D^{-1}(f - A u)
     * for k < d (no reference counterpart).  Step sizes default to Chebyshev
     * roots on [0.3, 1.1] x an estimate of lambda_max(D^{-1} A) per level and
     * can be replaced with jxamd_solver_set_poly (e.g. learned ones). */
    int poly_degree;
    /* Stage-2 block hybrid SOR (type 2) partitions: 0 = the one-rank sweep
     * (exact sequential order); c = 1..8 = the sweep as run on ceil(nb / c)
     * ranks, i.e. contiguous partitions of about c block rows, Gauss-Seidel
     * inside a partition and the other partitions' values at entry
     * (jxfpamg-0220/csrc/amg/par_relax_bsr.c:296-330, 523-557). */
    int stage2_chunk;
    /* Storage of the partitioned stage-2 sweeps' matrix blocks and inverted
     * diagonal blocks (stage2_solver_type 2 with stage2_chunk > 0), and of
     * the BILU(0) factors (stage2_solver_type 5): 0 FP64 (BILU factors FP32
     * when precond_precision > 0), 1 FP32, 2 FP16 blocks scaled by a power
     * of two per block row (per factor triangle for BILU) with FP32 inverted
     * diagonals; FP64 vectors and arithmetic throughout.  No reference
     * counterpart (its FP64 BILU keeps its factors in FP64,
     * src/bilu_double_helper.c:18-25; its mixed driver keeps stage 2 in FP64). */
    int stage2_precision;
    /* Arnoldi orthogonalisation kernels: 0 one cooperative launch per MGS
     * step with the new vector held in registers (when the device supports
     * cooperative launches and the vector fits the resident grid), 1 one
     * launch per projection, 2 (the default) the same one-launch step as 0
     * by an ordinary launch of a grid within the resident capacity (no
     * cooperative-launch synchronisation around it, which leaves the device
     * idle about 30 us per step; a workgroup that waits too long ends the
     * solve with an error instead of hanging).  Same arithmetic (MGS,
     * pgmres.c:1258-1290 order); the reductions group their partial sums
     * differently. */
    int arnoldi_kernel;
    /* Update of x at the end of a GMRES cycle: 0 x += Z y with the vectors
     * z_j = M^{-1} v_j that the Arnoldi steps formed kept in a second basis
     * (k_dim more vectors; no preconditioner application at the cycle's
     * end; the flexible-GMRES update, exact for the computed z_j), 1 x +=
     * M^{-1} (V y) as the reference (pgmres.c:1477-1503; one more
     * application per cycle).  For a linear FP64 preconditioner the two agree
     * to rounding. */
    int solution_update;
    /* GMRES stops when one step's residual estimate grows by more than this
     * factor, JXF_GMRESSetConvFacThreshold (pgmres.c:201, default 1e4 at
     * :449, test at :1325-1333) */
    double convfac_threshold;
    /* V-cycle depth: 0 = every level of the imported hierarchy (the
     * reference's coarsest level, at most 1000 rows, jxf_cpr.c:346); c > 0
     * stops at the first level with at most c rows (c <= 8192) and solves it
     * with an explicit dense inverse (one GEMV per cycle), dropping the
     * levels below it.  No reference counterpart. */
    int coarse_rows;
} jxamd_params;

/* Values of src/test_bsr_cprgmres.c:317-367 with tol 1e-6 and one-thread
 * smoothing (relax_threads = relax_chunk = 0). */
void jxamd_params_default(jxamd_params *p);

typedef struct {
    int iterations;              /* total Arnoldi steps (pgmres.c:1644) */
    int converged;
    int num_precond;             /* CPR applications */
    int num_restarts;
    double final_rel_res;        /* as JXF_GMRESGetFinalRelativeResidualNorm */
    double true_rel_res;         /* ||b - A x|| / ||b|| recomputed at exit, FP64 */
    double b_norm;
    double t_solve_ms;           /* device time of the whole solve (HIP events) */
    double t_precond_ms;         /* device time inside CPR applications, measured when print_level > 0
                                  * or profiling is on (each application then ends with a sync); else 0 */
    int num_false_conv;          /* residual estimates at the tolerance whose true residual was not
                                  * ("false convergence", pgmres.c:1200-1241, 1583-1608) */
    int diverged;                /* stopped by the convergence-factor test (pgmres.c:1325-1333) */
} jxamd_stats;

typedef struct jxamd_solver_s jxamd_solver;

/* Upload A and the hierarchy into HBM (SELL-64 layouts, level schedules,
 * coarse inverse).  Replaces the device-side part of JXF_CPRSetup
 * (jxf_cpr.c:268-422) + JXF_GMRESSetup (pgmres.c:544-622). */
int jxamd_solver_create(jxamd_solver **s, const jxamd_bsr *A, const jxamd_hierarchy *H, const jxamd_params *p);
int jxamd_solver_destroy(jxamd_solver *s);
/* CPR-GMRES solve, host vectors (x in: initial guess, out: solution).
 * Replaces JXF_GMRESSolve with the CPR callback (pgmres.c:950-1713). */
int jxamd_solver_solve(jxamd_solver *s, const double *b, double *x, double *res_history, jxamd_stats *st);
/* Same with b and x already resident in HBM. */
int jxamd_solver_solve_device(jxamd_solver *s, const double *d_b, double *d_x, double *res_history,
                              jxamd_stats *st);
/* x = M^{-1} r, device vectors.  Replaces JXF_CPRPrecond (jxf_cpr.c:424-610). */
int jxamd_cpr_precond_device(jxamd_solver *s, const double *d_r, double *d_x);
/* The CPR application as a preconditioner callback of the reference's GMRES:
 * the shape of JXF_PtrToSolverFcn, int (*)(void *precond_data, void *A,
 * void *b, void *x) (registered by JXF_GMRESSetPrecond, pgmres.c:267-281;
 * called at pgmres.c:1258, 1400, 1444, 1500, 1538).  precond_data points to a
 * jxamd_precond_ctx; b and x are the caller's vector objects, whose host
 * arrays (nb*b doubles) vector_data returns (for the reference:
 * jxf_VectorData(jxf_ParVectorLocalVector(v))).  A is ignored, as
 * JXF_CPRPrecond ignores its matrix argument (jxf_cpr.c:427).  x = M^{-1} b
 * through HBM (two PCIe copies); returns 0 or an error flag. */
typedef struct {
    jxamd_solver *solver;
    double *(*vector_data)(void *vec);
} jxamd_precond_ctx;
int jxamd_cpr_precond_callback(void *precond_data, void *A, void *b, void *x);

/* The CSR-array solve entry with the arguments of JXFPAMG_Solver_Interface
 * (src/jxpamg_interface.c:118-257) minus the MPI communicator (one GPU, one
 * process), for solver_id 220 = CPR-GMRES (src/jxpamg_interface.c:1967-2043).
 * row = nb*num_functions node-major rows; ia/ja may be 1-based (decremented
 * in place, :143-150) and each row's diagonal is swapped to the front in
 * place (jxf_CSRMatrixReorder, seq_csr_matop.c:189-231).  Regroups into BSR
 * with b = num_functions (2..4), builds the pressure-AMG hierarchy on the
 * host, solves from x = 0 on the GPU and writes ser_x and *iter.  argv:
 * -kdim, -maxiter (reference options), -rchunk, -rthreads (relax-6 chunking).
 * Any other solver_id returns JXAMD_ERROR_ARG. */
int jxamd_solver_interface(int argc, char *argv[], int row, int *ia, int *ja, double *a, double *ser_x,
                           double *ser_b, double tol, int solver_id, int num_functions, int *iter);

/* Polynomial smoother step sizes of one level (degree must equal
 * poly_degree); get returns the ones in use and the level's lambda_max
 * estimate of D^{-1} A. */
int jxamd_solver_set_poly(jxamd_solver *s, int level, int degree, const double *omega);
int jxamd_solver_get_poly(jxamd_solver *s, int level, int degree, double *omega, double *lambda_max);
int jxamd_solver_num_levels(jxamd_solver *s, int *num_levels);

/* ------------------------------------------------------------------------
 * Learned polynomial smoother (no reference counterpart).  A 3-layer MLP
 * (JXAMD_MLP_NFEAT level features -> 64 -> 64 -> JXAMD_MLP_NOUT, tanh hidden
 * units) predicts the roots of each level's polynomial smoother of degree
 * m <= 4 (its first m outputs) in units of lambda_max(D^{-1} A): step
 * w_k = 1 / (tau_k lambda_max).  One network serves one degree.
 * BF16 operands on the matrix cores, FP32 accumulation and master weights.
 * Parameter layout (FP32): W1[64][16] b1[64] W2[64][64] b2[64] W3[4][64] b3[4].
 * ---------------------------------------------------------------------- */
#define JXAMD_MLP_NFEAT  16
#define JXAMD_MLP_NOUT   4
#define JXAMD_MLP_NPARAM 5508
/* The training labels are roots of degree-m Chebyshev polynomials on
 * [lo, hi] x lambda_max with lo in [JXAMD_MLP_LO_MIN, JXAMD_MLP_LO_MAX] and
 * hi in [JXAMD_MLP_HI_MIN, JXAMD_MLP_HI_MAX]; a predicted root is clamped
 * into the span of that family for its degree (its trust range). */
#define JXAMD_MLP_LO_MIN 0.03
#define JXAMD_MLP_LO_MAX 0.45
#define JXAMD_MLP_HI_MIN 0.9
#define JXAMD_MLP_HI_MAX 1.4
/* y[n][4] = MLP(x[n][16]), device pointers */
int jxamd_mlp_forward(const float *d_params, const float *d_x, int n, float *d_y, void *stream);
/* steps Adam steps (lr, betas 0.9/0.999, eps 1e-8; update = 0 only computes
 * the gradient) on the mean squared error of the first nout (1..4) outputs
 * over mini-batches of 32 samples taken cyclically from (x[n][16],
 * t[n][nout]) starting at batch t0; d_grad receives the last step's
 * gradient, d_loss[steps] (optional) the losses */
int jxamd_mlp_train(float *d_params, float *d_m, float *d_v, float *d_grad, const float *d_x, const float *d_t, int n,
                    int nout, int steps, int t0, float lr, int update, float *d_loss, void *stream);
/* the features of one smoothing level (solver created with poly_degree > 0) */
int jxamd_solver_poly_features(jxamd_solver *s, int level, float *feat);
/* V-cycle contraction factor: ||e_k|| / ||e_0|| to the power 1/ncycles for
 * ncycles V-cycles on f = 0 from a fixed pseudo-random e_0 */
int jxamd_solver_vcycle_factor(jxamd_solver *s, int ncycles, double *factor);
/* The learned smoother and its judge.  The plain smoother of the reference
 * (relax 6 as configured by relax_chunk / relax_threads) is both the
 * fallback and the judge.  Each smoothing level has up to two candidate
 * polynomials: the network's roots (apply_mlp only) and the default
 * Chebyshev roots.  From relax 6 on every smoothing level, each level in
 * turn takes its candidate with the lower measured V-cycle contraction
 * factor (jxamd_solver_vcycle_factor, 4 cycles) unless that is worse than
 * relax 6 there, in which case it stays on relax 6 (a second pass retries
 * the levels left on relax 6).  The result never contracts worse than
 * relax 6 everywhere.
 * apply_mlp: the network's roots tau_k (d_params on the device; w_k =
 * 1 / (tau_k lambda_max), the first poly_degree <= 4 outputs, each clamped
 * into the span of the training labels' roots for that degree,
 * JXAMD_MLP_LO_* / JXAMD_MLP_HI_*); *used = levels on the network's roots.
 * judge_poly: the default Chebyshev roots alone, i.e. the same method
 * without the network; *used = levels on the polynomial.
 * The final factor and that of relax 6 on every level are reported if
 * non-NULL. */
int jxamd_solver_apply_mlp(jxamd_solver *s, const float *d_params, int *used, double *factor_learned,
                           double *factor_plain);
int jxamd_solver_judge_poly(jxamd_solver *s, int *used, double *factor, double *factor_plain);
/* a level's smoother when poly_degree > 0: set plain = 1 relax 6, 0 the
 * polynomial with the steps in use; get: 1 relax 6, 0 a polynomial on the
 * default (or set_poly) steps, 2 a polynomial on the network's roots, 3 the
 * coarsest level (a direct solve) */
int jxamd_solver_set_level_smoother(jxamd_solver *s, int level, int plain);
int jxamd_solver_get_level_smoother(jxamd_solver *s, int level, int *plain);

/* Single operators on device vectors (parity tests, benchmarks). */
/* y = alpha*A*x + beta*y, jxf_BSRMatrixMatvec (seq_bsr_matvec.c:21-134) */
int jxamd_op_bsr_matvec(jxamd_solver *s, double alpha, const double *d_x, double beta, double *d_y);
/* one relax-6 call (forward + backward GS) on level l, jxf_PAMGRelax6 (par_relax_6.c:224-273) */
int jxamd_op_hsgs(jxamd_solver *s, int level, const double *d_f, double *d_u);
/* coarsest-level solve, jxf_PAMGRelax9 (par_relax_9.c:20-142) */
int jxamd_op_coarse_solve(jxamd_solver *s, const double *d_f, double *d_u);
/* one V-cycle from u (jxf_PAMGSolve with max_iter 1, par_amg_solve.c:36) */
int jxamd_op_vcycle(jxamd_solver *s, const double *d_f, double *d_u);
/* z = (LU)^{-1} r with the exact BILU(0) factors of A (stage2_solver_type 5;
 * the reference's iterative BILU: jxfpamg-0220/csrc/ilu/par_bilu_solve.c:360-366) */
int jxamd_op_bilu_solve(jxamd_solver *s, const double *d_r, double *d_z);
/* one block hybrid-SOR sweep, jxf_ParBSRHGSRelax forward (par_relax_bsr.c:193-573),
 * one-rank or partitioned as set by stage2_chunk */
int jxamd_op_bhgs(jxamd_solver *s, double *d_x, const double *d_b, double w, double omega);
/* y = A*x launched reps times back to back on the solver's stream; the mean
 * kernel time per launch from HIP events around the whole batch */
int jxamd_op_bsr_matvec_timed(jxamd_solver *s, const double *d_x, double *d_y, int reps, double *ms_per_launch);
/* one CPR application r -> x launched reps times; mean device time per apply */
int jxamd_cpr_precond_timed(jxamd_solver *s, const double *d_r, double *d_x, int reps, double *ms_per_apply);
/* Per-kernel-class device timing.  While enabled, every launch of the
 * solver is bracketed by HIP events on its stream and charged to a class
 * together with its algorithmic HBM bytes (the bytes the algorithm must move
 * once: matrix entries, indices, each vector read or written once). */
enum {
    JXAMD_PROF_GS_SWEEP = 0,   /* AMG smoother (relax-6 sweeps or polynomial steps), all levels */
    JXAMD_PROF_BHGS_SWEEP = 1, /* stage-2 block sweeps */
    JXAMD_PROF_BSR_SPMV = 2,   /* y = A x, residuals */
    JXAMD_PROF_CSR_SPMV = 3,   /* level residual, restriction, prolongation */
    JXAMD_PROF_COARSE = 4,     /* coarsest-level dense solve */
    JXAMD_PROF_MGS = 5,        /* fused Gram-Schmidt dots / norms */
    JXAMD_PROF_VEC = 6,        /* scale, axpy, linear combinations, copies */
    JXAMD_PROF_BILU = 7,       /* stage-2 BILU(0) block triangular solves */
    JXAMD_PROF_KINDS = 8
};
int jxamd_solver_profile(jxamd_solver *s, int enable); /* enable also clears the counters */
int jxamd_solver_profile_read(jxamd_solver *s, int kind, long *launches, double *ms, double *bytes);
/* the smoother launches (JXAMD_PROF_GS_SWEEP) of one level: one kernel per
 * level, so this is a per-kernel figure */
int jxamd_solver_profile_read_level(jxamd_solver *s, int level, long *launches, double *ms, double *bytes);
/* nonzero when a sync-free wait hit its spin bound since creation or since
 * the last solve that reported it (0 when healthy).  A solve during which a
 * wait gave up returns JXAMD_ERROR_GENERIC and clears the solver's
 * synchronisation state (arrival counters, flags, this flag). */
int jxamd_solver_error_flag(jxamd_solver *s, int *flag);
/* polls after which a waiting workgroup of the cooperative MGS step gives up
 * (default 2^24).  No reference counterpart: failure-path testing forces the
 * give-up from the host with a small limit, so every kernel still exits. */
int jxamd_solver_set_spin_limit(jxamd_solver *s, int limit);

/* ------------------------------------------------------------------------
 * Standalone device operations on the reference BSR layout (device arrays
 * d_ia[nb+1], d_ja[nnzb], d_a[nnzb*b*b]; no solver handle).
 * ---------------------------------------------------------------------- */
/* d_norms[0] = ||b - A x||_2, d_norms[1] = ||b||_2 (device memory, FP64),
 * b = 1..4: the true-residual check GMRES makes on convergence
 * (jxfpamg-0220/csrc/krylov/pgmres.c:1505-1511: CopyVector b -> r, Matvec
 * alpha -1 beta 1, sqrt(InnerProd(r, r)); ||b|| as at pgmres.c:1062).
 * Deterministic (fixed-order reduction).  Allocates a 33 KB work buffer per
 * call and returns after the stream has finished the launch. */
int jxamd_bsr_residual_norm(int nb, int b, const int *d_ia, const int *d_ja, const double *d_a, const double *d_x,
                            const double *d_b, double *d_norms, void *stream);
/* The same computation with no allocation and no synchronisation: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its work buffer is a static device array of 33,024
 * bytes per GPU (partial sums of at most 2048 workgroups plus a counter),
 * held from the first call on that GPU until the process ends; the counter
 * is cleared by every call.  The launch is enqueued on `stream` and the
 * call returns at once.  Calls that overlap on one GPU (several host threads
 * or streams) share the array and race.  nb*b above 2^31 - 1 is rejected
 * with JXAMD_ERROR_ARG (every smaller system fits the buffer: the grid never
 * exceeds 2048 workgroups). */
int jxamd_bsr_residual_norm_noalloc(int nb, int b, const int *d_ia, const int *d_ja, const double *d_a,
                                    const double *d_x, const double *d_b, double *d_norms, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* JXAMD_H */
