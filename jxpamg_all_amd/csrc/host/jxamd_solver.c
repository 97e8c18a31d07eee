/*
 * jxamd_solver.c -- host side of the MI355X CPR-GMRES solve phase.
 *
 * Builds the device layouts once (SELL-64 matrices, dependency-level sweep
 * orders, explicit inverses of the diagonal blocks and of the coarsest
 * matrix), then runs the reference's algorithms with every vector resident
 * in HBM:
 *   GMRES(k) with modified Gram-Schmidt     jxfpamg-0220/csrc/krylov/pgmres.c:950-1713
 *   CPR two-stage preconditioner            jxfpamg-0220/csrc/cpr/jxf_cpr.c:424-610
 *   pressure AMG V-cycle                    jxfpamg-0220/csrc/amg/par_cycle.c:20-969
 *   hybrid SSOR (relax 6), 2 per visit      jxfpamg-0220/csrc/operation/relaxations/par_relax_6.c:224-273
 *   coarsest direct solve (relax 9)         jxfpamg-0220/csrc/operation/relaxations/par_relax_9.c:20-142
 *   block hybrid SOR (stage 2)              jxfpamg-0220/csrc/amg/par_relax_bsr.c:193-573
 * The Hessenberg/Givens recurrences run on the host in FP64, one column
 * read back per Arnoldi step.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>
#include <time.h>

#include "jxamd.h"
#include "../hip/jxamd_dev.h"

#define SELLW 64
#define HSM_WORDS 8 /* pinned words after the two Arnoldi column slots */

#define CK(x)                                                                                                          \
    do {                                                                                                               \
        int e__ = (x);                                                                                                 \
        if (e__) return e__;                                                                                           \
    } while (0)

/* ---------------------------------------------------------------------- */
/* device matrix containers                                                */
/* ---------------------------------------------------------------------- */
typedef struct {
    int n, ncols, nsl;
    int *sptr, *col, *perm, *flag;
    int *dep_ptr, *dep_list; /* sweep layouts: dependency slices per slice */
    void *val, *diag;        /* val: storage type vt; diag: vector type xt */
    float *rsc;              /* vt == F16: power-of-two scale per position */
    int vt, xt;
    int G;                   /* > 1: G lanes per row over SELL-G (sellg) or plain CSR (sptr = ia, col = ja) */
    int sellg;
    int pk16;                /* col: int16 offsets column - row, two slots per word; word offsets after sptr */
    double abytes;           /* algorithmic HBM bytes of one launch over it */
} dsell;

typedef struct {
    int nb, b, nsl;
    int *sptr, *col;
    double *val;
    /* forward block-HGS order */
    int gnsl, gnpos, K;      /* K: stage-2 sweeps fused into one launch */
    int *gperm, *gsptr, *gcol, *flag, *gdep_ptr, *gdep_list;
    int *gall_ptr, *gall_list, *gvorder; /* neighbourhood slices, sweep-overlapped order */
    double *gval, *dinv, *xs;  /* xs: x^(k) of the K sweeps in slice order, then x^(-1) */
    void *val32, *dinv32;      /* FP32 (stage2_precision 1) or FP16 blocks (2; FP32 dinv) for the partitioned stage 2 */
    float *rsc16;              /* stage2_precision 2: power-of-two scale per block row */
    int *pptr16;               /* stage2_precision 2: first slot pair of each slice (blocks stored two slots per word) */
    double spmv_bytes, sweep_bytes; /* algorithmic bytes: y = A x, one fused one-rank block-HGS sweep */
    int unit_diag;             /* every diagonal block is exactly the identity: the partitioned sweeps skip Dinv */
    double part_bytes, part_bytes32; /* algorithmic bytes of one partitioned sweep (FP64, narrow blocks) */
} dbsr;

/* one BILU(0) triangle on the device: level-ordered block SELL of the
 * dependency blocks (L: j < i, U: j > i), U's inverted diagonal per row */
typedef struct {
    int nsl, vt;
    int *perm, *sptr, *col, *dep_ptr, *dep_list, *flag;
    void *val, *dinv;        /* dinv: FP32 when the blocks are FP16 */
    float *rsc;              /* FP16 blocks: power-of-two scale per position (block row) */
    double bytes;
} dtri;

/* per-kernel-class device timing (HIP events around each launch) */
#define PROF_POOL 4096
typedef struct {
    int on, n;
    void *ev0[PROF_POOL], *ev1[PROF_POOL];
    int kind[PROF_POOL], level[PROF_POOL]; /* level: smoother launches (one kernel per level), else -1 */
    double bytes[PROF_POOL];
    long count[JXAMD_PROF_KINDS];
    double ms[JXAMD_PROF_KINDS], tbytes[JXAMD_PROF_KINDS];
    long lcount[JXAMD_MAX_LEVELS];
    double lms[JXAMD_MAX_LEVELS], lbytes[JXAMD_MAX_LEVELS];
} prof_t;

struct jxamd_solver_s {
    jxamd_params prm;
    void *stream;
    void *ev0, *ev1, *ev2, *ev3;
    void *evh[2];          /* Arnoldi columns copied to hpin[slot] (completion events) */
    double *hpin;          /* pinned host memory: 2 x (k_dim + 2) doubles */
    double *hpin_dev;      /* its device address (the cooperative MGS step writes the column there) */
    double *hsm, *hsm_dev; /* pinned words after the columns (host / device address): squared norms, error flag */
    double hseq;           /* last sequence number handed to a one-launch MGS step */
    double hwant[2];       /* per slot: the sequence number its column ends with (0: wait on evh[slot]) */
    int nb, b;
    size_t N;
    dbsr A;
    int L;
    int nlev[JXAMD_MAX_LEVELS];
    dsell Ares[JXAMD_MAX_LEVELS], Gf[JXAMD_MAX_LEVELS], Gb[JXAMD_MAX_LEVELS];
    dsell Gc[JXAMD_MAX_LEVELS]; /* threaded relax-6 layout (chunked mode) */
    int *wstep[JXAMD_MAX_LEVELS];
    int nchunks[JXAMD_MAX_LEVELS]; /* 0: sequential sweeps on this level */
    dsell R[JXAMD_MAX_LEVELS], P[JXAMD_MAX_LEVELS];
    void *cinv;            /* coarsest inverse, type xt */
    int nc, ncld;          /* coarsest size, row stride of its inverse */
    int vt, xt;            /* V-cycle storage / vector types (JXAMD_T_*) */
    size_t xsz;            /* bytes per V-cycle vector entry */
    void *F[JXAMD_MAX_LEVELS], *U[JXAMD_MAX_LEVELS], *T[JXAMD_MAX_LEVELS], *Vt[JXAMD_MAX_LEVELS];
    double *up64;          /* FP64 copy of the stage-1 result when xt != F64 */
    dtri Lf, Ub;           /* stage-2 BILU(0) factors (stage2_solver_type 5) */
    void *dinvl[JXAMD_MAX_LEVELS];          /* polynomial smoother: 1/a_ii (0 if a_ii = 0), type xt */
    int plain[JXAMD_MAX_LEVELS];            /* with poly_degree > 0: 1 = this level smooths with relax 6 */
    int learned[JXAMD_MAX_LEVELS];          /* 1 = the level's polynomial uses the network's roots */
    double poly_def[JXAMD_MAX_LEVELS][8];   /* the default Chebyshev steps */
    double poly_w[JXAMD_MAX_LEVELS][8];     /* its step sizes */
    double lam[JXAMD_MAX_LEVELS];           /* lambda_max(D^{-1} A_l) estimate */
    float feat[JXAMD_MAX_LEVELS][JXAMD_MLP_NFEAT]; /* level features of the learned step sizes */
    double *bz;            /* BILU(0) intermediate vector, N */
    double *xtmp;          /* stage-2 ping-pong partner, N */
    double *V;             /* (k_dim+1) Krylov vectors, N each */
    double *Z;             /* k_dim preconditioned vectors M^{-1} v_j (solution_update 0), N each */
    double *Rv, *Wv, *Bd, *Xd; /* GMRES work, host-solve staging */
    double *dsc;           /* device scalars: h[0..k+1], y[0..k], tmp */
    double *part;
    unsigned int *count;       /* [0] k_dot_mgs, [1] k_mgs_coop arrivals */
    int mgs_grid, mgs_tm, mgs_base; /* cooperative MGS step: grid (0: off), register tile, flag base */
    int spin_limit;            /* polls before a waiting workgroup of the MGS step gives up */
    double *mpart;             /* its partial sums, (k_dim + 1) * mgs_grid */
    int *mflag;
    unsigned long long *ticket;
    unsigned long long tbase;
    int *err;
    int epoch;
    double t_prec_ms;
    prof_t prof;
};

static int prof_flush(jxamd_solver *s)
{
    prof_t *p = &s->prof;
    if (!p->n) return 0;
    CK(jxamd_stream_sync(s->stream));
    for (int q = 0; q < p->n; q++) {
        float ms = 0.f;
        CK(jxamd_event_elapsed(&ms, p->ev0[q], p->ev1[q]));
        p->count[p->kind[q]]++;
        p->ms[p->kind[q]] += ms;
        p->tbytes[p->kind[q]] += p->bytes[q];
        const int l = p->level[q];
        if (l >= 0 && l < JXAMD_MAX_LEVELS) {
            p->lcount[l]++;
            p->lms[l] += ms;
            p->lbytes[l] += p->bytes[q];
        }
    }
    p->n = 0;
    return 0;
}

/* returns the slot to close with prof_end, or -1 when profiling is off */
static int prof_begin(jxamd_solver *s)
{
    prof_t *p = &s->prof;
    if (!p->on) return -1;
    if (p->n == PROF_POOL && prof_flush(s)) return -1;
    if (jxamd_event_record(p->ev0[p->n], s->stream)) return -1;
    return p->n;
}

static void prof_end_level(jxamd_solver *s, int slot, int kind, int level, double bytes)
{
    prof_t *p = &s->prof;
    if (slot < 0) return;
    if (jxamd_event_record(p->ev1[slot], s->stream)) return;
    p->kind[slot] = kind;
    p->level[slot] = level;
    p->bytes[slot] = bytes;
    p->n++;
}

static void prof_end(jxamd_solver *s, int slot, int kind, double bytes) { prof_end_level(s, slot, kind, -1, bytes); }


void jxamd_params_default(jxamd_params *p)
{
    memset(p, 0, sizeof(*p));
    p->k_dim = 30;
    p->max_iter = 100;
    p->is_check_restarted = 1;
    p->tol = 1e-6;
    p->a_tol = 0.0;
    p->pressure_index = 0;
    p->stage1_maxit = 1;
    p->stage2_maxit = 3;
    p->stage1_solver_type = 1;
    p->stage2_solver_type = 2;
    p->print_level = 0;
    p->threshold = 1e-15;
    p->stage2_weight = 0.7;
    p->stage2_omega = 0.9;
    p->convfac_threshold = 1.0e4;
    p->arnoldi_kernel = 2;
}

/* ---------------------------------------------------------------------- */
/* host-side layout construction                                           */
/* ---------------------------------------------------------------------- */

/* Sweep schedule of one GS direction: forward (dir>0) depends on the strictly
 * lower part, backward on the strictly upper part.  Rows are sorted by
 * dependency level and every level is padded to a multiple of W positions
 * (perm = -1; W = 64 rows per slice, or 64/G for row groups of G lanes), so
 * no slice holds a row together with one of its dependencies.
 * dep_ptr/dep_list: per slice, the distinct earlier slices that hold its
 * rows' dependencies. */
typedef struct {
    int npos, nsl, nlevels;
    int *perm, *dep_ptr, *dep_list, *slev; /* slev: level of each slice */
} hsched;

static void hsched_free(hsched *h)
{
    free(h->perm); free(h->dep_ptr); free(h->dep_list); free(h->slev);
    memset(h, 0, sizeof(*h));
}

static int build_schedule_w(int n, const int *ia, const int *ja, int dir, int W, hsched *o)
{
    memset(o, 0, sizeof(*o));
    int *lev = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    if (!lev) return JXAMD_ERROR_MEMORY;
    int maxlev = -1;
    for (int t = 0; t < n; t++) {
        int i = dir > 0 ? t : n - 1 - t;
        int l = 0;
        for (int k = ia[i]; k < ia[i + 1]; k++) {
            int j = ja[k];
            if ((dir > 0 ? j < i : j > i) && lev[j] + 1 > l) l = lev[j] + 1;
        }
        lev[i] = l;
        if (l > maxlev) maxlev = l;
    }
    const int nl = maxlev + 1;
    long *start = (long *)calloc((size_t)nl + 1, sizeof(long));
    int *cnt = (int *)calloc((size_t)nl + 1, sizeof(int));
    if (!start || !cnt) { free(lev); free(start); free(cnt); return JXAMD_ERROR_MEMORY; }
    for (int i = 0; i < n; i++) cnt[lev[i]]++;
    for (int l = 0; l < nl; l++) start[l + 1] = start[l] + ((cnt[l] + W - 1) / W) * W;
    if (start[nl] > 0x7fffffffL) { free(lev); free(start); free(cnt); return JXAMD_ERROR_ARG; }
    o->npos = (int)start[nl];
    o->nsl = o->npos / W;
    o->nlevels = nl;
    o->perm = (int *)malloc(sizeof(int) * (o->npos > 0 ? o->npos : 1));
    o->slev = (int *)malloc(sizeof(int) * (o->nsl > 0 ? o->nsl : 1));
    int *sl = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    if (!o->perm || !o->slev || !sl) { free(lev); free(start); free(cnt); free(sl); return JXAMD_ERROR_MEMORY; }
    for (int l = 0; l < nl; l++)
        for (long q = start[l] / W; q < start[l + 1] / W; q++) o->slev[q] = l;
    for (int p = 0; p < o->npos; p++) o->perm[p] = -1;
    for (int l = 0; l < nl; l++) cnt[l] = 0;
    for (int i = 0; i < n; i++) { /* rows in ascending order inside a level */
        long p = start[lev[i]] + cnt[lev[i]]++;
        o->perm[p] = i;
        sl[i] = (int)(p / W);
    }
    /* dependency slices */
    o->dep_ptr = (int *)malloc(sizeof(int) * ((size_t)o->nsl + 1));
    int *mark = (int *)malloc(sizeof(int) * ((size_t)o->nsl + 1));
    size_t cap = (size_t)o->nsl * 4 + 16, used = 0;
    o->dep_list = (int *)malloc(sizeof(int) * cap);
    if (!o->dep_ptr || !mark || !o->dep_list) {
        free(lev); free(start); free(cnt); free(sl); free(mark);
        return JXAMD_ERROR_MEMORY;
    }
    for (int q = 0; q < o->nsl; q++) mark[q] = -1;
    int bad = 0;
    for (int q = 0; q < o->nsl; q++) {
        o->dep_ptr[q] = (int)used;
        for (int l = 0; l < W; l++) {
            int i = o->perm[(size_t)q * W + l];
            if (i < 0) continue;
            for (int k = ia[i]; k < ia[i + 1]; k++) {
                int j = ja[k];
                if (!(dir > 0 ? j < i : j > i)) continue;
                int d = sl[j];
                if (d >= q) bad = 1; /* cannot happen: dependencies sit on lower levels */
                if (mark[d] == q) continue;
                mark[d] = q;
                if (used == cap) {
                    cap *= 2;
                    int *t = (int *)realloc(o->dep_list, sizeof(int) * cap);
                    if (!t) { bad = 2; break; }
                    o->dep_list = t;
                }
                o->dep_list[used++] = d;
            }
        }
    }
    o->dep_ptr[o->nsl] = (int)used;
    free(lev); free(start); free(cnt); free(sl); free(mark);
    if (bad) { hsched_free(o); return bad == 2 ? JXAMD_ERROR_MEMORY : JXAMD_ERROR_GENERIC; }
    return 0;
}

static int build_schedule(int n, const int *ia, const int *ja, int dir, hsched *o)
{
    return build_schedule_w(n, ia, ja, dir, SELLW, o);
}

/* Per slice of a schedule (rows by position in perm): the distinct slices
 * holding any neighbour of its rows, and the slice itself. */
static int build_nbhd(int n, const int *ia, const int *ja, const hsched *sc, const int *perm, int **optr,
                      int **olist)
{
    const int nsl = sc->nsl;
    int *sl = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    int *mark = (int *)malloc(sizeof(int) * (nsl > 0 ? nsl : 1));
    int *ptr = (int *)malloc(sizeof(int) * ((size_t)nsl + 1));
    size_t cap = (size_t)nsl * 8 + 16, used = 0;
    int *list = (int *)malloc(sizeof(int) * cap);
    if (!sl || !mark || !ptr || !list) { free(sl); free(mark); free(ptr); free(list); return JXAMD_ERROR_MEMORY; }
    for (int p = 0; p < sc->npos; p++)
        if (perm[p] >= 0) sl[perm[p]] = p / SELLW;
    for (int q = 0; q < nsl; q++) mark[q] = -1;
    for (int q = 0; q < nsl; q++) {
        ptr[q] = (int)used;
        for (int l = -1; l < SELLW; l++) {
            int i = l < 0 ? -1 : perm[q * SELLW + l];
            if (l >= 0 && i < 0) continue;
            int k0 = l < 0 ? 0 : ia[i], k1 = l < 0 ? 1 : ia[i + 1];
            for (int k = k0; k < k1; k++) {
                int d = l < 0 ? q : sl[ja[k]];
                if (mark[d] == q) continue;
                mark[d] = q;
                if (used == cap) {
                    cap *= 2;
                    int *t = (int *)realloc(list, sizeof(int) * cap);
                    if (!t) { free(sl); free(mark); free(ptr); free(list); return JXAMD_ERROR_MEMORY; }
                    list = t;
                }
                list[used++] = d;
            }
        }
    }
    ptr[nsl] = (int)used;
    free(sl);
    free(mark);
    *optr = ptr;
    *olist = list;
    return 0;
}

/* Order of the K*nsl virtual slices (k, s) of K overlapped sweeps: by
 * slev[s] + lag*k, then k, then s, with lag = 1 + the largest level step
 * from a slice to a slice in its neighbourhood.  Every dependency of (k, s)
 * -- (k, s') for s' in dep(s) and (k-1, s'') for s'' in nbhd(s) -- comes
 * earlier; this is checked. */
static int build_vorder(const hsched *sc, const int *nb_ptr, const int *nb_list, int K, int **out)
{
    const int nsl = sc->nsl;
    int lag = 1;
    for (int q = 0; q < nsl; q++)
        for (int e = nb_ptr[q]; e < nb_ptr[q + 1]; e++)
            if (sc->slev[nb_list[e]] - sc->slev[q] + 1 > lag) lag = sc->slev[nb_list[e]] - sc->slev[q] + 1;
    const long nkeys = (long)sc->nlevels + (long)lag * (K - 1);
    const size_t nv = (size_t)nsl * K;
    long *cnt = (long *)calloc((size_t)nkeys + 1, sizeof(long));
    int *ord = (int *)malloc(sizeof(int) * (nv > 0 ? nv : 1));
    int *pos = (int *)malloc(sizeof(int) * (nv > 0 ? nv : 1));
    if (!cnt || !ord || !pos) { free(cnt); free(ord); free(pos); return JXAMD_ERROR_MEMORY; }
    for (int k = 0; k < K; k++)
        for (int q = 0; q < nsl; q++) cnt[sc->slev[q] + (long)lag * k + 1]++;
    for (long t = 0; t < nkeys; t++) cnt[t + 1] += cnt[t];
    for (int k = 0; k < K; k++) /* k outer, q inner: stable within a key */
        for (int q = 0; q < nsl; q++) {
            long at = cnt[sc->slev[q] + (long)lag * k]++;
            ord[at] = k * nsl + q;
        }
    for (size_t t = 0; t < nv; t++) pos[ord[t]] = (int)t;
    int bad = 0;
    for (int k = 0; k < K && !bad; k++)
        for (int q = 0; q < nsl && !bad; q++) {
            const int me = pos[k * nsl + q];
            for (int e = sc->dep_ptr[q]; e < sc->dep_ptr[q + 1]; e++)
                if (pos[k * nsl + sc->dep_list[e]] >= me) bad = 1;
            if (k > 0)
                for (int e = nb_ptr[q]; e < nb_ptr[q + 1]; e++)
                    if (pos[(k - 1) * nsl + nb_list[e]] >= me) bad = 1;
        }
    free(cnt);
    free(pos);
    if (bad) { free(ord); return JXAMD_ERROR_GENERIC; }
    *out = ord;
    return 0;
}

/* Scalar SELL-64 on the host.  Rows in the order given by perm (identity if
 * NULL).  offdiag: drop the first (diagonal) entry of every row into diag[].
 * pad_self: pad with the row's own index (square smoother matrices), else 0. */
typedef struct {
    int n, nsl;
    int *sptr, *col, *perm;
    double *val, *diag;
} hsell;

static void hsell_free(hsell *h)
{
    free(h->sptr); free(h->col); free(h->perm); free(h->val); free(h->diag);
    memset(h, 0, sizeof(*h));
}

/* n is the number of positions; with perm_owned, perm_owned[p] < 0 marks a
 * padding position (empty row). */
static int hsell_build(hsell *h, int n, const int *ia, const int *ja, const double *a, int *perm_owned, int offdiag,
                       int pad_self)
{
    memset(h, 0, sizeof(*h));
    h->n = n;
    h->nsl = (n + SELLW - 1) / SELLW;
    h->perm = perm_owned;
    h->sptr = (int *)malloc(sizeof(int) * (h->nsl + 1));
    if (offdiag) h->diag = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    if (!h->sptr || (offdiag && !h->diag)) return JXAMD_ERROR_MEMORY;
    const int skip = offdiag ? 1 : 0;
    long total = 0;
    h->sptr[0] = 0;
    for (int s = 0; s < h->nsl; s++) {
        int w = 0;
        for (int l = 0; l < SELLW; l++) {
            int p = s * SELLW + l;
            if (p >= n) break;
            int i = perm_owned ? perm_owned[p] : p;
            if (i < 0) continue;
            int len = ia[i + 1] - ia[i] - skip;
            if (len > w) w = len;
        }
        total += w;
        if (total > 0x7fffffffL / SELLW) return JXAMD_ERROR_ARG;
        h->sptr[s + 1] = (int)total;
    }
    h->col = (int *)malloc(sizeof(int) * (size_t)(total > 0 ? total : 1) * SELLW);
    h->val = (double *)malloc(sizeof(double) * (size_t)(total > 0 ? total : 1) * SELLW);
    if (!h->col || !h->val) return JXAMD_ERROR_MEMORY;
    for (int s = 0; s < h->nsl; s++) {
        int k0 = h->sptr[s], w = h->sptr[s + 1] - k0;
        for (int l = 0; l < SELLW; l++) {
            int p = s * SELLW + l;
            int i = -1, len = 0, beg = 0;
            if (p < n) {
                i = perm_owned ? perm_owned[p] : p;
                if (i >= 0) {
                    beg = ia[i] + skip;
                    len = ia[i + 1] - beg;
                    if (offdiag) h->diag[p] = a[ia[i]];
                } else if (offdiag) {
                    h->diag[p] = 1.0;
                }
            }
            for (int k = 0; k < w; k++) {
                size_t e = (size_t)(k0 + k) * SELLW + l;
                if (k < len) {
                    h->col[e] = ja[beg + k];
                    h->val[e] = a[beg + k];
                } else {
                    h->col[e] = (pad_self && i >= 0) ? i : 0;
                    h->val[e] = 0.0;
                }
            }
        }
    }
    return 0;
}

static int upload_ints(int **dst, const int *src, size_t n, void *st)
{
    CK(jxamd_dev_malloc((void **)dst, sizeof(int) * (n ? n : 1)));
    return jxamd_h2d(*dst, src, sizeof(int) * n, st);
}

/* Packed int16 column offsets of a square level matrix in a slice layout
 * (slices of R rows, G lanes per row; R = 64, G = 1 for SELL-64 in natural
 * order): lane ln of slot k of slice q belongs to row q*R + ln/G.  Slice q's
 * slots [sp[q], sp[q+1]) take ceil(ns/2) word rows from wp[q]; word t of a
 * lane holds the offsets of slots 2t (low half) and 2t+1 (high half).
 * Zero entries (padding) get offset 0; lanes of rows past n point at row
 * n - 1.  PK16_PAD word rows of zeros follow (the kernels load a batch's
 * words whole).  Returns 1 with *words (nw rows of 64) and *wp when every
 * offset fits in int16, else 0. */
#define PK16_PAD 8
static int pack16_slices(const int *col, const double *val, const int *sp, int nsl, int R, int G, int n,
                         int32_t **words, int **wp, size_t *nw)
{
    int *w = (int *)malloc(sizeof(int) * ((size_t)nsl + 1));
    if (!w) return 0;
    w[0] = 0;
    for (int q = 0; q < nsl; q++) w[q + 1] = w[q] + (sp[q + 1] - sp[q] + 1) / 2;
    const size_t rows = (size_t)w[nsl] + PK16_PAD;
    int32_t *o = (int32_t *)calloc(rows * SELLW, sizeof(int32_t));
    if (!o) { free(w); return 0; }
    for (int q = 0; q < nsl; q++)
        for (int k = sp[q]; k < sp[q + 1]; k++)
            for (int ln = 0; ln < SELLW; ln++) {
                const size_t e = (size_t)k * SELLW + ln;
                const int row = q * R + ln / G;
                long off;
                if (row >= n) off = (long)(n - 1) - row;
                else if (val[e] == 0.0) off = 0;
                else off = (long)col[e] - row;
                if (off < -32768 || off > 32767) { free(o); free(w); return 0; }
                const int t = k - sp[q];
                const size_t we = ((size_t)w[q] + t / 2) * SELLW + ln;
                const uint32_t h = (uint32_t)(uint16_t)(int16_t)off;
                o[we] = (int32_t)((uint32_t)o[we] | (t % 2 ? h << 16 : h));
            }
    *words = o;
    *wp = w;
    *nw = rows;
    return 1;
}

/* replace d->sptr / d->col by [sptr | word offsets] and the packed words;
 * h16 (FP16 layouts: the values' bits, slot-major as uploaded): the values
 * are stored again as pairs in the words' order (slot 2t of a slice in the
 * low half of word t, slot 2t + 1 in the high half, zero past the slice's
 * last slot), so the kernels load them whole with the column words */
static int upload_pk16(dsell *d, const int *sp, int nsl, const int32_t *o, const int *wp, size_t rows, void *st,
                       const uint16_t *h16)
{
    if (h16) {
        uint16_t *pv = (uint16_t *)calloc(rows * SELLW * 2, sizeof(uint16_t));
        if (!pv) return JXAMD_ERROR_MEMORY;
        for (int q = 0; q < nsl; q++)
            for (int k = sp[q]; k < sp[q + 1]; k++) {
                const int t = k - sp[q];
                for (int ln = 0; ln < SELLW; ln++)
                    pv[(((size_t)wp[q] + t / 2) * SELLW + ln) * 2 + (t & 1)] = h16[(size_t)k * SELLW + ln];
            }
        void *dv = NULL;
        int rc = jxamd_dev_malloc(&dv, sizeof(uint32_t) * rows * SELLW);
        if (!rc) rc = jxamd_h2d(dv, pv, sizeof(uint32_t) * rows * SELLW, st);
        if (!rc) rc = jxamd_stream_sync(st);
        free(pv);
        if (rc) { jxamd_dev_free(dv); return rc; }
        jxamd_dev_free(d->val);
        d->val = dv;
    }
    int *sp2 = (int *)malloc(sizeof(int) * 2 * ((size_t)nsl + 1));
    if (!sp2) return JXAMD_ERROR_MEMORY;
    memcpy(sp2, sp, sizeof(int) * ((size_t)nsl + 1));
    memcpy(sp2 + nsl + 1, wp, sizeof(int) * ((size_t)nsl + 1));
    int *dsp = NULL, *dcol = NULL;
    int rc = upload_ints(&dsp, sp2, 2 * ((size_t)nsl + 1), st);
    if (!rc) rc = upload_ints(&dcol, (const int *)o, rows * SELLW, st);
    if (!rc) rc = jxamd_stream_sync(st);
    free(sp2);
    if (rc) { jxamd_dev_free(dsp); jxamd_dev_free(dcol); return rc; }
    jxamd_dev_free(d->sptr);
    jxamd_dev_free(d->col);
    d->sptr = dsp;
    d->col = dcol;
    d->pk16 = 1;
    return 0;
}

static int zero_ints(int **dst, size_t n, void *st)
{
    int *z = (int *)calloc(n ? n : 1, sizeof(int));
    if (!z) return JXAMD_ERROR_MEMORY;
    int e = jxamd_dev_malloc((void **)dst, sizeof(int) * (n ? n : 1));
    if (!e) e = jxamd_h2d(*dst, z, sizeof(int) * (n ? n : 1), st);
    if (!e) e = jxamd_stream_sync(st);
    free(z);
    return e;
}

/* sched != NULL: a sweep layout; uploads its dependency lists and a zeroed
 * per-slice flag array */
/* IEEE binary16 bits of f, round to nearest even (values here are scaled
 * into [-1, 1], so overflow cannot occur; tiny values become subnormal or 0) */
static uint16_t half_bits(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* rounds past 65504 */
    if (ax < 0x38800000u) { /* below 2^-14: subnormal half */
        if (ax < 0x33000000u) return (uint16_t)sign; /* below 2^-25 */
        const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
        const int shift = 126 - (int)(ax >> 23); /* 14..24 */
        uint32_t q = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (q & 1u))) q++;
        return (uint16_t)(sign | q);
    }
    uint32_t h = ((ax >> 13) - (112u << 10));
    const uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

static size_t type_size(int t) { return t == JXAMD_T_F64 ? 8 : (t == JXAMD_T_F32 ? 4 : 2); }

/* algorithmic bytes of one pass over a level matrix: values (+ F16 row
 * scales), indices, row pointer, plus nvec vector entries per row */
/* storage code of a level matrix for its kernels (with the packed-index flag) */
static int lvt(const dsell *d) { return d->vt | (d->pk16 ? JXAMD_T_PK16 : 0); }

static double level_bytes(const dsell *d, long nnz, int n, double nvec)
{
    double b = (double)nnz * (type_size(d->vt) + (d->pk16 ? 2.0 : 4.0)) + 4.0 * (n + 1) + nvec * type_size(d->xt) * n;
    if (d->vt == JXAMD_T_F16) b += 4.0 * n;
    return b;
}

/* Upload a host SELL layout, values converted to storage type vt and the
 * diagonal to vector type xt. */
static int dsell_upload(dsell *d, const hsell *h, int ncols, void *st, const hsched *sched, int vt, int xt,
                        int pack)
{
    memset(d, 0, sizeof(*d));
    uint16_t *h16 = NULL; /* FP16 values' bits, kept for the packed layout */
    d->n = h->n;
    d->ncols = ncols;
    d->nsl = h->nsl;
    d->vt = vt;
    d->xt = xt;
    size_t slots = (size_t)(h->nsl ? h->sptr[h->nsl] : 0) * SELLW;
    CK(jxamd_dev_malloc((void **)&d->sptr, sizeof(int) * (h->nsl + 1)));
    CK(jxamd_dev_malloc((void **)&d->col, sizeof(int) * (slots ? slots : 1)));
    CK(jxamd_dev_malloc(&d->val, type_size(vt) * (slots ? slots : 1)));
    CK(jxamd_h2d(d->sptr, h->sptr, sizeof(int) * (h->nsl + 1), st));
    CK(jxamd_h2d(d->col, h->col, sizeof(int) * slots, st));
    if (vt == JXAMD_T_F64) {
        CK(jxamd_h2d(d->val, h->val, sizeof(double) * slots, st));
    } else {
        void *tmp = malloc(type_size(vt) * (slots ? slots : 1));
        float *sc = NULL;
        if (!tmp) return JXAMD_ERROR_MEMORY;
        if (vt == JXAMD_T_F32) {
            for (size_t e = 0; e < slots; e++) ((float *)tmp)[e] = (float)h->val[e];
        } else {
            const size_t npos = (size_t)h->nsl * SELLW;
            sc = (float *)malloc(sizeof(float) * (npos ? npos : 1));
            if (!sc) { free(tmp); return JXAMD_ERROR_MEMORY; }
            for (int q = 0; q < h->nsl; q++)
                for (int l = 0; l < SELLW; l++) {
                    double mx = 0.0;
                    for (int k = h->sptr[q]; k < h->sptr[q + 1]; k++) {
                        double v = fabs(h->val[(size_t)k * SELLW + l]);
                        if (v > mx) mx = v;
                    }
                    /* smallest power of two >= the row's largest magnitude */
                    int ex = 0;
                    if (mx > 0.0) {
                        double m = frexp(mx, &ex); /* mx = m * 2^ex, m in [0.5, 1) */
                        if (m == 0.5) ex -= 1;
                    }
                    const float scale = mx > 0.0 ? (float)ldexp(1.0, ex) : 1.0f;
                    sc[(size_t)q * SELLW + l] = scale;
                    for (int k = h->sptr[q]; k < h->sptr[q + 1]; k++) {
                        size_t e = (size_t)k * SELLW + l;
                        ((uint16_t *)tmp)[e] = half_bits((float)(h->val[e] / scale));
                    }
                }
            int rc = jxamd_dev_malloc((void **)&d->rsc, sizeof(float) * (npos ? npos : 1));
            if (!rc) rc = jxamd_h2d(d->rsc, sc, sizeof(float) * npos, st);
            if (rc) { free(tmp); free(sc); return rc; }
        }
        int rc = jxamd_h2d(d->val, tmp, type_size(vt) * slots, st);
        if (!rc) rc = jxamd_stream_sync(st);
        free(sc);
        if (rc || vt != JXAMD_T_F16) free(tmp);
        else h16 = (uint16_t *)tmp;
        CK(rc);
    }
    if (pack && !h->perm && ncols == h->n) {
        int32_t *o = NULL;
        int *wp = NULL;
        size_t rows = 0;
        if (pack16_slices(h->col, h->val, h->sptr, h->nsl, SELLW, 1, h->n, &o, &wp, &rows)) {
            const int rc = upload_pk16(d, h->sptr, h->nsl, o, wp, rows, st, h16);
            free(o);
            free(wp);
            if (rc) free(h16);
            CK(rc);
        }
    }
    free(h16);
    if (h->perm) {
        CK(jxamd_dev_malloc((void **)&d->perm, sizeof(int) * (h->n ? h->n : 1)));
        CK(jxamd_h2d(d->perm, h->perm, sizeof(int) * h->n, st));
    }
    if (h->diag) {
        CK(jxamd_dev_malloc(&d->diag, type_size(xt) * (h->n ? h->n : 1)));
        if (xt == JXAMD_T_F64) {
            CK(jxamd_h2d(d->diag, h->diag, sizeof(double) * h->n, st));
        } else {
            float *t = (float *)malloc(sizeof(float) * (h->n ? h->n : 1));
            if (!t) return JXAMD_ERROR_MEMORY;
            for (int p = 0; p < h->n; p++) t[p] = (float)h->diag[p];
            int rc = jxamd_h2d(d->diag, t, sizeof(float) * h->n, st);
            if (!rc) rc = jxamd_stream_sync(st);
            free(t);
            CK(rc);
        }
    }
    if (sched) {
        CK(upload_ints(&d->dep_ptr, sched->dep_ptr, (size_t)sched->nsl + 1, st));
        CK(upload_ints(&d->dep_list, sched->dep_list, (size_t)sched->dep_ptr[sched->nsl], st));
        CK(zero_ints(&d->flag, (size_t)sched->nsl, st));
    }
    return jxamd_stream_sync(st);
}

/* lanes per row for a matrix of mean row length nnz/n: one (SELL-64) for
 * short rows, else the smallest power of two from 4 to 32 that leaves each
 * lane at most one batch of the kernels' 8 in-flight entries of a mean row */
static int choose_group(long nnz, int n)
{
    const double m = n > 0 ? (double)nnz / n : 0.0;
    if (m < 12.0) return 1;
    int G = 4;
    while (G < 32 && G * 8 < m) G *= 2;
    return G;
}

/* Upload a CSR matrix for the row-group kernels in the SELL-G layout: a
 * slice holds R = 64/G consecutive rows, one group of G lanes per row; slot
 * row u of the slice holds entries u*G .. u*G+G-1 of every row (row order,
 * the diagonal first), so lane g of a group sums entries g, g+G, ... and a
 * wavefront's load of one slot row is one contiguous transaction.  A slice
 * has as many slot rows as its longest row needs; short rows are padded with
 * a zero value and a valid column (the row itself when it is a column).
 * F16 rows are scaled by the smallest power of two at or above their largest
 * off-diagonal magnitude (the whole row if it has no off-diagonal).
 * Used for the residual, restriction, prolongation and polynomial-smoother
 * operators.  Square matrices whose couplings lie within 32767 rows of the
 * diagonal get packed int16 column offsets (pack16_slices). */
static int dsellg_upload(dsell *d, int n, int ncols, const int *ia, const int *ja, const double *a, int vt, int xt,
                         int G, void *st)
{
    memset(d, 0, sizeof(*d));
    d->n = n;
    d->ncols = ncols;
    d->vt = vt;
    d->xt = xt;
    d->G = G;
    d->sellg = 1;
    if (G < 2 || G > SELLW || SELLW % G) return JXAMD_ERROR_ARG;
    const int R = SELLW / G;
    const int nsl = (n + R - 1) / R;
    d->nsl = nsl;
    int *sp = (int *)malloc(sizeof(int) * ((size_t)nsl + 1));
    if (!sp) return JXAMD_ERROR_MEMORY;
    sp[0] = 0;
    for (int q = 0; q < nsl; q++) {
        int steps = 0;
        for (int i = q * R; i < n && i < (q + 1) * R; i++) {
            const int t = (ia[i + 1] - ia[i] + G - 1) / G;
            if (t > steps) steps = t;
        }
        sp[q + 1] = sp[q] + steps;
    }
    const size_t tot = (size_t)sp[nsl] * SELLW;
    int *cl = (int *)malloc(sizeof(int) * (tot ? tot : 1));
    double *vl = (double *)calloc(tot ? tot : 1, sizeof(double));
    float *sc = vt == JXAMD_T_F16 ? (float *)malloc(sizeof(float) * (n > 0 ? n : 1)) : NULL;
    if (!cl || !vl || (vt == JXAMD_T_F16 && !sc)) { free(sp); free(cl); free(vl); free(sc); return JXAMD_ERROR_MEMORY; }
    for (int q = 0; q < nsl; q++)
        for (int u = sp[q]; u < sp[q + 1]; u++)
            for (int ln = 0; ln < SELLW; ln++) {
                const int i = q * R + ln / G;
                cl[(size_t)u * SELLW + ln] = (i < n && i < ncols) ? i : 0;
            }
    for (int i = 0; i < n; i++) {
        const int q = i / R, lr = i % R;
        double scale = 1.0;
        if (sc) {
            double mx = 0.0;
            const int k0 = ia[i + 1] - ia[i] > 1 ? ia[i] + 1 : ia[i];
            for (int k = k0; k < ia[i + 1]; k++)
                if (fabs(a[k]) > mx) mx = fabs(a[k]);
            int ex = 0;
            if (mx > 0.0) {
                double m = frexp(mx, &ex);
                if (m == 0.5) ex -= 1;
            }
            scale = mx > 0.0 ? ldexp(1.0, ex) : 1.0;
            sc[i] = (float)scale;
        }
        for (int pp = 0; pp < ia[i + 1] - ia[i]; pp++) {
            const size_t e = (size_t)(sp[q] + pp / G) * SELLW + (size_t)lr * G + pp % G;
            cl[e] = ja[ia[i] + pp];
            vl[e] = sc ? a[ia[i] + pp] / scale : a[ia[i] + pp];
        }
    }
    int rc = upload_ints(&d->sptr, sp, (size_t)nsl + 1, st);
    if (!rc) rc = upload_ints(&d->col, cl, tot, st);
    if (!rc) rc = jxamd_dev_malloc(&d->val, type_size(vt) * (tot ? tot : 1));
    void *tmp = NULL;
    if (!rc && vt != JXAMD_T_F64) {
        tmp = malloc(type_size(vt) * (tot ? tot : 1));
        if (!tmp) rc = JXAMD_ERROR_MEMORY;
        else if (vt == JXAMD_T_F32)
            for (size_t k = 0; k < tot; k++) ((float *)tmp)[k] = (float)vl[k];
        else
            for (size_t k = 0; k < tot; k++) ((uint16_t *)tmp)[k] = half_bits((float)vl[k]);
    }
    if (!rc) rc = jxamd_h2d(d->val, tmp ? tmp : (void *)vl, type_size(vt) * tot, st);
    if (!rc && sc) {
        rc = jxamd_dev_malloc((void **)&d->rsc, sizeof(float) * (n > 0 ? n : 1));
        if (!rc) rc = jxamd_h2d(d->rsc, sc, sizeof(float) * n, st);
    }
    if (!rc) rc = jxamd_stream_sync(st);
    int32_t *o = NULL;
    int *wp = NULL;
    size_t rows = 0;
    if (!rc && n == ncols && pack16_slices(cl, vl, sp, nsl, R, G, n, &o, &wp, &rows))
        rc = upload_pk16(d, sp, nsl, o, wp, rows, st, vt == JXAMD_T_F16 ? (const uint16_t *)tmp : NULL);
    free(o); free(wp);
    free(sp); free(cl); free(vl); free(sc); free(tmp);
    CK(rc);
    return 0;
}

/* Upload a CSR matrix as it is (sptr = ia, col = ja) for the chunked
 * row-group smoother, whose blocks cover chunk-aligned row ranges; F16 rows
 * are scaled as in dsellg_upload; want_diag: also the diagonal (first entry
 * of each row) as a vector of type xt. */
static int dcsr_upload(dsell *d, int n, int ncols, const int *ia, const int *ja, const double *a, int vt, int xt,
                       int G, int want_diag, void *st)
{
    memset(d, 0, sizeof(*d));
    d->n = n;
    d->ncols = ncols;
    d->vt = vt;
    d->xt = xt;
    d->G = G;
    const size_t nnz = (size_t)ia[n];
    CK(upload_ints(&d->sptr, ia, (size_t)n + 1, st));
    CK(upload_ints(&d->col, ja, nnz, st));
    CK(jxamd_dev_malloc(&d->val, type_size(vt) * (nnz ? nnz : 1)));
    if (vt == JXAMD_T_F64) {
        CK(jxamd_h2d(d->val, a, sizeof(double) * nnz, st));
    } else {
        void *tmp = malloc(type_size(vt) * (nnz ? nnz : 1));
        float *sc = NULL;
        if (!tmp) return JXAMD_ERROR_MEMORY;
        if (vt == JXAMD_T_F32) {
            for (size_t k = 0; k < nnz; k++) ((float *)tmp)[k] = (float)a[k];
        } else {
            sc = (float *)malloc(sizeof(float) * (n > 0 ? n : 1));
            if (!sc) { free(tmp); return JXAMD_ERROR_MEMORY; }
            for (int i = 0; i < n; i++) {
                double mx = 0.0;
                const int k0 = ia[i + 1] - ia[i] > 1 ? ia[i] + 1 : ia[i];
                for (int k = k0; k < ia[i + 1]; k++)
                    if (fabs(a[k]) > mx) mx = fabs(a[k]);
                int ex = 0;
                if (mx > 0.0) {
                    double m = frexp(mx, &ex);
                    if (m == 0.5) ex -= 1;
                }
                const float scale = mx > 0.0 ? (float)ldexp(1.0, ex) : 1.0f;
                sc[i] = scale;
                for (int k = ia[i]; k < ia[i + 1]; k++) ((uint16_t *)tmp)[k] = half_bits((float)(a[k] / scale));
            }
            int rc = jxamd_dev_malloc((void **)&d->rsc, sizeof(float) * (n > 0 ? n : 1));
            if (!rc) rc = jxamd_h2d(d->rsc, sc, sizeof(float) * n, st);
            if (rc) { free(tmp); free(sc); return rc; }
        }
        int rc = jxamd_h2d(d->val, tmp, type_size(vt) * nnz, st);
        if (!rc) rc = jxamd_stream_sync(st);
        free(tmp);
        free(sc);
        CK(rc);
    }
    if (want_diag) {
        double *dg = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
        if (!dg) return JXAMD_ERROR_MEMORY;
        for (int i = 0; i < n; i++) dg[i] = a[ia[i]];
        int rc = jxamd_dev_malloc(&d->diag, type_size(xt) * (n > 0 ? n : 1));
        if (!rc) {
            if (xt == JXAMD_T_F64) rc = jxamd_h2d(d->diag, dg, sizeof(double) * n, st);
            else {
                float *f = (float *)dg; /* in place: float k lies before double k */
                for (int i = 0; i < n; i++) f[i] = (float)dg[i];
                rc = jxamd_h2d(d->diag, f, sizeof(float) * n, st);
            }
        }
        if (!rc) rc = jxamd_stream_sync(st);
        free(dg);
        CK(rc);
    }
    return jxamd_stream_sync(st);
}

static void dsell_free(dsell *d)
{
    jxamd_dev_free(d->sptr); jxamd_dev_free(d->col); jxamd_dev_free(d->val);
    jxamd_dev_free(d->perm); jxamd_dev_free(d->diag); jxamd_dev_free(d->flag);
    jxamd_dev_free(d->dep_ptr); jxamd_dev_free(d->dep_list); jxamd_dev_free(d->rsc);
    memset(d, 0, sizeof(*d));
}

/* In-place inverse of a dense n x n row-major matrix by Gauss-Jordan with
 * partial pivoting.  Returns nonzero if a pivot is exactly zero. */
static int dense_inverse(double *M, int n)
{
    int *piv = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    if (!piv) return JXAMD_ERROR_MEMORY;
    for (int i = 0; i < n; i++) piv[i] = i;
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = fabs(M[(size_t)k * n + k]);
        for (int i = k + 1; i < n; i++) {
            double v = fabs(M[(size_t)i * n + k]);
            if (v > best) { best = v; p = i; }
        }
        if (best == 0.0) { free(piv); return JXAMD_ERROR_GENERIC; }
        if (p != k) {
            for (int j = 0; j < n; j++) {
                double t = M[(size_t)k * n + j];
                M[(size_t)k * n + j] = M[(size_t)p * n + j];
                M[(size_t)p * n + j] = t;
            }
            int t = piv[k]; piv[k] = piv[p]; piv[p] = t;
        }
        double d = 1.0 / M[(size_t)k * n + k];
        M[(size_t)k * n + k] = 1.0;
        for (int j = 0; j < n; j++) M[(size_t)k * n + j] *= d;
        /* the row updates of one pivot step are independent: host threads
         * share them on large coarsest levels (same arithmetic per entry) */
#pragma omp parallel for schedule(static) if (n >= 512)
        for (int i = 0; i < n; i++) {
            if (i == k) continue;
            double f = M[(size_t)i * n + k];
            if (f == 0.0) continue;
            M[(size_t)i * n + k] = 0.0;
            for (int j = 0; j < n; j++) M[(size_t)i * n + j] -= f * M[(size_t)k * n + j];
        }
    }
    /* undo the row permutation as a column permutation */
    double *tmp = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    if (!tmp) { free(piv); return JXAMD_ERROR_MEMORY; }
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) tmp[piv[j]] = M[(size_t)i * n + j];
        memcpy(M + (size_t)i * n, tmp, sizeof(double) * n);
    }
    free(tmp);
    free(piv);
    return 0;
}

/* The stage-2 block solve v -> D^{-1} v as the reference computes it,
 * jxf_BSRMatrixBlockInvMatvec_Stable (seq_bsr_matrix.c:1054-1097), written
 * out as a b x b matrix: the routine is linear in v for a fixed block, so
 * column j of M is its image of the unit vector e_j.  b = 2: closed form,
 * or v / (a_ii + 1e-12) componentwise when the block is ill-conditioned
 * (:1067-1078).  b >= 3 (_MixedPrec, :948-1051): partial pivoting LU; when a
 * pivot column is below 1e-15 the elimination stops there (the diagonal
 * estimate formed at :985-991 is overwritten by the substitutions that
 * follow), and the substitutions then run on the partly eliminated block,
 * giving 0 for every component whose pivot is below 1e-15. */
static void block_solve_matrix(const double *D, int b, double *M)
{
    for (int col = 0; col < b; col++) {
        double m[JXAMD_MAX_BLOCK * JXAMD_MAX_BLOCK], v[JXAMD_MAX_BLOCK], o[JXAMD_MAX_BLOCK];
        for (int i = 0; i < b; i++) v[i] = i == col ? 1.0 : 0.0;
        if (b == 2) {
            const double a11 = D[0], a12 = D[1], a21 = D[2], a22 = D[3];
            const double det = a11 * a22 - a12 * a21;
            if (fabs(det) < 1e-12 * (fabs(a11 * a22) + fabs(a12 * a21))) {
                o[0] = v[0] / (a11 + 1e-12);
                o[1] = v[1] / (a22 + 1e-12);
            } else {
                const double inv_det = 1.0 / det;
                o[0] = (a22 * v[0] - a12 * v[1]) * inv_det;
                o[1] = (-a21 * v[0] + a11 * v[1]) * inv_det;
            }
        } else {
            memcpy(m, D, sizeof(double) * b * b);
            for (int i = 0; i < b; i++) {
                int mr = i;
                double mv = fabs(m[i * b + i]);
                for (int k = i + 1; k < b; k++)
                    if (fabs(m[k * b + i]) > mv) { mv = fabs(m[k * b + i]); mr = k; }
                if (mv < 1e-15) break;
                if (mr != i) {
                    for (int j = 0; j < b; j++) { double t = m[i * b + j]; m[i * b + j] = m[mr * b + j]; m[mr * b + j] = t; }
                    double t = v[i]; v[i] = v[mr]; v[mr] = t;
                }
                const double piv = m[i * b + i];
                for (int k = i + 1; k < b; k++) {
                    const double f = m[k * b + i] / piv;
                    m[k * b + i] = f;
                    for (int j = i + 1; j < b; j++) m[k * b + j] -= f * m[i * b + j];
                }
            }
            for (int i = 0; i < b; i++) {
                o[i] = v[i];
                for (int j = 0; j < i; j++) o[i] -= m[i * b + j] * o[j];
            }
            for (int i = b - 1; i >= 0; i--) {
                for (int j = i + 1; j < b; j++) o[i] -= m[i * b + j] * o[j];
                const double dg = m[i * b + i];
                o[i] = fabs(dg) > 1e-15 ? o[i] / dg : 0.0;
            }
        }
        for (int r = 0; r < b; r++) M[r * b + col] = o[r];
    }
}

/* transpose of a CSR matrix (rows of R sorted by the source row, which keeps
 * the accumulation order of jxf_CSRMatrixMatvecT, seq_csr_matvec.c:211-330) */
static int csr_transpose(int n, int m, const int *ia, const int *ja, const double *a, int **tia, int **tja,
                         double **ta)
{
    int nnz = ia[n];
    int *ri = (int *)calloc((size_t)m + 1, sizeof(int));
    int *rj = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    double *ra = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    if (!ri || !rj || !ra) { free(ri); free(rj); free(ra); return JXAMD_ERROR_MEMORY; }
    for (int k = 0; k < nnz; k++) ri[ja[k] + 1]++;
    for (int c = 0; c < m; c++) ri[c + 1] += ri[c];
    int *pos = (int *)malloc(sizeof(int) * ((size_t)m + 1));
    if (!pos) { free(ri); free(rj); free(ra); return JXAMD_ERROR_MEMORY; }
    memcpy(pos, ri, sizeof(int) * ((size_t)m + 1));
    for (int i = 0; i < n; i++)
        for (int k = ia[i]; k < ia[i + 1]; k++) {
            int q = pos[ja[k]]++;
            rj[q] = i;
            ra[q] = a[k];
        }
    free(pos);
    *tia = ri; *tja = rj; *ta = ra;
    return 0;
}

static int check_csr(const jxamd_csr *c, int square)
{
    if (!c->ia || !c->ja || !c->a || c->n < 0 || c->ncols < 0) return JXAMD_ERROR_ARG;
    if (c->ia[0] != 0) return JXAMD_ERROR_ARG;
    for (int i = 0; i < c->n; i++) {
        if (c->ia[i + 1] < c->ia[i]) return JXAMD_ERROR_ARG;
        for (int k = c->ia[i]; k < c->ia[i + 1]; k++)
            if (c->ja[k] < 0 || c->ja[k] >= c->ncols) return JXAMD_ERROR_ARG;
        if (square && (c->ia[i + 1] == c->ia[i] || c->ja[c->ia[i]] != i)) return JXAMD_ERROR_ARG;
    }
    return 0;
}

/* upload n doubles as type vt (F64 or F32) */
static int upload_real(void **dst, const double *src, size_t n, int vt, void *st)
{
    CK(jxamd_dev_malloc(dst, type_size(vt) * (n ? n : 1)));
    if (vt == JXAMD_T_F64) return jxamd_h2d(*dst, src, sizeof(double) * n, st);
    float *t = (float *)malloc(sizeof(float) * (n ? n : 1));
    if (!t) return JXAMD_ERROR_MEMORY;
    for (size_t k = 0; k < n; k++) t[k] = (float)src[k];
    int rc = jxamd_h2d(*dst, t, sizeof(float) * n, st);
    if (!rc) rc = jxamd_stream_sync(st);
    free(t);
    return rc;
}

/* storage type of the partitioned stage-2 blocks for stage2_precision p */
static int stage2_vt(int p) { return p == 2 ? JXAMD_T_F16 : (p == 1 ? JXAMD_T_F32 : JXAMD_T_F64); }

/* FP16 copy of the stage-2 SELL blocks v (slot-major, b*b planes of 64
 * lanes): block row i is scaled by the smallest power of two at or above the
 * largest magnitude of its off-diagonal blocks, so every stored entry is at
 * most 1 in magnitude; the scales go to d->rsc16 */
static int upload_blocks_f16(dbsr *d, const hsell *h, const double *v, int nb, int bb, void *st)
{
    /* slice s's slots k0..k1-1 are stored in pairs (k0 + 2p, k0 + 2p + 1),
     * element e of both in one 32-bit word (low half the first slot; a
     * zero high half past the slice's last slot), pairs of slice s from
     * pptr[s] on: word ((pptr[s] + p) * bb + e) * 64 + lane */
    const int nsl = h->nsl;
    int *pp = (int *)malloc(sizeof(int) * ((size_t)nsl + 1));
    float *sc = (float *)malloc(sizeof(float) * (nb > 0 ? nb : 1));
    if (!pp || !sc) { free(pp); free(sc); return JXAMD_ERROR_MEMORY; }
    pp[0] = 0;
    for (int sl = 0; sl < nsl; sl++) pp[sl + 1] = pp[sl] + (h->sptr[sl + 1] - h->sptr[sl] + 1) / 2;
    const size_t pairs = (size_t)pp[nsl];
    uint32_t *t = (uint32_t *)calloc(pairs * bb * SELLW + 1, sizeof(uint32_t));
    if (!t) { free(pp); free(sc); return JXAMD_ERROR_MEMORY; }
    for (int i = 0; i < nb; i++) {
        const int sl = i / SELLW, l = i % SELLW;
        double mx = 0.0;
        for (int k = h->sptr[sl]; k < h->sptr[sl + 1]; k++) {
            if (h->col[(size_t)k * SELLW + l] == i) continue; /* the diagonal block and padding */
            for (int e = 0; e < bb; e++) {
                const double a = fabs(v[((size_t)k * bb + e) * SELLW + l]);
                if (a > mx) mx = a;
            }
        }
        int ex = 0;
        if (mx > 0.0) {
            double m = frexp(mx, &ex);
            if (m == 0.5) ex -= 1;
        }
        const double scale = mx > 0.0 ? ldexp(1.0, ex) : 1.0;
        sc[i] = (float)scale;
        for (int k = h->sptr[sl]; k < h->sptr[sl + 1]; k++) {
            if (h->col[(size_t)k * SELLW + l] == i) continue;
            const int kk = k - h->sptr[sl];
            for (int e = 0; e < bb; e++) {
                const uint32_t hb = half_bits((float)(v[((size_t)k * bb + e) * SELLW + l] / scale));
                t[(((size_t)pp[sl] + kk / 2) * bb + e) * SELLW + l] |= hb << (16 * (kk & 1));
            }
        }
    }
    int rc = jxamd_dev_malloc(&d->val32, sizeof(uint32_t) * (pairs * bb * SELLW + 1));
    if (!rc) rc = jxamd_h2d(d->val32, t, sizeof(uint32_t) * pairs * bb * SELLW, st);
    if (!rc) rc = jxamd_dev_malloc((void **)&d->rsc16, sizeof(float) * (nb > 0 ? nb : 1));
    if (!rc) rc = jxamd_h2d(d->rsc16, sc, sizeof(float) * nb, st);
    if (!rc) rc = upload_ints(&d->pptr16, pp, (size_t)nsl + 1, st);
    if (!rc) rc = jxamd_stream_sync(st);
    free(pp);
    free(sc);
    free(t);
    return rc;
}

/* ---------------------------------------------------------------------- */
/* create / destroy                                                        */
/* ---------------------------------------------------------------------- */
static int build_bsr(jxamd_solver *s, const jxamd_bsr *A)
{
    const int nb = A->nb, b = A->b, bb = b * b;
    dbsr *d = &s->A;
    d->nb = nb;
    d->b = b;
    /* every diagonal block exactly the identity (any ABF-decoupled system):
     * the SpMV layout then leaves them out (y_i starts from x_i, which is
     * A_ii x_i to the bit) and the partitioned sweeps skip Dinv */
    d->unit_diag = 1;
    for (int i = 0; i < nb && d->unit_diag; i++)
        for (int r = 0; r < b; r++)
            for (int c2 = 0; c2 < b; c2++)
                if (A->a[(size_t)A->ia[i] * bb + r * b + c2] != (r == c2 ? 1.0 : 0.0)) d->unit_diag = 0;
    const int skip = d->unit_diag;
    /* natural-order SELL over all blocks but the identity diagonal ones (SpMV) */
    {
        hsell h;
        jxamd_csr pat = {nb, nb, A->nnzb, A->ia, A->ja, NULL};
        /* reuse the scalar builder for the pattern, then place block values */
        double *dummy = (double *)calloc((size_t)(A->nnzb > 0 ? A->nnzb : 1), sizeof(double));
        if (!dummy) return JXAMD_ERROR_MEMORY;
        int e = hsell_build(&h, nb, pat.ia, pat.ja, dummy, NULL, skip, 1); /* padding: own row, zero block */
        free(dummy);
        if (e) { hsell_free(&h); return e; }
        size_t slots = (size_t)(h.nsl ? h.sptr[h.nsl] : 0);
        double *v = (double *)calloc(slots * bb * SELLW + 1, sizeof(double));
        if (!v) { hsell_free(&h); return JXAMD_ERROR_MEMORY; }
        for (int sl = 0; sl < h.nsl; sl++) {
            int k0 = h.sptr[sl], w = h.sptr[sl + 1] - k0;
            for (int l = 0; l < SELLW; l++) {
                int i = sl * SELLW + l;
                if (i >= nb) break;
                int len = A->ia[i + 1] - A->ia[i] - skip;
                for (int k = 0; k < w && k < len; k++) {
                    const double *blk = A->a + (size_t)(A->ia[i] + skip + k) * bb;
                    for (int e2 = 0; e2 < bb; e2++) v[((size_t)(k0 + k) * bb + e2) * SELLW + l] = blk[e2];
                }
            }
        }
        d->nsl = h.nsl;
        /* bytes the SpMV reads and writes: the stored blocks (without the
         * identity diagonal ones) with their column indices, row pointers,
         * x once, y once */
        d->spmv_bytes = (double)(A->nnzb - (skip ? nb : 0)) * (8.0 * bb + 4.0) + 4.0 * (nb + 1) + 16.0 * nb * b;
        /* one fused one-rank sweep (k_bhgs_multi): off-diagonal blocks and
         * their columns, row pointers, x read once, rhs read, x written, Dinv */
        d->sweep_bytes = (double)(A->nnzb - nb) * (8.0 * bb + 4.0) + 4.0 * (nb + 1) + 24.0 * nb * b + 8.0 * nb * bb;
        int rc = jxamd_dev_malloc((void **)&d->sptr, sizeof(int) * (h.nsl + 1));
        if (!rc) rc = jxamd_dev_malloc((void **)&d->col, sizeof(int) * (slots * SELLW + 1));
        if (!rc) rc = jxamd_dev_malloc((void **)&d->val, sizeof(double) * (slots * bb * SELLW + 1));
        if (!rc) rc = jxamd_h2d(d->sptr, h.sptr, sizeof(int) * (h.nsl + 1), s->stream);
        if (!rc) rc = jxamd_h2d(d->col, h.col, sizeof(int) * slots * SELLW, s->stream);
        if (!rc) rc = jxamd_h2d(d->val, v, sizeof(double) * slots * bb * SELLW, s->stream);
        if (!rc && s->prm.stage2_chunk > 0 && s->prm.stage2_precision == 1)
            rc = upload_real(&d->val32, v, slots * bb * SELLW, JXAMD_T_F32, s->stream);
        if (!rc && s->prm.stage2_chunk > 0 && s->prm.stage2_precision == 2)
            rc = upload_blocks_f16(d, &h, v, nb, bb, s->stream);
        if (!rc) rc = jxamd_stream_sync(s->stream);
        free(v);
        hsell_free(&h);
        CK(rc);
    }
    /* forward block-HGS order: off-diagonal blocks in level order + Dinv */
    {
        hsched sc;
        CK(build_schedule(nb, A->ia, A->ja, 1, &sc));
        int *perm = sc.perm;
        const int npos = sc.npos;
        sc.perm = NULL; /* owned by h below */
        hsell h;
        double *dummy = (double *)calloc((size_t)(A->nnzb > 0 ? A->nnzb : 1), sizeof(double));
        if (!dummy) { free(perm); hsched_free(&sc); return JXAMD_ERROR_MEMORY; }
        int e = hsell_build(&h, npos, A->ia, A->ja, dummy, perm, 1, 1);
        free(dummy);
        if (e) { hsell_free(&h); hsched_free(&sc); return e; }
        size_t slots = (size_t)(h.nsl ? h.sptr[h.nsl] : 0);
        double *v = (double *)calloc(slots * bb * SELLW + 1, sizeof(double));
        double *dinv = (double *)malloc(sizeof(double) * (size_t)nb * bb + 8);
        if (!v || !dinv) { free(v); free(dinv); hsell_free(&h); return JXAMD_ERROR_MEMORY; }
        for (int sl = 0; sl < h.nsl; sl++) {
            int k0 = h.sptr[sl], w = h.sptr[sl + 1] - k0;
            for (int l = 0; l < SELLW; l++) {
                int p = sl * SELLW + l;
                int i = perm[p];
                if (i < 0) continue;
                int beg = A->ia[i] + 1, len = A->ia[i + 1] - beg;
                for (int k = 0; k < w && k < len; k++) {
                    const double *blk = A->a + (size_t)(beg + k) * bb;
                    for (int e2 = 0; e2 < bb; e2++) v[((size_t)(k0 + k) * bb + e2) * SELLW + l] = blk[e2];
                }
            }
        }
        for (int i = 0; i < nb; i++) block_solve_matrix(A->a + (size_t)A->ia[i] * bb, b, dinv + (size_t)i * bb);
        /* one partitioned sweep (k_bhgs_part) reads the SpMV layout (every
         * stored block but the identity diagonal ones, with their column
         * indices), row pointers, x once, rhs, writes x, and reads Dinv
         * unless every diagonal block is the identity */
        {
            const double nblk = (double)A->nnzb - (d->unit_diag ? (double)nb : 0.0);
            const double common = 4.0 * nblk + 4.0 * (nb + 1) + 24.0 * nb * b;
            const double sv = s->prm.stage2_precision == 2 ? 2.0 : 4.0;
            d->part_bytes = common + nblk * 8.0 * bb + (d->unit_diag ? 0.0 : 8.0 * nb * bb);
            d->part_bytes32 = common + nblk * sv * bb + (d->unit_diag ? 0.0 : 4.0 * nb * bb) +
                              (s->prm.stage2_precision == 2 ? 4.0 * nb : 0.0);
        }
        /* columns as positions: ~pos for lower neighbours, the own position for padding */
        {
            int *inv = (int *)malloc(sizeof(int) * (size_t)nb);
            if (!inv) { free(v); free(dinv); hsell_free(&h); hsched_free(&sc); return JXAMD_ERROR_MEMORY; }
            for (int q = 0; q < npos; q++)
                if (perm[q] >= 0) inv[perm[q]] = q;
            for (int sl = 0; sl < h.nsl; sl++)
                for (int k = h.sptr[sl]; k < h.sptr[sl + 1]; k++)
                    for (int l = 0; l < SELLW; l++) {
                        const int q = sl * SELLW + l, i = perm[q];
                        int *c = &h.col[(size_t)k * SELLW + l];
                        if (i < 0 || *c == i) *c = q;
                        else *c = *c < i ? ~inv[*c] : inv[*c];
                    }
            free(inv);
        }
        d->gnsl = h.nsl;
        d->gnpos = npos;
        int rc = jxamd_dev_malloc((void **)&d->gsptr, sizeof(int) * (h.nsl + 1));
        if (!rc) rc = jxamd_dev_malloc((void **)&d->gcol, sizeof(int) * (slots * SELLW + 1));
        if (!rc) rc = jxamd_dev_malloc((void **)&d->gval, sizeof(double) * (slots * bb * SELLW + 1));
        if (!rc) rc = jxamd_dev_malloc((void **)&d->dinv, sizeof(double) * ((size_t)nb * bb + 1));
        if (!rc) rc = jxamd_h2d(d->gsptr, h.sptr, sizeof(int) * (h.nsl + 1), s->stream);
        if (!rc) rc = jxamd_h2d(d->gcol, h.col, sizeof(int) * slots * SELLW, s->stream);
        if (!rc) rc = jxamd_h2d(d->gval, v, sizeof(double) * slots * bb * SELLW, s->stream);
        if (!rc) rc = upload_ints(&d->gperm, perm, (size_t)npos, s->stream);
        if (!rc) rc = jxamd_h2d(d->dinv, dinv, sizeof(double) * (size_t)nb * bb, s->stream);
        if (!rc && s->prm.stage2_chunk > 0 && s->prm.stage2_precision >= 1)
            rc = upload_real(&d->dinv32, dinv, (size_t)nb * bb, JXAMD_T_F32, s->stream);
        if (!rc) rc = upload_ints(&d->gdep_ptr, sc.dep_ptr, (size_t)sc.nsl + 1, s->stream);
        if (!rc) rc = upload_ints(&d->gdep_list, sc.dep_list, (size_t)sc.dep_ptr[sc.nsl], s->stream);
        d->K = s->prm.stage2_maxit > 1 ? s->prm.stage2_maxit : 1;
        if (!rc) rc = jxamd_dev_malloc((void **)&d->xs, sizeof(double) * (size_t)(d->K + 1) * npos * b);
        if (!rc) rc = zero_ints(&d->flag, (size_t)sc.nsl * d->K, s->stream);
        if (!rc && d->K > 1) {
            int *nptr = NULL, *nlist = NULL, *ord = NULL;
            rc = build_nbhd(nb, A->ia, A->ja, &sc, perm, &nptr, &nlist);
            if (!rc) rc = build_vorder(&sc, nptr, nlist, d->K, &ord);
            if (!rc) rc = upload_ints(&d->gall_ptr, nptr, (size_t)sc.nsl + 1, s->stream);
            if (!rc) rc = upload_ints(&d->gall_list, nlist, (size_t)nptr[sc.nsl], s->stream);
            if (!rc) rc = upload_ints(&d->gvorder, ord, (size_t)sc.nsl * d->K, s->stream);
            if (!rc) rc = jxamd_stream_sync(s->stream);
            free(nptr);
            free(nlist);
            free(ord);
        }
        if (!rc) rc = jxamd_stream_sync(s->stream);
        free(v);
        free(dinv);
        hsell_free(&h); /* frees perm */
        hsched_free(&sc);
        CK(rc);
    }
    return 0;
}

/* FP16 copy of one BILU(0) triangle in its slice layout (v: slot-major, b*b
 * planes of 64 lanes): position p's blocks are scaled by the smallest power
 * of two at or above their largest magnitude, which goes to t->rsc[p];
 * elements 2w and 2w+1 of a block share 32-bit word w (low half first, a zero
 * high half past the last element), slot k's word w of lane l at
 * (k * nw + w) * 64 + l, nw = ceil(b*b / 2) */
static int upload_tri_f16(dtri *t, const hsell *h, const double *v, int bb, void *st)
{
    const size_t slots = (size_t)(h->nsl ? h->sptr[h->nsl] : 0);
    const size_t npos = (size_t)h->nsl * SELLW;
    const int nw = (bb + 1) / 2;
    uint16_t *hv = (uint16_t *)calloc(slots * nw * 2 * SELLW + 1, sizeof(uint16_t));
    float *sc = (float *)malloc(sizeof(float) * (npos ? npos : 1));
    if (!hv || !sc) { free(hv); free(sc); return JXAMD_ERROR_MEMORY; }
    for (int q = 0; q < h->nsl; q++)
        for (int l = 0; l < SELLW; l++) {
            double mx = 0.0;
            for (int k = h->sptr[q]; k < h->sptr[q + 1]; k++)
                for (int e = 0; e < bb; e++) {
                    const double a = fabs(v[((size_t)k * bb + e) * SELLW + l]);
                    if (a > mx) mx = a;
                }
            int ex = 0;
            if (mx > 0.0) {
                const double m = frexp(mx, &ex); /* mx = m * 2^ex, m in [0.5, 1) */
                if (m == 0.5) ex -= 1;
            }
            const double scale = mx > 0.0 ? ldexp(1.0, ex) : 1.0;
            sc[(size_t)q * SELLW + l] = (float)scale;
            for (int k = h->sptr[q]; k < h->sptr[q + 1]; k++)
                for (int e = 0; e < bb; e++) {
                    const size_t x = ((size_t)k * bb + e) * SELLW + l;
                    const size_t word = ((size_t)k * nw + e / 2) * SELLW + l;
                    hv[2 * word + (e & 1)] = half_bits((float)(v[x] / scale));
                }
        }
    int rc = jxamd_dev_malloc(&t->val, sizeof(uint32_t) * (slots * nw * SELLW + 1));
    if (!rc) rc = jxamd_h2d(t->val, hv, sizeof(uint32_t) * slots * nw * SELLW, st);
    if (!rc) rc = jxamd_dev_malloc((void **)&t->rsc, sizeof(float) * (npos ? npos : 1));
    if (!rc) rc = jxamd_h2d(t->rsc, sc, sizeof(float) * npos, st);
    if (!rc) rc = jxamd_stream_sync(st);
    free(hv);
    free(sc);
    return rc;
}

static void dtri_free(dtri *t)
{
    jxamd_dev_free(t->perm); jxamd_dev_free(t->sptr); jxamd_dev_free(t->col); jxamd_dev_free(t->dep_ptr);
    jxamd_dev_free(t->dep_list); jxamd_dev_free(t->flag); jxamd_dev_free(t->val); jxamd_dev_free(t->dinv);
    jxamd_dev_free(t->rsc);
    memset(t, 0, sizeof(*t));
}

/* one triangle of the BILU(0) factors: dir > 0 the unit lower L (forward
 * level order), dir < 0 the upper U with its inverted diagonal (backward).
 * Storage: stage2_precision 1 FP32, 2 FP16 with each block row scaled by the
 * smallest power of two at or above the largest magnitude of its blocks in
 * this triangle (FP32 inverted diagonals), 0 FP64 (FP32 when stage 1 runs
 * narrow, precond_precision > 0). */
static int build_tri(jxamd_solver *s, const jxamd_bsr *A, const double *lu, const double *dinv, int dir, dtri *t)
{
    const int nb = A->nb, b = A->b, bb = b * b;
    memset(t, 0, sizeof(*t));
    const int sp = s->prm.stage2_precision;
    t->vt = sp == 2 ? JXAMD_T_F16 : (sp == 1 || s->prm.precond_precision > 0) ? JXAMD_T_F32 : JXAMD_T_F64;
    int *fia = (int *)malloc(sizeof(int) * ((size_t)nb + 1));
    int *fja = (int *)malloc(sizeof(int) * (size_t)(A->nnzb > 0 ? A->nnzb : 1));
    int *fidx = (int *)malloc(sizeof(int) * (size_t)(A->nnzb > 0 ? A->nnzb : 1));
    double *dummy = (double *)calloc((size_t)(A->nnzb > 0 ? A->nnzb : 1), sizeof(double));
    if (!fia || !fja || !fidx || !dummy) { free(fia); free(fja); free(fidx); free(dummy); return JXAMD_ERROR_MEMORY; }
    long nf = 0;
    fia[0] = 0;
    for (int i = 0; i < nb; i++) {
        for (int k = A->ia[i]; k < A->ia[i + 1]; k++)
            if (dir > 0 ? A->ja[k] < i : A->ja[k] > i) { fja[nf] = A->ja[k]; fidx[nf] = k; nf++; }
        fia[i + 1] = (int)nf;
    }
    hsched sc;
    int e = build_schedule(nb, A->ia, A->ja, dir, &sc);
    hsell h;
    memset(&h, 0, sizeof(h));
    int *perm = NULL;
    if (!e) {
        perm = sc.perm;
        sc.perm = NULL; /* owned by h */
        e = hsell_build(&h, sc.npos, fia, fja, dummy, perm, 0, 1);
    }
    double *v = NULL;
    size_t slots = 0;
    if (!e) {
        slots = (size_t)(h.nsl ? h.sptr[h.nsl] : 0);
        v = (double *)calloc(slots * bb * SELLW + 1, sizeof(double));
        if (!v) e = JXAMD_ERROR_MEMORY;
    }
    if (!e) {
        for (int q = 0; q < h.nsl; q++) {
            const int k0 = h.sptr[q], w = h.sptr[q + 1] - k0;
            for (int l = 0; l < SELLW; l++) {
                const int i = perm[q * SELLW + l];
                if (i < 0) continue;
                const int beg = fia[i], len = fia[i + 1] - beg;
                for (int k = 0; k < w && k < len; k++) {
                    const double *blk = lu + (size_t)fidx[beg + k] * bb;
                    for (int e2 = 0; e2 < bb; e2++) v[((size_t)(k0 + k) * bb + e2) * SELLW + l] = blk[e2];
                }
            }
        }
        t->nsl = h.nsl;
        e = upload_ints(&t->perm, perm, (size_t)sc.npos, s->stream);
        if (!e) e = upload_ints(&t->sptr, h.sptr, (size_t)h.nsl + 1, s->stream);
        if (!e) e = upload_ints(&t->col, h.col, slots * SELLW, s->stream);
        if (!e && t->vt == JXAMD_T_F16) e = upload_tri_f16(t, &h, v, bb, s->stream);
        else if (!e) e = upload_real(&t->val, v, slots * bb * SELLW, t->vt, s->stream);
        if (!e && dir < 0)
            e = upload_real(&t->dinv, dinv, (size_t)nb * bb, t->vt == JXAMD_T_F16 ? JXAMD_T_F32 : t->vt, s->stream);
        if (!e) e = upload_ints(&t->dep_ptr, sc.dep_ptr, (size_t)sc.nsl + 1, s->stream);
        if (!e) e = upload_ints(&t->dep_list, sc.dep_list, (size_t)sc.dep_ptr[sc.nsl], s->stream);
        if (!e) e = zero_ints(&t->flag, (size_t)sc.nsl, s->stream);
        if (!e) e = jxamd_stream_sync(s->stream);
        /* dependency blocks + indices, rhs read, result written, dependency
         * values read; U: inverted diagonals; FP16: a scale per block row */
        const double sd = t->vt == JXAMD_T_F16 ? 4.0 : (double)type_size(t->vt);
        const double sb = t->vt == JXAMD_T_F16 ? 4.0 * ((bb + 1) / 2) : (double)type_size(t->vt) * bb;
        t->bytes = (double)nf * (sb + 4.0) + 24.0 * nb * b + (dir < 0 ? sd * nb * bb : 0.0) +
                   (t->vt == JXAMD_T_F16 ? 4.0 * nb : 0.0);
    }
    free(v);
    hsell_free(&h); /* frees perm */
    hsched_free(&sc);
    free(fia); free(fja); free(fidx); free(dummy);
    return e;
}

static int build_bilu(jxamd_solver *s, const jxamd_bsr *A)
{
    const size_t bb = (size_t)A->b * A->b;
    double *lu = (double *)malloc(sizeof(double) * (size_t)A->nnzb * bb);
    double *dinv = (double *)malloc(sizeof(double) * (size_t)A->nb * bb);
    if (!lu || !dinv) { free(lu); free(dinv); return JXAMD_ERROR_MEMORY; }
    int e = jxamd_bilu0_factor(A->nb, A->b, A->ia, A->ja, A->a, lu, dinv);
    if (!e) e = build_tri(s, A, lu, dinv, 1, &s->Lf);
    if (!e) e = build_tri(s, A, lu, dinv, -1, &s->Ub);
    if (!e) e = jxamd_dev_malloc((void **)&s->bz, sizeof(double) * s->N);
    free(lu);
    free(dinv);
    return e;
}

/* level l's transfer operators (P, R = P^T) or, on the coarsest level, the
 * explicit inverse */
static int build_transfer(jxamd_solver *s, int l, const jxamd_hierarchy *H)
{
    const jxamd_csr *A = &H->A[l];
    const int n = A->n;
    hsell h;
    int e;
    if (l < s->L - 1) {
        /* interpolation P (n x nc) and restriction R = P^T (nc x n) */
        const jxamd_csr *Pm = &H->P[l];
        const int GP = choose_group(Pm->ia[Pm->n], Pm->n);
        if (GP > 1) {
            e = dsellg_upload(&s->P[l], Pm->n, Pm->ncols, Pm->ia, Pm->ja, Pm->a, s->vt, s->xt, GP, s->stream);
            memset(&h, 0, sizeof(h));
        } else {
            e = hsell_build(&h, Pm->n, Pm->ia, Pm->ja, Pm->a, NULL, 0, 0);
            if (!e) e = dsell_upload(&s->P[l], &h, Pm->ncols, s->stream, NULL, s->vt, s->xt, 0);
        }
        /* y += P x: x once, y read and written */
        s->P[l].abytes = level_bytes(&s->P[l], Pm->ia[Pm->n], Pm->n, 2.0) + (double)type_size(s->xt) * Pm->ncols;
        if (!e) e = jxamd_stream_sync(s->stream);
        hsell_free(&h);
        CK(e);
        int *ri, *rj;
        double *ra;
        CK(csr_transpose(Pm->n, Pm->ncols, Pm->ia, Pm->ja, Pm->a, &ri, &rj, &ra));
        const int GR = choose_group(ri[Pm->ncols], Pm->ncols);
        if (GR > 1) {
            e = dsellg_upload(&s->R[l], Pm->ncols, Pm->n, ri, rj, ra, s->vt, s->xt, GR, s->stream);
            memset(&h, 0, sizeof(h));
        } else {
            e = hsell_build(&h, Pm->ncols, ri, rj, ra, NULL, 0, 0);
            if (!e) e = dsell_upload(&s->R[l], &h, Pm->n, s->stream, NULL, s->vt, s->xt, 0);
        }
        s->R[l].abytes = level_bytes(&s->R[l], Pm->ia[Pm->n], Pm->ncols, 1.0) + (double)type_size(s->xt) * Pm->n;
        if (!e) e = jxamd_stream_sync(s->stream);
        hsell_free(&h);
        free(ri); free(rj); free(ra);
        CK(e);
    } else if (s->L > 1) {
        /* coarsest: explicit inverse of the densified matrix */
        double *M = (double *)calloc((size_t)n * n + 1, sizeof(double));
        if (!M) return JXAMD_ERROR_MEMORY;
        for (int i = 0; i < n; i++)
            for (int k = A->ia[i]; k < A->ia[i + 1]; k++) M[(size_t)i * n + A->ja[k]] = A->a[k];
        e = dense_inverse(M, n);
        /* rows padded to ld entries (a multiple of four, zeros beyond n) so
         * the GEMV reads 16-byte vectors */
        const int ld = (n + 3) & ~3;
        void *Mp = e ? NULL : calloc((size_t)ld * n + 1, s->xsz);
        if (!e && !Mp) e = JXAMD_ERROR_MEMORY;
        if (!e) {
            for (int i = 0; i < n; i++)
                for (int j = 0; j < n; j++) {
                    const double v = M[(size_t)i * n + j];
                    if (s->xt == JXAMD_T_F64) ((double *)Mp)[(size_t)i * ld + j] = v;
                    else ((float *)Mp)[(size_t)i * ld + j] = (float)v;
                }
        }
        if (!e) e = jxamd_dev_malloc(&s->cinv, s->xsz * ((size_t)ld * n + 1));
        if (!e) e = jxamd_h2d(s->cinv, Mp, s->xsz * (size_t)ld * n, s->stream);
        if (!e) e = jxamd_stream_sync(s->stream);
        free(M);
        free(Mp);
        s->nc = n;
        s->ncld = ld;
        CK(e);
    }
    return 0;
}


/* polynomial smoother data of level l: 1/a_ii, a lambda_max(D^{-1} A)
 * estimate by 20 power iterations from a fixed start vector, and default
 * Chebyshev step sizes on [0.3, 1.1] x lambda */
static int build_poly(jxamd_solver *s, int l, const jxamd_hierarchy *H)
{
    const jxamd_csr *A = &H->A[l];
    const int n = A->n;
    double *di = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *v = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *w = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    if (!di || !v || !w) { free(di); free(v); free(w); return JXAMD_ERROR_MEMORY; }
    for (int i = 0; i < n; i++) {
        const double d = A->a[A->ia[i]];
        di[i] = d != 0.0 ? 1.0 / d : 0.0;
    }
    uint64_t st = 0x9e3779b97f4a7c15ull + (uint64_t)l;
    for (int i = 0; i < n; i++) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        v[i] = (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
    double lam = 1.0;
    for (int it = 0; it < 20; it++) {
        double nv = 0.0, nw = 0.0;
        for (int i = 0; i < n; i++) {
            double acc = 0.0;
            for (int k = A->ia[i]; k < A->ia[i + 1]; k++) acc += A->a[k] * v[A->ja[k]];
            w[i] = di[i] * acc;
            nv += v[i] * v[i];
            nw += w[i] * w[i];
        }
        if (nw == 0.0 || nv == 0.0) break;
        lam = sqrt(nw / nv);
        const double sc = 1.0 / sqrt(nw);
        for (int i = 0; i < n; i++) v[i] = w[i] * sc;
    }
    s->lam[l] = lam;
    /* features of the level for the learned step sizes */
    {
        double gmax = 0.0, rsum = 0.0, rmin = 1e300, rmean = 0.0, pos = 0.0, noff = 0.0, lsum = 0.0, lsq = 0.0,
               rsratio = 0.0;
        int maxrow = 0;
        for (int i = 0; i < n; i++) {
            const double d = A->a[A->ia[i]];
            double off = 0.0, sum = 0.0;
            for (int k = A->ia[i]; k < A->ia[i + 1]; k++) {
                sum += A->a[k];
                if (k == A->ia[i]) continue;
                off += fabs(A->a[k]);
                pos += A->a[k] > 0.0;
                noff += 1.0;
            }
            const double ad = fabs(d) > 0.0 ? fabs(d) : 1.0;
            const double ratio = off / ad;
            if (1.0 + ratio > gmax) gmax = 1.0 + ratio;
            if (ratio < rmin) rmin = ratio;
            rmean += ratio;
            rsratio += sum / (d != 0.0 ? d : 1.0);
            const double lg = log10(ad);
            lsum += lg;
            lsq += lg * lg;
            if (A->ia[i + 1] - A->ia[i] > maxrow) maxrow = A->ia[i + 1] - A->ia[i];
        }
        (void)rsum;
        const double nn = n > 0 ? (double)n : 1.0;
        const double lmean = lsum / nn, lvar = lsq / nn - lmean * lmean;
        const int last = s->L == 1 ? 0 : s->L - 2;
        float *f = s->feat[l];
        f[0] = (float)(l / 10.0);
        f[1] = (float)(log10(nn) / 7.0);
        f[2] = (float)(log10(nn / (double)H->A[0].n) / 3.0);
        f[3] = (float)((double)A->ia[n] / nn / 30.0);
        f[4] = (float)(maxrow / 100.0);
        f[5] = (float)(lam / 2.0);
        f[6] = (float)(gmax / 2.0);
        f[7] = (float)(rmean / nn);
        f[8] = (float)(n > 0 ? rmin : 0.0);
        f[9] = (float)(noff > 0.0 ? pos / noff : 0.0);
        f[10] = (float)(sqrt(lvar > 0.0 ? lvar : 0.0) / 5.0);
        f[11] = (float)(l + 1 < s->L ? (double)H->A[l + 1].n / nn : 0.0);
        f[12] = (float)(l == 0);
        f[13] = (float)(l == last);
        f[14] = (float)(rsratio / nn);
        f[15] = 1.0f;
    }
    const int m = s->prm.poly_degree;
    const double lo = 0.3 * lam, hi = 1.1 * lam, c = 0.5 * (hi + lo), d = 0.5 * (hi - lo);
    for (int k = 0; k < m; k++) s->poly_w[l][k] = s->poly_def[l][k] = 1.0 / (c + d * cos(M_PI * (2.0 * k + 1.0) / (2.0 * m)));
    int e;
    if (s->xt == JXAMD_T_F64) {
        e = jxamd_dev_malloc(&s->dinvl[l], sizeof(double) * (n > 0 ? n : 1));
        if (!e) e = jxamd_h2d(s->dinvl[l], di, sizeof(double) * n, s->stream);
    } else {
        float *f = (float *)v; /* reuse: float view of the scratch vector */
        for (int i = 0; i < n; i++) f[i] = (float)di[i];
        e = jxamd_dev_malloc(&s->dinvl[l], sizeof(float) * (n > 0 ? n : 1));
        if (!e) e = jxamd_h2d(s->dinvl[l], f, sizeof(float) * n, s->stream);
    }
    if (!e) e = jxamd_stream_sync(s->stream);
    free(di); free(v); free(w);
    return e;
}

static int build_level(jxamd_solver *s, int l, const jxamd_hierarchy *H)
{
    const jxamd_csr *A = &H->A[l];
    const int n = A->n;
    hsell h;
    int e;
    /* natural order, full rows: residual F - A U (and the polynomial smoother) */
    const int GA = choose_group(A->ia[n], n);
    if (GA > 1) {
        CK(dsellg_upload(&s->Ares[l], n, n, A->ia, A->ja, A->a, s->vt, s->xt, GA, s->stream));
    } else {
        e = hsell_build(&h, n, A->ia, A->ja, A->a, NULL, 0, 0);
        if (!e) e = dsell_upload(&s->Ares[l], &h, n, s->stream, NULL, s->vt, s->xt, 1);
        if (!e) e = jxamd_stream_sync(s->stream);
        hsell_free(&h);
        CK(e);
    }
    /* CSR SpMV bytes: nnz (value + index) + row pointer + x + rhs + y */
    s->Ares[l].abytes = level_bytes(&s->Ares[l], A->ia[n], n, 3.0);
    /* the polynomial smoother, and relax 6 beside it (the plain smoother a
     * level falls back to) */
    if (s->prm.poly_degree > 0 && (l < s->L - 1 || s->L == 1)) CK(build_poly(s, l, H));
    int T = 0;
    if (s->prm.relax_threads > 1) T = s->prm.relax_threads;
    else if (s->prm.relax_chunk > 0) T = (n + s->prm.relax_chunk - 1) / s->prm.relax_chunk;
    if (T > n) T = n;
    if (T > 1 && (l < s->L - 1 || s->L == 1) && jxamd_gs_chunk_lds_rows(n, T) <= JXAMD_CHUNK_LDS_MAX) {
        /* short chunks: one pass, off-diagonals in natural row order (SELL),
         * or full CSR rows read by row groups */
        if (GA > 1) {
            CK(dcsr_upload(&s->Gc[l], n, n, A->ia, A->ja, A->a, s->vt, s->xt, GA, 1, s->stream));
        } else {
            e = hsell_build(&h, n, A->ia, A->ja, A->a, NULL, 1, 1);
            if (!e) e = dsell_upload(&s->Gc[l], &h, n, s->stream, NULL, s->vt, s->xt, 1);
            if (!e) e = jxamd_stream_sync(s->stream);
            hsell_free(&h);
            CK(e);
        }
        s->nchunks[l] = T;
        s->wstep[l] = NULL;
        /* matrix once, f + entry values + results */
        s->Gc[l].abytes = level_bytes(&s->Gc[l], A->ia[n], n, 3.0);
    } else if (T > 1 && (l < s->L - 1 || s->L == 1)) {
        /* long chunks: wave w holds chunks 64w..64w+63; slice wstep[w]+t
         * holds row ns(c)+t of each chunk c of the wave */
        const int size = n / T, rest = n - size * T;
        const int waves = (T + SELLW - 1) / SELLW;
        int *wstep = (int *)malloc(sizeof(int) * ((size_t)waves + 1));
        if (!wstep) return JXAMD_ERROR_MEMORY;
        wstep[0] = 0;
        for (int w = 0; w < waves; w++) wstep[w + 1] = wstep[w] + size + (w * SELLW < rest ? 1 : 0);
        const int npos = wstep[waves] * SELLW;
        int *perm = (int *)malloc(sizeof(int) * (npos > 0 ? npos : 1));
        if (!perm) { free(wstep); return JXAMD_ERROR_MEMORY; }
        for (int w = 0; w < waves; w++)
            for (int t = 0; t < wstep[w + 1] - wstep[w]; t++)
                for (int ln = 0; ln < SELLW; ln++) {
                    int c = w * SELLW + ln, row = -1;
                    if (c < T) {
                        int ns = c < rest ? c * size + c : c * size + rest;
                        int len = size + (c < rest ? 1 : 0);
                        if (t < len) row = ns + t;
                    }
                    perm[(size_t)(wstep[w] + t) * SELLW + ln] = row;
                }
        e = hsell_build(&h, npos, A->ia, A->ja, A->a, perm, 1, 1);
        if (!e) e = dsell_upload(&s->Gc[l], &h, n, s->stream, NULL, s->vt, s->xt, 0);
        if (!e) e = upload_ints(&s->wstep[l], wstep, (size_t)waves + 1, s->stream);
        if (!e) e = jxamd_stream_sync(s->stream);
        hsell_free(&h);
        free(wstep);
        CK(e);
        s->nchunks[l] = T;
        /* forward + backward pass over the matrix, old + new + rhs vectors */
        s->Gc[l].abytes = 2.0 * level_bytes(&s->Gc[l], A->ia[n], n, 0.0) + 3.0 * type_size(s->xt) * n;
    } else if ((l < s->L - 1 || s->L == 1) && GA > 1) {
        /* long rows: sweep orders with slices of 64/GA rows of one dependency
         * level, each row summed by a group of GA lanes over its plain CSR
         * row (k_gs_sweep_grp) */
        for (int dir = 1; dir >= -1; dir -= 2) {
            dsell *d = dir > 0 ? &s->Gf[l] : &s->Gb[l];
            hsched sc;
            CK(build_schedule_w(n, A->ia, A->ja, dir, SELLW / GA, &sc));
            e = dcsr_upload(d, n, n, A->ia, A->ja, A->a, s->vt, s->xt, GA, 1, s->stream);
            if (!e) e = upload_ints(&d->perm, sc.perm, (size_t)sc.npos, s->stream);
            if (!e) e = upload_ints(&d->dep_ptr, sc.dep_ptr, (size_t)sc.nsl + 1, s->stream);
            if (!e) e = upload_ints(&d->dep_list, sc.dep_list, (size_t)sc.dep_ptr[sc.nsl], s->stream);
            if (!e) e = zero_ints(&d->flag, (size_t)sc.nsl, s->stream);
            if (!e) e = jxamd_stream_sync(s->stream);
            d->nsl = sc.nsl;
            hsched_free(&sc);
            CK(e);
            /* sweep: matrix once + old x + new x + rhs */
            d->abytes = level_bytes(d, A->ia[n], n, 3.0);
        }
    } else if (l < s->L - 1 || s->L == 1) {
        /* forward and backward sweep orders, off-diagonal entries */
        for (int dir = 1; dir >= -1; dir -= 2) {
            hsched sc;
            CK(build_schedule(n, A->ia, A->ja, dir, &sc));
            int *perm = sc.perm;
            sc.perm = NULL; /* owned by h below */
            e = hsell_build(&h, sc.npos, A->ia, A->ja, A->a, perm, 1, 1);
            if (!e) e = dsell_upload(dir > 0 ? &s->Gf[l] : &s->Gb[l], &h, n, s->stream, &sc, s->vt, s->xt, 0);
            hsell_free(&h);
            hsched_free(&sc);
            CK(e);
            /* sweep: matrix once + old x + new x + rhs */
            (dir > 0 ? &s->Gf[l] : &s->Gb[l])->abytes = level_bytes(dir > 0 ? &s->Gf[l] : &s->Gb[l], A->ia[n], n, 3.0);
        }
    }
    return build_transfer(s, l, H);
}

int jxamd_solver_create(jxamd_solver **out, const jxamd_bsr *A, const jxamd_hierarchy *H, const jxamd_params *p)
{
    if (!out || !A || !H || !p) return JXAMD_ERROR_ARG;
    *out = NULL;
    if (A->b < 2 || A->b > JXAMD_MAX_BLOCK || A->nb < 1 || A->ncols != A->nb) return JXAMD_ERROR_ARG;
    if (H->num_levels < 1 || H->num_levels > JXAMD_MAX_LEVELS) return JXAMD_ERROR_ARG;
    if (p->k_dim < 1 || p->pressure_index < 0 || p->pressure_index >= A->b) return JXAMD_ERROR_ARG;
    if (p->stage1_maxit < 0 || p->stage2_maxit < 0) return JXAMD_ERROR_ARG;
    if (p->relax_threads < 0 || p->relax_chunk < 0) return JXAMD_ERROR_ARG;
    if (p->precond_precision < 0 || p->precond_precision > 2) return JXAMD_ERROR_ARG;
    if (p->stage2_solver_type != 2 && p->stage2_solver_type != 5) return JXAMD_ERROR_ARG;
    if (p->poly_degree < 0 || p->poly_degree > 8) return JXAMD_ERROR_ARG;
    if (p->stage2_chunk < 0 || p->stage2_chunk > 8) return JXAMD_ERROR_ARG;
    if (p->stage2_precision < 0 || p->stage2_precision > 2) return JXAMD_ERROR_ARG;
    if (p->arnoldi_kernel < 0 || p->arnoldi_kernel > 2) return JXAMD_ERROR_ARG;
    if (p->solution_update < 0 || p->solution_update > 1) return JXAMD_ERROR_ARG;
    if (!(p->convfac_threshold > 0.0)) return JXAMD_ERROR_ARG;
    if (p->coarse_rows < 0 || p->coarse_rows > 8192) return JXAMD_ERROR_ARG;
    {
        jxamd_csr pat = {A->nb, A->nb, A->nnzb, A->ia, A->ja, A->a};
        if (A->ia[A->nb] != A->nnzb) return JXAMD_ERROR_ARG;
        CK(check_csr(&pat, 1));
    }
    if (H->A[0].n != A->nb) return JXAMD_ERROR_ARG;
    for (int l = 0; l < H->num_levels; l++) {
        if (H->A[l].ncols != H->A[l].n) return JXAMD_ERROR_ARG;
        CK(check_csr(&H->A[l], 1));
        if (l < H->num_levels - 1) {
            if (H->P[l].n != H->A[l].n || H->P[l].ncols != H->A[l + 1].n) return JXAMD_ERROR_ARG;
            CK(check_csr(&H->P[l], 0));
        }
    }
    int ndev = 0;
    jxamd_dev_count(&ndev);
    if (ndev < 1) return JXAMD_ERROR_DEVICE;

    jxamd_solver *s = (jxamd_solver *)calloc(1, sizeof(jxamd_solver));
    if (!s) return JXAMD_ERROR_MEMORY;
    s->prm = *p;
    s->nb = A->nb;
    s->b = A->b;
    s->N = (size_t)A->nb * A->b;
    s->L = H->num_levels;
    if (p->coarse_rows > 0)
        for (int l = 0; l < H->num_levels; l++)
            if (H->A[l].n <= p->coarse_rows) {
                s->L = l + 1; /* this level becomes the coarsest: explicit inverse */
                break;
            }
    s->vt = p->precond_precision == 2 ? JXAMD_T_F16 : (p->precond_precision == 1 ? JXAMD_T_F32 : JXAMD_T_F64);
    s->xt = p->precond_precision == 0 ? JXAMD_T_F64 : JXAMD_T_F32;
    s->xsz = type_size(s->xt);
    int e = jxamd_stream_create(&s->stream);
    if (!e) e = jxamd_event_create(&s->ev0);
    if (!e) e = jxamd_event_create(&s->ev1);
    if (!e) e = jxamd_event_create(&s->ev2);
    if (!e) e = jxamd_event_create(&s->ev3);
    if (!e) e = jxamd_event_create_sync(&s->evh[0]);
    if (!e) e = jxamd_event_create_sync(&s->evh[1]);
    if (!e) e = jxamd_host_alloc((void **)&s->hpin, sizeof(double) * (2 * ((size_t)p->k_dim + 2) + HSM_WORDS));
    if (!e) e = jxamd_host_dev_ptr((void **)&s->hpin_dev, s->hpin);
    if (!e) {
        memset(s->hpin, 0, sizeof(double) * (2 * ((size_t)p->k_dim + 2) + HSM_WORDS));
        s->hsm = s->hpin + 2 * ((size_t)p->k_dim + 2);
        s->hsm_dev = s->hpin_dev + 2 * ((size_t)p->k_dim + 2);
    }
    if (!e && getenv("JXAMD_DEBUG"))
        fprintf(stderr, "[jxamd] solver %p: pinned column buffer host %p device %p\n", (void *)s, (void *)s->hpin,
                (void *)s->hpin_dev);
    if (!e) e = build_bsr(s, A);
    if (!e && p->stage2_solver_type == 5) e = build_bilu(s, A);
    for (int l = 0; l < s->L && !e; l++) {
        s->nlev[l] = H->A[l].n;
        e = build_level(s, l, H);
    }
    for (int l = 0; l < s->L && !e; l++) {
        /* four entries beyond the level size, zero (the coarsest level's
         * GEMV reads its right-hand side in 16-byte vectors up to a multiple
         * of four entries) */
        size_t n = (size_t)s->nlev[l] + 4;
        e = jxamd_dev_malloc(&s->F[l], s->xsz * n);
        if (!e) e = jxamd_dev_malloc(&s->U[l], s->xsz * n);
        if (!e) e = jxamd_dev_malloc(&s->T[l], s->xsz * n);
        if (!e) e = jxamd_dev_malloc(&s->Vt[l], s->xsz * n);
        if (!e) e = s->xt == JXAMD_T_F64 ? jxamd_k_fill(n, (double *)s->F[l], 0.0, s->stream)
                                         : jxamd_k_fill_f32((int)n, (float *)s->F[l], 0.0f, s->stream);
    }
    if (!e && s->xt != JXAMD_T_F64) e = jxamd_dev_malloc((void **)&s->up64, sizeof(double) * ((size_t)s->nb + 1));
    const size_t N = s->N;
    if (!e) e = jxamd_dev_malloc((void **)&s->xtmp, sizeof(double) * N);
    if (!e) e = jxamd_dev_malloc((void **)&s->V, sizeof(double) * N * (size_t)(p->k_dim + 1));
    if (!e && p->solution_update == 0) e = jxamd_dev_malloc((void **)&s->Z, sizeof(double) * N * (size_t)p->k_dim);
    if (!e) e = jxamd_dev_malloc((void **)&s->Rv, sizeof(double) * N);
    if (!e) e = jxamd_dev_malloc((void **)&s->Wv, sizeof(double) * N);
    if (!e) e = jxamd_dev_malloc((void **)&s->Bd, sizeof(double) * N);
    if (!e) e = jxamd_dev_malloc((void **)&s->Xd, sizeof(double) * N);
    if (!e) e = jxamd_dev_malloc((void **)&s->dsc, sizeof(double) * (2 * (size_t)p->k_dim + 16));
    if (!e) e = jxamd_dev_malloc((void **)&s->part, sizeof(double) * 4096); /* >= jxamd_dot_grid() blocks */
    if (!e && p->arnoldi_kernel != 1) e = jxamd_mgs_coop_plan(N, p->arnoldi_kernel == 0, &s->mgs_grid, &s->mgs_tm);
    if (!e && s->mgs_grid > 0) {
        e = jxamd_dev_malloc((void **)&s->mpart, sizeof(double) * (size_t)(p->k_dim + 1) * s->mgs_grid);
        if (!e) e = zero_ints(&s->mflag, 1, s->stream);
    }
    if (!e) e = jxamd_dev_malloc((void **)&s->count, sizeof(unsigned int) * 4);
    if (!e) e = jxamd_dev_malloc((void **)&s->ticket, sizeof(unsigned long long));
    if (!e) e = jxamd_dev_malloc((void **)&s->err, sizeof(int) * 4);
    if (!e) {
        unsigned int zc[4] = {0, 0, 0, 0};
        unsigned long long zt = 0;
        int ze[4] = {0, 0, 0, 0};
        e = jxamd_h2d(s->count, zc, sizeof(zc), s->stream);
        if (!e) e = jxamd_h2d(s->ticket, &zt, sizeof(zt), s->stream);
        if (!e) e = jxamd_h2d(s->err, ze, sizeof(ze), s->stream);
        if (!e) e = jxamd_stream_sync(s->stream);
    }
    s->tbase = 0;
    s->epoch = 0;
    s->spin_limit = JXAMD_SPIN_LIMIT;
    if (e) {
        jxamd_solver_destroy(s);
        return e;
    }
    *out = s;
    return 0;
}

int jxamd_solver_destroy(jxamd_solver *s)
{
    if (!s) return 0;
    if (s->stream) jxamd_stream_sync(s->stream);
    if (getenv("JXAMD_DEBUG")) fprintf(stderr, "[jxamd] solver %p destroyed\n", (void *)s);
    dbsr *d = &s->A;
    jxamd_dev_free(d->sptr); jxamd_dev_free(d->col); jxamd_dev_free(d->val);
    jxamd_dev_free(d->gperm); jxamd_dev_free(d->gsptr); jxamd_dev_free(d->gcol);
    jxamd_dev_free(d->gval); jxamd_dev_free(d->dinv); jxamd_dev_free(d->flag);
    jxamd_dev_free(d->gdep_ptr); jxamd_dev_free(d->gdep_list);
    jxamd_dev_free(d->gall_ptr); jxamd_dev_free(d->gall_list); jxamd_dev_free(d->gvorder); jxamd_dev_free(d->xs);
    jxamd_dev_free(d->val32); jxamd_dev_free(d->dinv32); jxamd_dev_free(d->rsc16); jxamd_dev_free(d->pptr16);
    for (int l = 0; l < JXAMD_MAX_LEVELS; l++) {
        dsell_free(&s->Ares[l]); dsell_free(&s->Gf[l]); dsell_free(&s->Gb[l]); dsell_free(&s->Gc[l]);
        jxamd_dev_free(s->wstep[l]);
        jxamd_dev_free(s->dinvl[l]);
        dsell_free(&s->R[l]); dsell_free(&s->P[l]);
        jxamd_dev_free(s->F[l]); jxamd_dev_free(s->U[l]); jxamd_dev_free(s->T[l]); jxamd_dev_free(s->Vt[l]);
    }
    jxamd_dev_free(s->cinv); jxamd_dev_free(s->up64);
    dtri_free(&s->Lf); dtri_free(&s->Ub); jxamd_dev_free(s->bz);
    jxamd_dev_free(s->xtmp); jxamd_dev_free(s->V); jxamd_dev_free(s->Z); jxamd_dev_free(s->Rv); jxamd_dev_free(s->Wv);
    jxamd_dev_free(s->Bd); jxamd_dev_free(s->Xd);
    jxamd_dev_free(s->dsc); jxamd_dev_free(s->part); jxamd_dev_free(s->count);
    jxamd_dev_free(s->mpart); jxamd_dev_free(s->mflag);
    jxamd_dev_free(s->ticket); jxamd_dev_free(s->err);
    for (int q = 0; q < PROF_POOL; q++) {
        jxamd_event_destroy(s->prof.ev0[q]);
        jxamd_event_destroy(s->prof.ev1[q]);
    }
    jxamd_event_destroy(s->ev0); jxamd_event_destroy(s->ev1);
    jxamd_event_destroy(s->ev2); jxamd_event_destroy(s->ev3);
    jxamd_event_destroy(s->evh[0]); jxamd_event_destroy(s->evh[1]);
    jxamd_host_free(s->hpin);
    jxamd_stream_destroy(s->stream);
    free(s);
    return 0;
}

int jxamd_solver_error_flag(jxamd_solver *s, int *flag)
{
    if (!s || !flag) return JXAMD_ERROR_ARG;
    CK(jxamd_d2h(flag, s->err, sizeof(int), s->stream));
    return jxamd_stream_sync(s->stream);
}

int jxamd_solver_set_spin_limit(jxamd_solver *s, int limit)
{
    if (!s || limit < 0) return JXAMD_ERROR_ARG;
    s->spin_limit = limit;
    return 0;
}

/* After a bounded wait gave up, the reductions' arrival counters and the MGS
 * flag may hold stale counts; with the stream idle, clear them, the flag
 * base and the error flag so the next launch starts from a clean state. */
static int reset_sync_state(jxamd_solver *s)
{
    unsigned int zc[4] = {0, 0, 0, 0};
    int zi[4] = {0, 0, 0, 0};
    CK(jxamd_stream_sync(s->stream));
    CK(jxamd_h2d(s->count, zc, sizeof(zc), s->stream));
    CK(jxamd_h2d(s->err, zi, sizeof(zi), s->stream));
    if (s->mflag) CK(jxamd_h2d(s->mflag, zi, sizeof(int), s->stream));
    s->mgs_base = 0;
    return jxamd_stream_sync(s->stream);
}

/* ---------------------------------------------------------------------- */
/* preconditioner pieces                                                   */
/* ---------------------------------------------------------------------- */

/* one relax-6 call: forward sweep u -> T, backward sweep T -> u.
 * uo_zero: the incoming u is known to be zero (first visit of a level). */
static int hsgs(jxamd_solver *s, int l, const void *f, int fs, void *u, int uo_zero)
{
    for (int pass = 0; pass < 2; pass++) {
        dsell *g = pass == 0 ? &s->Gf[l] : &s->Gb[l];
        const void *uo = pass == 0 ? (uo_zero ? NULL : u) : s->T[l];
        void *un = pass == 0 ? s->T[l] : u;
        s->epoch++;
        int pq = prof_begin(s);
        if (g->G > 1)
            CK(jxamd_k_gs_sweep_grp(g->G, g->vt, g->xt, g->nsl, g->perm, g->sptr, g->col, g->val, g->rsc, g->diag, f,
                                    fs, uo, un, g->dep_ptr, g->dep_list, g->flag, s->epoch, pass == 0 ? 1 : -1,
                                    s->ticket, &s->tbase, s->err, s->stream));
        else
            CK(jxamd_k_gs_sweep(g->vt, g->xt, g->nsl, g->perm, g->sptr, g->col, g->val, g->rsc, g->diag, f, fs, uo,
                                un, g->dep_ptr, g->dep_list, g->flag, s->epoch, pass == 0 ? 1 : -1, s->ticket,
                                &s->tbase, s->err, s->stream));
        prof_end_level(s, pq, JXAMD_PROF_GS_SWEEP, l, g->abytes);
    }
    return 0;
}

/* one threaded relax-6 call on level l: entry values uo (NULL: zeros) -> un */
static int hsgs_chunk(jxamd_solver *s, int l, const void *f, int fs, const void *uo, void *un)
{
    dsell *g = &s->Gc[l];
    int pq = prof_begin(s);
    if (s->wstep[l])
        CK(jxamd_k_gs_chunk(g->vt, g->xt, s->nlev[l], s->nchunks[l], s->wstep[l], g->sptr, g->col, g->val, g->rsc,
                            g->diag, f, fs, uo, un, s->stream));
    else if (g->G > 1)
        CK(jxamd_k_gs_chunk_grp(g->G, g->vt, g->xt, s->nlev[l], s->nchunks[l], g->sptr, g->col, g->val, g->rsc,
                                g->diag, f, fs, uo, un, s->stream));
    else
        CK(jxamd_k_gs_chunk_lds(lvt(g), g->xt, s->nlev[l], s->nchunks[l], g->sptr, g->col, g->val, g->rsc, g->diag, f,
                                fs, uo, un, s->stream));
    prof_end_level(s, pq, JXAMD_PROF_GS_SWEEP, l, g->abytes);
    return 0;
}

/* the level smooths with its polynomial (poly_degree > 0 and no fallback to relax 6) */
static int level_poly(const jxamd_solver *s, int l) { return s->prm.poly_degree > 0 && !s->plain[l]; }

/* one application of the level's polynomial smoother; steps alternate
 * between T and Vt and the last one lands in u.  uo_zero: 0 u holds the
 * entry values, 1 they are zero, 2 they are zero and the first step is
 * already in T[l] */
static int poly_apply(jxamd_solver *s, int l, const void *f, int fs, void *u, int uo_zero)
{
    const dsell *A = &s->Ares[l];
    const int m = s->prm.poly_degree;
    const void *src = uo_zero ? NULL : u;
    int k0 = 0;
    if (uo_zero == 2) {
        /* the restriction already wrote the first step into T[l] */
        src = s->T[l];
        k0 = 1;
    }
    for (int k = k0; k < m; k++) {
        void *dst = (k == m - 1 && m > 1) ? u : ((k % 2 == 0) ? s->T[l] : s->Vt[l]);
        int pq = prof_begin(s);
        if (!src)
            CK(jxamd_k_poly_first(A->xt, A->n, s->dinvl[l], f, fs, dst, s->poly_w[l][k], s->stream));
        else if (A->G > 1)
            CK(jxamd_k_poly_step_grp(A->G, lvt(A), A->xt, A->n, A->sptr, A->col, A->val, A->rsc, s->dinvl[l], f, fs,
                                     src, dst, s->poly_w[l][k], s->stream));
        else
            CK(jxamd_k_poly_step(lvt(A), A->xt, A->n, A->sptr, A->col, A->val, A->rsc, s->dinvl[l], f, fs, src, dst,
                                 s->poly_w[l][k], s->stream));
        /* a step from zeros reads only f and dinv */
        prof_end_level(s, pq, JXAMD_PROF_GS_SWEEP, l, src ? A->abytes : 3.0 * (double)s->xsz * A->n);
        src = dst;
    }
    if (m == 1) return jxamd_d2d(u, s->T[l], s->xsz * s->nlev[l], s->stream);
    return 0;
}

static int csr_launch(jxamd_solver *s, const dsell *A, const void *x, const void *rhs, int rs, void *y, int mode);

/* F_{l+1} = R_l Vt_l; with the polynomial smoother the same launch writes the
 * first step of level l+1's pre-smoothing (from zero) into T[l+1] */
static int restrict_launch(jxamd_solver *s, int l)
{
    const dsell *R = &s->R[l];
    const int c = l + 1;
    if (!level_poly(s, c) || c >= s->L - 1)
        return csr_launch(s, R, s->Vt[l], NULL, 1, s->F[c], 0);
    int pq = prof_begin(s);
    CK(jxamd_k_restrict_first(R->G > 1 ? R->G : 1, lvt(R), R->xt, R->n, R->sptr, R->col, R->val, R->rsc, s->Vt[l],
                              s->F[c], s->dinvl[c], s->T[c], s->poly_w[c][0], s->stream));
    prof_end(s, pq, JXAMD_PROF_CSR_SPMV, R->abytes + 2.0 * (double)s->xsz * R->n);
    return 0;
}

/* the two relax-6 calls of one level visit (relax_order 1, par_relax_if.c:36-60),
 * or one polynomial smoother application; the result stays in u */
static int relax_pair(jxamd_solver *s, int l, const void *f, int fs, void *u, int uo_zero)
{
    if (level_poly(s, l)) return poly_apply(s, l, f, fs, u, uo_zero);
    if (s->nchunks[l] > 1) {
        CK(hsgs_chunk(s, l, f, fs, uo_zero ? NULL : u, s->T[l]));
        return hsgs_chunk(s, l, f, fs, s->T[l], u);
    }
    CK(hsgs(s, l, f, fs, u, uo_zero));
    return hsgs(s, l, f, fs, u, 0);
}

/* a single relax-6 call (or polynomial smoother application); the result stays in u */
static int relax_once(jxamd_solver *s, int l, const void *f, int fs, void *u, int uo_zero)
{
    if (level_poly(s, l)) return poly_apply(s, l, f, fs, u, uo_zero);
    if (s->nchunks[l] > 1) {
        CK(hsgs_chunk(s, l, f, fs, uo_zero ? NULL : u, s->T[l]));
        return jxamd_d2d(u, s->T[l], s->xsz * s->nlev[l], s->stream);
    }
    return hsgs(s, l, f, fs, u, uo_zero);
}

static int csr_launch(jxamd_solver *s, const dsell *A, const void *x, const void *rhs, int rs, void *y, int mode)
{
    int pq = prof_begin(s);
    if (A->G > 1)
        CK(jxamd_k_csr_spmv_grp(A->G, lvt(A), A->xt, A->n, A->sptr, A->col, A->val, A->rsc, x, 1, rhs, rs, y, mode,
                                s->stream));
    else
        CK(jxamd_k_csr_spmv(lvt(A), A->xt, A->n, A->sptr, A->col, A->val, A->rsc, x, 1, rhs, rs, y, mode,
                            s->stream));
    prof_end(s, pq, JXAMD_PROF_CSR_SPMV, A->abytes);
    return 0;
}

/* V-cycle on f0 (level-0 right-hand side of the V-cycle's vector type,
 * possibly a strided view) into U[0]. */
static int vcycle(jxamd_solver *s, const void *f0, int fs0, int u0_zero)
{
    const int L = s->L;
    if (L == 1) {
        /* no coarsening: one relax-6 call (par_cycle.c:236-255) */
        return relax_once(s, 0, f0, fs0, s->U[0], u0_zero);
    }
    for (int l = 0; l < L - 1; l++) {
        const void *f = l == 0 ? f0 : s->F[l];
        const int fs = l == 0 ? fs0 : 1;
        /* coarse levels start from zero; with the polynomial smoother their
         * first step came with the restriction (2) */
        const int zero = l == 0 ? u0_zero : (level_poly(s, l) ? 2 : 1);
        CK(relax_pair(s, l, f, fs, s->U[l], zero));
        /* Vtemp = F - A U; F_{l+1} = P^T Vtemp */
        CK(csr_launch(s, &s->Ares[l], s->U[l], f, fs, s->Vt[l], 1));
        CK(restrict_launch(s, l));
    }
    int pc = prof_begin(s);
    CK(jxamd_k_dense_gemv(s->xt, s->nc, s->ncld, s->cinv, s->F[L - 1], s->U[L - 1], s->stream));
    prof_end(s, pc, JXAMD_PROF_COARSE, (double)s->xsz * s->nc * ((double)s->nc + 2.0));
    for (int l = L - 2; l >= 0; l--) {
        CK(csr_launch(s, &s->P[l], s->U[l + 1], NULL, 1, s->U[l], 2));
        const void *f = l == 0 ? f0 : s->F[l];
        const int fs = l == 0 ? fs0 : 1;
        CK(relax_pair(s, l, f, fs, s->U[l], 0));
    }
    return 0;
}

/* K stage-2 sweeps from x0 (or from the prolongated xp if x0 is NULL) into
 * xn; K is 1 or the fused count d->K */
static int bhgs_run(jxamd_solver *s, int K, const double *rhs, const double *x0, const double *xp, double *xn,
                    double w, double omega)
{
    dbsr *d = &s->A;
    const long long xstride = (long long)d->gnpos * d->b;
    double *xsl0 = d->xs + (size_t)d->K * xstride;
    s->epoch++;
    int pq = prof_begin(s);
    int e = jxamd_k_bhgs_gather(d->gnpos, d->b, d->gperm, x0, xp, s->prm.pressure_index, xsl0, s->stream);
    if (!e)
        e = jxamd_k_bhgs_multi(d->b, d->gnsl, K, K > 1 ? d->gvorder : NULL, d->gperm, d->gsptr, d->gcol, d->gval,
                               d->dinv, rhs, xsl0, d->xs, xstride, xn, d->gdep_ptr, d->gdep_list, d->gall_ptr,
                               d->gall_list, d->flag, s->epoch, w, omega, s->ticket, &s->tbase, s->err, s->stream);
    prof_end(s, pq, JXAMD_PROF_BHGS_SWEEP, K * d->sweep_bytes);
    return e;
}

static int bsr_apply(jxamd_solver *s, const double *x, const double *rhs, double *y, int mode);

static int tri_run(jxamd_solver *s, const dtri *t, const double *rhs, double *out)
{
    s->epoch++;
    int pq = prof_begin(s);
    CK(jxamd_k_btri(s->b, t->vt, t->nsl, t->perm, t->sptr, t->col, t->val, t->rsc, t->dinv, rhs, out, t->dep_ptr,
                    t->dep_list, t->flag, s->epoch, s->ticket, &s->tbase, s->err, s->stream));
    prof_end(s, pq, JXAMD_PROF_BILU, t->bytes);
    return 0;
}

/* z = (LU)^{-1} r; r and z may not alias s->bz */
static int bilu_solve(jxamd_solver *s, const double *r, double *z)
{
    CK(tri_run(s, &s->Lf, r, s->bz));
    return tri_run(s, &s->Ub, s->bz, z);
}

/* x = M^{-1} r (jxf_cpr.c:424-610) */
static int cpr_apply(jxamd_solver *s, const double *r, double *x)
{
    const int b = s->b, pi = s->prm.pressure_index;
    /* stage 1: the restricted residual is read in place (stride b), or cast
     * into the V-cycle's precision */
    const void *f0 = r + pi;
    int fs0 = b;
    int zero = 1;
    if (s->xt != JXAMD_T_F64 && s->prm.stage1_maxit > 0) {
        if (level_poly(s, 0) && s->L > 1) {
            /* the cast and level 0's first polynomial step from zero in one launch */
            int pq = prof_begin(s);
            CK(jxamd_k_cast_first(s->nlev[0], r, b, pi, (float *)s->F[0], (const float *)s->dinvl[0],
                                  (float *)s->T[0], s->poly_w[0][0], s->stream));
            prof_end_level(s, pq, JXAMD_PROF_GS_SWEEP, 0, (8.0 + 4.0 * 3.0) * s->nlev[0]);
            zero = 2;
        } else {
            CK(jxamd_k_gather_cast(s->nlev[0], r, b, pi, (float *)s->F[0], s->stream));
        }
        f0 = s->F[0];
        fs0 = 1;
    }
    for (int it = 0; it < s->prm.stage1_maxit; it++) {
        CK(vcycle(s, f0, fs0, zero));
        zero = 0;
    }
    /* prolongation (type 0) + stage-2 sweeps; the first sweep reads the
     * prolongated vector on the fly and the last one lands in x */
    const int m = s->prm.stage2_maxit;
    /* the partitioned sweeps read an FP32 stage-1 result directly */
    const int xp_f32 = s->xt != JXAMD_T_F64 && s->prm.stage1_maxit > 0 && m > 0 && s->prm.stage2_chunk > 0 &&
                       s->prm.stage2_solver_type == 2;
    double *xp = (double *)s->U[0];
    if (s->xt != JXAMD_T_F64) {
        xp = s->up64;
        if (s->prm.stage1_maxit > 0 && !xp_f32) CK(jxamd_k_widen(s->nlev[0], (const float *)s->U[0], xp, s->stream));
    }
    if (s->prm.stage1_maxit == 0) CK(jxamd_k_fill(s->nlev[0], xp, 0.0, s->stream));
    if (m == 0 || s->prm.stage2_solver_type == 5) CK(jxamd_k_prolong(s->nb, b, pi, xp, x, s->stream));
    if (m == 0) return 0;
    if (s->prm.stage2_chunk > 0 && s->prm.stage2_solver_type == 2) {
        /* sweeps as on ceil(nb / c) ranks: row-parallel launches, ping-pong
         * between xtmp and x so that the last one lands in x */
        dbsr *d = &s->A;
        const int parts = (s->nb + s->prm.stage2_chunk - 1) / s->prm.stage2_chunk;
        double *dst = (m % 2 == 1) ? x : s->xtmp;
        const double *src = NULL;
        for (int it = 0; it < m; it++) {
            int pq = prof_begin(s);
            const int nar = s->prm.stage2_precision; /* 0 FP64, 1 FP32, 2 FP16 blocks */
            CK(jxamd_k_bhgs_part(b, stage2_vt(nar), s->nb, parts, d->sptr, d->pptr16, d->col,
                                 nar ? d->val32 : (const void *)d->val, d->rsc16,
                                 d->unit_diag ? NULL : (nar ? d->dinv32 : (const void *)d->dinv), r, src,
                                 xp_f32 ? s->U[0] : (const void *)xp, xp_f32, pi, dst, s->prm.stage2_weight,
                                 s->prm.stage2_omega, s->stream));
            prof_end(s, pq, JXAMD_PROF_BHGS_SWEEP, nar ? d->part_bytes32 : d->part_bytes);
            src = dst;
            dst = (dst == x) ? s->xtmp : x;
        }
        return 0;
    }
    if (s->prm.stage2_solver_type == 5) {
        /* BILU(0) corrections: x += (LU)^{-1} (r - A x) */
        for (int it = 0; it < m; it++) {
            CK(bsr_apply(s, x, r, s->xtmp, 1));
            CK(bilu_solve(s, s->xtmp, s->xtmp));
            CK(jxamd_k_axpy(s->N, NULL, 1.0, 0, s->xtmp, x, s->stream));
        }
        return 0;
    }
    return bhgs_run(s, m, r, NULL, xp, x, s->prm.stage2_weight, s->prm.stage2_omega);
}

int jxamd_cpr_precond_device(jxamd_solver *s, const double *d_r, double *d_x)
{
    if (!s || !d_r || !d_x) return JXAMD_ERROR_ARG;
    CK(cpr_apply(s, d_r, d_x));
    return jxamd_stream_sync(s->stream);
}

int jxamd_cpr_precond_callback(void *precond_data, void *A, void *b, void *x)
{
    (void)A;
    jxamd_precond_ctx *c = (jxamd_precond_ctx *)precond_data;
    if (!c || !c->solver || !c->vector_data || !b || !x) return JXAMD_ERROR_ARG;
    jxamd_solver *s = c->solver;
    const double *hb = c->vector_data(b);
    double *hx = c->vector_data(x);
    if (!hb || !hx) return JXAMD_ERROR_ARG;
    CK(jxamd_h2d(s->Bd, hb, sizeof(double) * s->N, s->stream));
    CK(cpr_apply(s, s->Bd, s->Xd));
    CK(jxamd_d2h(hx, s->Xd, sizeof(double) * s->N, s->stream));
    return jxamd_stream_sync(s->stream);
}

int jxamd_op_bsr_matvec(jxamd_solver *s, double alpha, const double *d_x, double beta, double *d_y)
{
    if (!s) return JXAMD_ERROR_ARG;
    dbsr *d = &s->A;
    CK(jxamd_k_bsr_spmv(d->b, d->nb, d->unit_diag, d->sptr, d->col, d->val, d_x, NULL, d_y, 2, alpha, beta, s->stream));
    return jxamd_stream_sync(s->stream);
}

int jxamd_op_bsr_matvec_timed(jxamd_solver *s, const double *d_x, double *d_y, int reps, double *ms_per_launch)
{
    if (!s || reps < 1 || !ms_per_launch) return JXAMD_ERROR_ARG;
    dbsr *d = &s->A;
    CK(jxamd_event_record(s->ev0, s->stream));
    for (int r = 0; r < reps; r++)
        CK(jxamd_k_bsr_spmv(d->b, d->nb, d->unit_diag, d->sptr, d->col, d->val, d_x, NULL, d_y, 0, 1.0, 0.0, s->stream));
    CK(jxamd_event_record(s->ev1, s->stream));
    CK(jxamd_stream_sync(s->stream));
    float ms = 0.f;
    CK(jxamd_event_elapsed(&ms, s->ev0, s->ev1));
    *ms_per_launch = ms / reps;
    return 0;
}

int jxamd_cpr_precond_timed(jxamd_solver *s, const double *d_r, double *d_x, int reps, double *ms_per_apply)
{
    if (!s || reps < 1 || !ms_per_apply) return JXAMD_ERROR_ARG;
    CK(jxamd_event_record(s->ev0, s->stream));
    for (int r = 0; r < reps; r++) CK(cpr_apply(s, d_r, d_x));
    CK(jxamd_event_record(s->ev1, s->stream));
    CK(jxamd_stream_sync(s->stream));
    float ms = 0.f;
    CK(jxamd_event_elapsed(&ms, s->ev0, s->ev1));
    *ms_per_apply = ms / reps;
    return 0;
}

/* FP64 device vectors in and out of the V-cycle's precision */
static int to_level(jxamd_solver *s, int n, const double *src, void *dst)
{
    if (s->xt == JXAMD_T_F64) return jxamd_k_copy(n, src, (double *)dst, s->stream);
    return jxamd_k_gather_cast(n, src, 1, 0, (float *)dst, s->stream);
}

static int from_level(jxamd_solver *s, int n, const void *src, double *dst)
{
    if (s->xt == JXAMD_T_F64) return jxamd_k_copy(n, (const double *)src, dst, s->stream);
    return jxamd_k_widen(n, (const float *)src, dst, s->stream);
}

int jxamd_op_hsgs(jxamd_solver *s, int level, const double *d_f, double *d_u)
{
    if (!s || level < 0 || (s->L > 1 && level >= s->L - 1) || level >= s->L) return JXAMD_ERROR_ARG;
    const int n = s->nlev[level];
    CK(to_level(s, n, d_f, s->F[level]));
    CK(to_level(s, n, d_u, s->U[level]));
    CK(relax_once(s, level, s->F[level], 1, s->U[level], 0));
    CK(from_level(s, n, s->U[level], d_u));
    return jxamd_stream_sync(s->stream);
}

int jxamd_op_coarse_solve(jxamd_solver *s, const double *d_f, double *d_u)
{
    if (!s || s->L < 2) return JXAMD_ERROR_ARG;
    const int L = s->L;
    CK(to_level(s, s->nc, d_f, s->F[L - 1]));
    CK(jxamd_k_dense_gemv(s->xt, s->nc, s->ncld, s->cinv, s->F[L - 1], s->U[L - 1], s->stream));
    CK(from_level(s, s->nc, s->U[L - 1], d_u));
    return jxamd_stream_sync(s->stream);
}

int jxamd_op_vcycle(jxamd_solver *s, const double *d_f, double *d_u)
{
    if (!s) return JXAMD_ERROR_ARG;
    const int n = s->nlev[0];
    CK(to_level(s, n, d_f, s->F[0]));
    CK(to_level(s, n, d_u, s->U[0]));
    CK(vcycle(s, s->F[0], 1, 0));
    CK(from_level(s, n, s->U[0], d_u));
    return jxamd_stream_sync(s->stream);
}

int jxamd_solver_num_levels(jxamd_solver *s, int *num_levels)
{
    if (!s || !num_levels) return JXAMD_ERROR_ARG;
    *num_levels = s->L;
    return 0;
}

static int poly_level_ok(const jxamd_solver *s, int level, int degree)
{
    return s && s->prm.poly_degree > 0 && degree == s->prm.poly_degree && level >= 0 && level < s->L &&
           (level < s->L - 1 || s->L == 1);
}

int jxamd_solver_set_poly(jxamd_solver *s, int level, int degree, const double *omega)
{
    if (!poly_level_ok(s, level, degree) || !omega) return JXAMD_ERROR_ARG;
    for (int k = 0; k < degree; k++)
        if (!(omega[k] == omega[k])) return JXAMD_ERROR_ARG; /* NaN */
    for (int k = 0; k < degree; k++) s->poly_w[level][k] = omega[k];
    return 0;
}

int jxamd_solver_get_poly(jxamd_solver *s, int level, int degree, double *omega, double *lambda_max)
{
    if (!poly_level_ok(s, level, degree)) return JXAMD_ERROR_ARG;
    if (omega)
        for (int k = 0; k < degree; k++) omega[k] = s->poly_w[level][k];
    if (lambda_max) *lambda_max = s->lam[level];
    return 0;
}

int jxamd_solver_poly_features(jxamd_solver *s, int level, float *feat)
{
    if (!poly_level_ok(s, level, s ? s->prm.poly_degree : 0) || !feat) return JXAMD_ERROR_ARG;
    memcpy(feat, s->feat[level], sizeof(float) * JXAMD_MLP_NFEAT);
    return 0;
}

/* ||U[0]|| from the device (host copy; setup-time only) */
static int level0_norm(jxamd_solver *s, double *nrm)
{
    const int n = s->nlev[0];
    void *h = malloc(s->xsz * (n > 0 ? n : 1));
    if (!h) return JXAMD_ERROR_MEMORY;
    int e = jxamd_d2h(h, s->U[0], s->xsz * n, s->stream);
    if (!e) e = jxamd_stream_sync(s->stream);
    double acc = 0.0;
    for (int i = 0; i < n && !e; i++) {
        double v = s->xt == JXAMD_T_F64 ? ((double *)h)[i] : (double)((float *)h)[i];
        acc += v * v;
    }
    free(h);
    *nrm = sqrt(acc);
    return e;
}

int jxamd_solver_vcycle_factor(jxamd_solver *s, int ncycles, double *factor)
{
    if (!s || ncycles < 1 || !factor) return JXAMD_ERROR_ARG;
    const int n = s->nlev[0];
    void *h = malloc(s->xsz * (n > 0 ? n : 1));
    if (!h) return JXAMD_ERROR_MEMORY;
    uint64_t st = 0x2545f4914f6cdd1dull;
    for (int i = 0; i < n; i++) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        const double v = (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        if (s->xt == JXAMD_T_F64) ((double *)h)[i] = v;
        else ((float *)h)[i] = (float)v;
    }
    int e = jxamd_h2d(s->U[0], h, s->xsz * n, s->stream);
    if (!e) e = s->xt == JXAMD_T_F64 ? jxamd_k_fill(n, (double *)s->F[0], 0.0, s->stream)
                                     : jxamd_k_fill_f32(n, (float *)s->F[0], 0.0f, s->stream);
    if (!e) e = jxamd_stream_sync(s->stream);
    free(h);
    CK(e);
    double n0 = 0.0, n1 = 0.0;
    CK(level0_norm(s, &n0));
    for (int c = 0; c < ncycles; c++) CK(vcycle(s, s->F[0], 1, 0));
    CK(level0_norm(s, &n1));
    *factor = n0 > 0.0 ? pow(n1 / n0, 1.0 / ncycles) : 0.0;
    return 0;
}

/* Judge of the polynomial smoother against the plain smoother, which is also
 * the fallback.  Each level has up to two candidate step sets: w[l] (the
 * network's roots, where ok[l]; none when w is NULL) and the default
 * Chebyshev steps.  From relax 6 on every smoothing level, each level in
 * turn takes the candidate with the lower measured V-cycle contraction
 * factor unless that is worse than keeping relax 6 there (a second pass
 * retries the levels left on relax 6 in the final context).  The result
 * never contracts worse than relax 6 everywhere; a level whose candidates
 * are poor stays on relax 6 without disturbing the judgement of the others.
 * *used = levels on the network's roots, *used_default = on the defaults. */
static int judge_levels(jxamd_solver *s, double (*w)[8], const int *ok, int *used, int *used_default, double *factor,
                        double *fplain)
{
    const int m = s->prm.poly_degree;
    const int Ls = s->L > 1 ? s->L - 1 : 1;
    for (int l = 0; l < Ls; l++) {
        s->plain[l] = 1;
        s->learned[l] = 0;
    }
    double best = 0.0;
    CK(jxamd_solver_vcycle_factor(s, 4, &best));
    if (fplain) *fplain = best;
    for (int pass = 0; pass < 2; pass++) {
        for (int l = 0; l < Ls; l++) {
            if (!s->plain[l]) continue;
            double fbest = 0.0, wbest[8];
            int have = 0, lrn = 0;
            for (int cand = 0; cand < 2; cand++) {
                const double *wc = cand == 0 ? (w && ok[l] ? w[l] : NULL) : s->poly_def[l];
                if (!wc) continue;
                for (int k = 0; k < m; k++) s->poly_w[l][k] = wc[k];
                s->plain[l] = 0;
                double f = 0.0;
                CK(jxamd_solver_vcycle_factor(s, 4, &f));
                if (!have || f < fbest) {
                    fbest = f;
                    lrn = cand == 0;
                    for (int k = 0; k < m; k++) wbest[k] = wc[k];
                    have = 1;
                }
            }
            if (have && fbest <= best) {
                best = fbest;
                s->plain[l] = 0;
                s->learned[l] = lrn;
                for (int k = 0; k < m; k++) s->poly_w[l][k] = wbest[k];
            } else {
                s->plain[l] = 1; /* worse per cycle: stay on relax 6 */
                for (int k = 0; k < m; k++) s->poly_w[l][k] = s->poly_def[l][k];
            }
        }
    }
    int nl = 0, nd = 0;
    for (int l = 0; l < Ls; l++)
        if (!s->plain[l]) {
            if (s->learned[l]) nl++;
            else nd++;
        }
    if (used) *used = nl;
    if (used_default) *used_default = nd;
    if (factor) *factor = best;
    return 0;
}

int jxamd_solver_judge_poly(jxamd_solver *s, int *used, double *factor, double *factor_plain)
{
    if (!s || s->prm.poly_degree < 1) return JXAMD_ERROR_ARG;
    return judge_levels(s, NULL, NULL, NULL, used, factor, factor_plain);
}

int jxamd_solver_set_level_smoother(jxamd_solver *s, int level, int plain)
{
    if (!s || s->prm.poly_degree < 1 || level < 0 || level >= s->L) return JXAMD_ERROR_ARG;
    s->plain[level] = plain ? 1 : 0;
    s->learned[level] = 0;
    return 0;
}

int jxamd_solver_get_level_smoother(jxamd_solver *s, int level, int *plain)
{
    if (!s || !plain || level < 0 || level >= s->L) return JXAMD_ERROR_ARG;
    if (s->L > 1 && level == s->L - 1) *plain = 3; /* the coarsest level: direct solve */
    else *plain = (s->prm.poly_degree < 1 || s->plain[level]) ? 1 : (s->learned[level] ? 2 : 0);
    return 0;
}

int jxamd_solver_apply_mlp(jxamd_solver *s, const float *d_params, int *used, double *factor_learned,
                           double *factor_plain)
{
    if (!s || !d_params || s->prm.poly_degree < 1 || s->prm.poly_degree > JXAMD_MLP_NOUT) return JXAMD_ERROR_ARG;
    const int m = s->prm.poly_degree; /* the network's first m outputs */
    const int Ls = s->L > 1 ? s->L - 1 : 1;
    float hy[JXAMD_MAX_LEVELS * JXAMD_MLP_NOUT];
    float *dx = NULL, *dy = NULL;
    int e = jxamd_dev_malloc((void **)&dx, sizeof(float) * Ls * JXAMD_MLP_NFEAT);
    if (!e) e = jxamd_dev_malloc((void **)&dy, sizeof(float) * Ls * JXAMD_MLP_NOUT);
    if (!e) e = jxamd_h2d(dx, s->feat, sizeof(float) * Ls * JXAMD_MLP_NFEAT, s->stream);
    if (!e) e = jxamd_mlp_forward(d_params, dx, Ls, dy, s->stream);
    if (!e) e = jxamd_d2h(hy, dy, sizeof(float) * Ls * JXAMD_MLP_NOUT, s->stream);
    if (!e) e = jxamd_stream_sync(s->stream);
    jxamd_dev_free(dx);
    jxamd_dev_free(dy);
    CK(e);
    /* the network predicts the polynomial's roots in units of lambda_max:
     * w_k = 1 / (tau_k lambda_max) */
    double w[JXAMD_MAX_LEVELS][8];
    int ok[JXAMD_MAX_LEVELS];
    /* the span of the roots of the label family for degree m: the smallest
     * root of a Chebyshev polynomial on [lo, hi] is lo (1 + c) / 2 + hi (1 - c) / 2,
     * the largest lo (1 - c) / 2 + hi (1 + c) / 2, c = cos(pi / (2m)) */
    const double cm = cos(M_PI / (2.0 * m));
    const double tmin = JXAMD_MLP_LO_MIN * (1.0 + cm) / 2.0 + JXAMD_MLP_HI_MIN * (1.0 - cm) / 2.0;
    const double tmax = JXAMD_MLP_LO_MAX * (1.0 - cm) / 2.0 + JXAMD_MLP_HI_MAX * (1.0 + cm) / 2.0;
    for (int l = 0; l < Ls; l++) {
        ok[l] = 1;
        for (int k = 0; k < m; k++) {
            double tau = (double)hy[l * JXAMD_MLP_NOUT + k];
            if (tau != tau) ok[l] = 0; /* NaN: this level stays on relax 6 */
            tau = tau < tmin ? tmin : (tau > tmax ? tmax : tau);
            w[l][k] = 1.0 / (tau * s->lam[l]);
        }
    }
    return judge_levels(s, w, ok, used, NULL, factor_learned, factor_plain);
}

int jxamd_op_bilu_solve(jxamd_solver *s, const double *d_r, double *d_z)
{
    if (!s || s->prm.stage2_solver_type != 5 || !d_r || !d_z || d_r == d_z) return JXAMD_ERROR_ARG;
    CK(bilu_solve(s, d_r, d_z));
    return jxamd_stream_sync(s->stream);
}

int jxamd_op_bhgs(jxamd_solver *s, double *d_x, const double *d_b, double w, double omega)
{
    if (!s) return JXAMD_ERROR_ARG;
    if (s->prm.stage2_chunk > 0) {
        const int parts = (s->nb + s->prm.stage2_chunk - 1) / s->prm.stage2_chunk;
        const int nar = s->prm.stage2_precision;
        CK(jxamd_k_bhgs_part(s->b, stage2_vt(nar), s->nb, parts, s->A.sptr, s->A.pptr16, s->A.col,
                             nar ? s->A.val32 : (const void *)s->A.val, s->A.rsc16,
                             s->A.unit_diag ? NULL : (nar ? s->A.dinv32 : (const void *)s->A.dinv), d_b, d_x, NULL, 0, 0,
                             s->xtmp, w, omega, s->stream));
    } else {
        CK(bhgs_run(s, 1, d_b, d_x, NULL, s->xtmp, w, omega));
    }
    CK(jxamd_k_copy(s->N, s->xtmp, d_x, s->stream));
    return jxamd_stream_sync(s->stream);
}

/* ---------------------------------------------------------------------- */
/* GMRES (pgmres.c:950-1713)                                               */
/* ---------------------------------------------------------------------- */
static int dev_dot(jxamd_solver *s, const double *u, double *y, const double *vprev, const double *hprev, double *out,
                   double *hout)
{
    int pq = prof_begin(s);
    int e = jxamd_k_dot_mgs(s->N, u, y, vprev, hprev, s->part, s->count, out, hout, s->stream);
    /* reads u and y (and v_prev, writing y) */
    prof_end(s, pq, JXAMD_PROF_MGS, 8.0 * s->N * (vprev ? (u == y ? 3.0 : 4.0) : (u == y ? 1.0 : 2.0)));
    return e;
}

/* y = A x (mode 0) or y = rhs - A x (mode 1) with the stage-0 layout */
static int bsr_apply(jxamd_solver *s, const double *x, const double *rhs, double *y, int mode)
{
    dbsr *A = &s->A;
    int pq = prof_begin(s);
    int e = jxamd_k_bsr_spmv(A->b, A->nb, A->unit_diag, A->sptr, A->col, A->val, x, rhs, y, mode, 0.0, 0.0, s->stream);
    prof_end(s, pq, JXAMD_PROF_BSR_SPMV, A->spmv_bytes + (mode == 1 ? 8.0 * s->N : 0.0));
    return e;
}

int jxamd_solver_profile(jxamd_solver *s, int enable)
{
    if (!s) return JXAMD_ERROR_ARG;
    prof_t *p = &s->prof;
    CK(prof_flush(s));
    if (enable && !p->ev0[0]) {
        for (int q = 0; q < PROF_POOL; q++) {
            CK(jxamd_event_create(&p->ev0[q]));
            CK(jxamd_event_create(&p->ev1[q]));
        }
    }
    memset(p->count, 0, sizeof(p->count));
    memset(p->ms, 0, sizeof(p->ms));
    memset(p->tbytes, 0, sizeof(p->tbytes));
    memset(p->lcount, 0, sizeof(p->lcount));
    memset(p->lms, 0, sizeof(p->lms));
    memset(p->lbytes, 0, sizeof(p->lbytes));
    p->on = enable ? 1 : 0;
    return 0;
}

int jxamd_solver_profile_read(jxamd_solver *s, int kind, long *launches, double *ms, double *bytes)
{
    if (!s || kind < 0 || kind >= JXAMD_PROF_KINDS) return JXAMD_ERROR_ARG;
    CK(prof_flush(s));
    if (launches) *launches = s->prof.count[kind];
    if (ms) *ms = s->prof.ms[kind];
    if (bytes) *bytes = s->prof.tbytes[kind];
    return 0;
}

int jxamd_solver_profile_read_level(jxamd_solver *s, int level, long *launches, double *ms, double *bytes)
{
    if (!s || level < 0 || level >= JXAMD_MAX_LEVELS) return JXAMD_ERROR_ARG;
    CK(prof_flush(s));
    if (launches) *launches = s->prof.lcount[level];
    if (ms) *ms = s->prof.lms[level];
    if (bytes) *bytes = s->prof.lbytes[level];
    return 0;
}

static double mono_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Squared norms reach the host through pinned words (hsm): the dot
 * kernel's last workgroup writes its result there as well as to the device
 * slot, so a reading takes no copy, and the host learns of it as soon as
 * it is written, not when the stream drains.  A word holds an armed
 * pattern, a signalling NaN that no arithmetic result has, until its
 * kernel has written it. */
static const uint64_t HSM_ARMED = 0x7ff4a5a5a5a5a5a5ull;

static void hsm_arm(double *w) { memcpy(w, &HSM_ARMED, sizeof(double)); }

static int hsm_written(const double *w)
{
    const double d = *(volatile const double *)w;
    uint64_t u;
    memcpy(&u, &d, sizeof(u));
    return u != HSM_ARMED;
}

/* wait until hsm[0..k-1] are written: polling, then after 0.2 s a wait
 * for the stream, after which they must be there */
static int hsm_wait(jxamd_solver *s, int k)
{
    double t0 = -1.0;
    for (int q = 0; q < k; q++) {
        for (unsigned long spin = 1; !hsm_written(s->hsm + q); spin++) {
            if ((spin & 4095) == 0) {
                if (t0 < 0.0) t0 = mono_s();
                else if (mono_s() - t0 > 0.2) {
                    CK(jxamd_stream_sync(s->stream));
                    for (int r = q; r < k; r++)
                        if (!hsm_written(s->hsm + r)) return JXAMD_ERROR_GENERIC;
                    q = k;
                    break;
                }
            }
        }
    }
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    return 0;
}

static int host_norm(jxamd_solver *s, const double *v, double *nrm)
{
    double *slot = s->dsc + 2 * s->prm.k_dim + 8;
    hsm_arm(s->hsm);
    CK(dev_dot(s, v, (double *)v, NULL, NULL, slot, s->hsm_dev));
    CK(hsm_wait(s, 1));
    *nrm = sqrt(s->hsm[0]);
    return 0;
}

/* ||u|| and ||v|| with one wait, in two halves: the kernels, then the
 * reading.  The squared norm of v also stays in the device word
 * dsc[2 k_dim + 9]. */
static int host_norm2_enqueue(jxamd_solver *s, const double *u, const double *v)
{
    double *slot = s->dsc + 2 * s->prm.k_dim + 8;
    hsm_arm(s->hsm);
    hsm_arm(s->hsm + 1);
    CK(dev_dot(s, u, (double *)u, NULL, NULL, slot, s->hsm_dev));
    return dev_dot(s, v, (double *)v, NULL, NULL, slot + 1, s->hsm_dev + 1);
}

static int host_norm2_read(jxamd_solver *s, double *nu, double *nv)
{
    CK(hsm_wait(s, 2));
    *nu = sqrt(s->hsm[0]);
    *nv = sqrt(s->hsm[1]);
    return 0;
}

/* CPR application, timed with events when the solve is timed (print_level
 * > 0 or profiling: each timed application is followed by a stream sync) */
static int precond_timed(jxamd_solver *s, const double *r, double *x, int timed)
{
    if (timed) CK(jxamd_event_record(s->ev2, s->stream));
    CK(cpr_apply(s, r, x));
    if (timed) CK(jxamd_event_record(s->ev3, s->stream));
    return 0;
}

static int accumulate_precond_time(jxamd_solver *s)
{
    float ms = 0.f;
    CK(jxamd_stream_sync(s->stream));
    CK(jxamd_event_elapsed(&ms, s->ev2, s->ev3));
    s->t_prec_ms += ms;
    return 0;
}

/* Enqueue Arnoldi step i: v_i = A M^{-1} v_{i-1}, MGS against v_0..v_{i-1}
 * and normalisation (column in h on the device: projections, then the
 * squared norm), then the column's copy to hpin[slot].  The one-launch MGS
 * kernel writes the column there itself and then a fresh sequence number in
 * the slot's last word (hwant[slot]); the other paths record the slot's
 * completion event (hwant[slot] = 0).  Nothing here waits for the device. */
static int arnoldi_enqueue(jxamd_solver *s, int i, int slot, int timed)
{
    const size_t N = s->N;
    double *V = s->V, *h = s->dsc;
#define VEC(j) (V + (size_t)(j) * N)
    /* z = M^{-1} v_{i-1}, kept as column i-1 of Z for the cycle's update */
    double *z = s->Z ? s->Z + (size_t)(i - 1) * N : s->Rv;
    CK(precond_timed(s, VEC(i - 1), z, timed));
    CK(bsr_apply(s, z, NULL, VEC(i), 0));
    if (s->mgs_grid > 0) {
        /* the whole MGS step in one launch: v_0..v_{i-1} read once,
         * w = v_i read and written once */
        int pq = prof_begin(s);
        const double seq = s->hseq + 1.0;
        const int e = jxamd_k_mgs_coop(s->mgs_grid, s->mgs_tm, s->prm.arnoldi_kernel == 0, N, VEC(0), N, i, VEC(i), h,
                                       s->mpart, s->count + 1,
                                       s->mflag, s->mgs_base, s->err,
                                       s->hpin_dev + (size_t)slot * (s->prm.k_dim + 2), s->spin_limit,
                                       s->prm.k_dim + 1, seq, s->stream);
        if (!e) {
            prof_end(s, pq, JXAMD_PROF_MGS, 8.0 * N * (i + 2));
            s->mgs_base += i + 1;
            s->hseq = seq;
            s->hwant[slot] = seq; /* the kernel writes the column, then seq, to hpin[slot] */
            return 0;
        }
        /* the launch was refused (for a cooperative launch: the grid could
         * not be made resident): nothing ran; use the launch-per-projection
         * MGS from now on */
        (void)jxamd_last_error(); /* the opened timing slot is reused by the next prof_begin */
        s->mgs_grid = 0;
    }
    if (s->mgs_grid == 0) {
        /* fused MGS: kernel j subtracts h[j-1] v_{j-1} and forms h[j] = <v_j, w> */
        CK(dev_dot(s, VEC(0), VEC(i), NULL, NULL, h + 0, NULL));
        for (int j = 1; j < i; j++) CK(dev_dot(s, VEC(j), VEC(i), VEC(j - 1), h + j - 1, h + j, NULL));
        CK(dev_dot(s, VEC(i), VEC(i), VEC(i - 1), h + i - 1, h + i, NULL));
        int pq = prof_begin(s);
        CK(jxamd_k_scale(N, h + i, 0.0, VEC(i), s->stream));
        prof_end(s, pq, JXAMD_PROF_VEC, 16.0 * N);
    }
#undef VEC
    CK(jxamd_d2h(s->hpin + (size_t)slot * (s->prm.k_dim + 2), h, sizeof(double) * (i + 1), s->stream));
    s->hwant[slot] = 0.0;
    return jxamd_event_record(s->evh[slot], s->stream);
}

/* Wait until hpin[slot] holds the column arnoldi_enqueue last put there.
 * A one-launch MGS step announces its column by the sequence number in the
 * slot's last word: the host polls that word (no completion event, so the
 * stream carries no marker between the step and the next one, and the host
 * learns of the column as the kernel writes it).  After 0.2 s of polling the
 * host waits for the stream instead and then requires the number to be
 * there: a step whose bounded waits gave up may never write it, and the
 * solve then ends with JXAMD_ERROR_GENERIC (its caller finds the device
 * error flag set and resets the synchronisation state). */
static int column_wait(jxamd_solver *s, int slot)
{
    const double want = s->hwant[slot];
    if (want == 0.0) return jxamd_event_sync(s->evh[slot]);
    volatile const double *f = s->hpin + (size_t)slot * (s->prm.k_dim + 2) + s->prm.k_dim + 1;
    if (*f != want) {
        const double t0 = mono_s();
        for (unsigned long spin = 1; *f != want; spin++) {
            if ((spin & 4095) == 0 && mono_s() - t0 > 0.2) {
                CK(jxamd_stream_sync(s->stream));
                if (*f != want) return JXAMD_ERROR_GENERIC;
                break;
            }
        }
    }
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    return 0;
}

int jxamd_solver_solve_device(jxamd_solver *s, const double *bvec, double *x, double *res_history, jxamd_stats *st)
{
    if (!s || !bvec || !x) return JXAMD_ERROR_ARG;
    const int k_dim = s->prm.k_dim, max_iter = s->prm.max_iter;
    const size_t N = s->N;
    const double r_tol = s->prm.tol, a_tol = s->prm.a_tol, epsmac = 1.e-16;
    const double convfac_threshold = s->prm.convfac_threshold;
    double *V = s->V;
#define VEC(j) (V + (size_t)(j) * N)
    double *ydev = s->dsc + k_dim + 4; /* back-solved coefficients, k_dim+1 */
    double *rs = (double *)calloc(k_dim + 1, sizeof(double));
    double *c = (double *)calloc(k_dim, sizeof(double));
    double *sn = (double *)calloc(k_dim, sizeof(double));
    double *hcol = (double *)calloc(k_dim + 2, sizeof(double));
    double **hh = (double **)calloc(k_dim + 1, sizeof(double *));
    int err = 0, converged = 0, nprec = 0, nrestart = 0, nfalse = 0, diverged = 0;
    int eflag = -1; /* the device error flag, once read */
    if (!rs || !c || !sn || !hcol || !hh) err = JXAMD_ERROR_MEMORY;
    for (int q = 0; q <= k_dim && !err; q++)
        if (!(hh[q] = (double *)calloc(k_dim, sizeof(double)))) err = JXAMD_ERROR_MEMORY;
    if (err) {
        for (int q = 0; hh && q <= k_dim; q++) free(hh[q]);
        free(hh); free(rs); free(c); free(sn); free(hcol);
        return err;
    }
    int i = 0, j, k, iter = 0;
    double t, gamma, r_norm = 0.0, b_norm = 0.0, den_norm, epsilon, r_norm_0 = 0.0;
    double normup, normdown, conv_factor = 0.0;
    double true_norm = -1.0; /* ||b - A x|| of the current x when known */
    const int timed = s->prm.print_level > 0 || s->prof.on;
    s->t_prec_ms = 0.0;

#define TRY(x)                                                                                                         \
    do {                                                                                                               \
        err = (x);                                                                                                     \
        if (err) goto done;                                                                                            \
    } while (0)

    TRY(jxamd_event_record(s->ev0, s->stream));
    /* P0 = b - A x */
    TRY(bsr_apply(s, x, bvec, VEC(0), 1));
    TRY(host_norm2_enqueue(s, bvec, VEC(0)));
    /* Speculation: the first cycle's v_0 = r_0 / ||r_0|| (the norm read on
     * the device: 1 / sqrt of the same squared norm, as the host forms it)
     * and its step 1 go in before the norms reach the host, so the device
     * does not idle through that round trip.  A solve that stops before the
     * cycle (zero or converged initial residual) leaves the step unused. */
    int spec = 0;
    if (!timed && k_dim >= 1 && max_iter >= 1) {
        TRY(jxamd_k_scale(N, s->dsc + 2 * k_dim + 9, 0.0, VEC(0), s->stream));
        TRY(arnoldi_enqueue(s, 1, 1, timed));
        spec = 1;
    }
    TRY(host_norm2_read(s, &b_norm, &r_norm)); /* one wait for both */
    if (b_norm != 0.0 && (b_norm / b_norm) != (b_norm / b_norm)) { err = JXAMD_ERROR_GENERIC; goto done; }
    r_norm_0 = r_norm;
    if (r_norm != 0.0 && (r_norm / r_norm) != (r_norm / r_norm)) { err = JXAMD_ERROR_GENERIC; goto done; }
    normdown = r_norm;
    if (res_history) res_history[0] = r_norm;
    den_norm = b_norm > 0.0 ? b_norm : r_norm;
    epsilon = a_tol > r_tol * den_norm ? a_tol : r_tol * den_norm;

    while (iter < max_iter) {
        rs[0] = r_norm;
        if (r_norm == 0.0) goto done;
        if (r_norm <= epsilon) {
            /* false convergence 1 check (pgmres.c:1200-1241) */
            TRY(bsr_apply(s, x, bvec, s->Rv, 1));
            TRY(host_norm(s, s->Rv, &r_norm));
            if (r_norm <= epsilon) { converged = 1; true_norm = r_norm; break; }
            nfalse++;
            if (s->prm.is_check_restarted == 0) break;
            if (spec) {
                /* the speculative cycle start is not used: v_0 back to the
                 * unscaled b - A x (the same kernel, the same values) */
                TRY(bsr_apply(s, x, bvec, VEC(0), 1));
                spec = 0;
            }
        }
        if (!spec) TRY(jxamd_k_scale(N, NULL, 1.0 / r_norm, VEC(0), s->stream));
        i = 0;
        int ahead = spec; /* step i + 1 already enqueued */
        nprec += spec;
        spec = 0;
        while (i < k_dim && iter < max_iter) {
            i++;
            iter++;
            const int slot = i & 1;
            if (!ahead) {
                TRY(arnoldi_enqueue(s, i, slot, timed));
                nprec++;
            }
            ahead = 0;
            /* Look ahead: enqueue step i + 1 before reading step i's column
             * when step i is unlikely to end the cycle (the residual the last
             * contraction predicts stays 4x above the target), so the device
             * does not idle while the host runs the Givens update.  A step
             * enqueued in vain only writes v_{i+1}, which the update of x
             * does not read. */
            if (!timed && i < k_dim && iter < max_iter) {
                const double cf = conv_factor > 0.0 && conv_factor < 1.0 ? conv_factor : 0.5;
                if (r_norm * cf > 4.0 * epsilon) {
                    TRY(arnoldi_enqueue(s, i + 1, slot ^ 1, timed));
                    nprec++;
                    ahead = 1;
                }
            }
            if (timed) TRY(accumulate_precond_time(s)); /* also synchronises the stream */
            else TRY(column_wait(s, slot));
            memcpy(hcol, s->hpin + (size_t)slot * (k_dim + 2), sizeof(double) * (i + 1));
            for (j = 0; j < i; j++) hh[j][i - 1] = hcol[j];
            hh[i][i - 1] = sqrt(hcol[i]);
            for (j = 1; j < i; j++) {
                t = hh[j - 1][i - 1];
                hh[j - 1][i - 1] = sn[j - 1] * hh[j][i - 1] + c[j - 1] * t;
                hh[j][i - 1] = -sn[j - 1] * t + c[j - 1] * hh[j][i - 1];
            }
            t = hh[i][i - 1] * hh[i][i - 1];
            t += hh[i - 1][i - 1] * hh[i - 1][i - 1];
            gamma = sqrt(t);
            if (gamma == 0.0) gamma = epsmac;
            c[i - 1] = hh[i - 1][i - 1] / gamma;
            sn[i - 1] = hh[i][i - 1] / gamma;
            rs[i] = -hh[i][i - 1] * rs[i - 1];
            rs[i] /= gamma;
            rs[i - 1] = c[i - 1] * rs[i - 1];
            hh[i - 1][i - 1] = sn[i - 1] * hh[i][i - 1] + c[i - 1] * hh[i - 1][i - 1];
            r_norm = fabs(rs[i]);
            normup = r_norm;
            conv_factor = normup / normdown;
            normdown = normup;
            if (res_history) res_history[iter] = r_norm;
            if (s->prm.print_level > 0) printf("% 5d    %e    %e\n", iter, r_norm, b_norm > 0 ? r_norm / b_norm : r_norm);
            if (iter > 1 && conv_factor > convfac_threshold) { diverged = 1; break; }
            if (r_norm <= epsilon) break;
        }
        if (ahead) nprec--; /* a step enqueued ahead in vain is not an application of the method */
        /* back-solve and update x += M^{-1} (sum y_j v_j) (pgmres.c:1477-1503) */
        rs[i - 1] = rs[i - 1] / hh[i - 1][i - 1];
        for (k = i - 2; k >= 0; k--) {
            t = 0.0;
            for (j = k + 1; j < i; j++) t -= hh[k][j] * rs[j];
            t += rs[k];
            rs[k] = t / hh[k][k];
        }
        /* y goes to the device inside the launch when it is short enough */
        const int yargs = i <= JXAMD_COEFS_MAX;
        if (!yargs) TRY(jxamd_h2d(ydev, rs, sizeof(double) * i, s->stream));
        if (s->Z) {
            /* x += Z y (one pass: x + the sum, as the axpy of the sum gives) */
            int pq = prof_begin(s);
            if (yargs) {
                TRY(jxamd_k_lincomb_host(N, i, s->Z, N, rs, 1, x, s->stream));
                prof_end(s, pq, JXAMD_PROF_VEC, 8.0 * N * (i + 2));
            } else {
                TRY(jxamd_k_lincomb(N, i, s->Z, N, ydev, s->Wv, s->stream));
                prof_end(s, pq, JXAMD_PROF_VEC, 8.0 * N * (i + 1));
                TRY(jxamd_k_axpy(N, NULL, 1.0, 0, s->Wv, x, s->stream));
            }
        } else {
            int pq = prof_begin(s);
            if (yargs) TRY(jxamd_k_lincomb_host(N, i, V, N, rs, 0, s->Wv, s->stream));
            else TRY(jxamd_k_lincomb(N, i, V, N, ydev, s->Wv, s->stream));
            prof_end(s, pq, JXAMD_PROF_VEC, 8.0 * N * (i + 1));
            TRY(precond_timed(s, s->Wv, s->Rv, timed));
            nprec++;
            TRY(jxamd_k_axpy(N, NULL, 1.0, 0, s->Rv, x, s->stream));
            if (timed) TRY(accumulate_precond_time(s));
        }
        if (r_norm <= epsilon) {
            TRY(bsr_apply(s, x, bvec, s->Rv, 1));
            TRY(host_norm(s, s->Rv, &r_norm));
            if (r_norm <= epsilon) { converged = 1; true_norm = r_norm; break; }
            nfalse++;
            if (s->prm.is_check_restarted == 1) {
                /* false convergence 2: restart from the true residual */
                TRY(jxamd_k_copy(N, s->Rv, VEC(0), s->stream));
                i = 0;
                nrestart++;
            } else {
                break;
            }
        }
        if (iter > 1 && conv_factor > convfac_threshold) break;
        /* restart residual from the Givens recurrence (pgmres.c:1617-1637) */
        for (j = i; j > 0; j--) {
            rs[j - 1] = -sn[j - 1] * rs[j];
            rs[j] = c[j - 1] * rs[j];
        }
        if (i) TRY(jxamd_k_axpy(N, NULL, rs[i] - 1.0, 0, VEC(i), VEC(i), s->stream));
        for (j = i - 1; j > 0; j--) TRY(jxamd_k_axpy(N, NULL, rs[j], 0, VEC(j), VEC(i), s->stream));
        if (i) {
            TRY(jxamd_k_axpy(N, NULL, rs[0] - 1.0, 0, VEC(0), VEC(0), s->stream));
            TRY(jxamd_k_axpy(N, NULL, 1.0, 0, VEC(i), VEC(0), s->stream));
            nrestart++;
        }
    }
done:
    if (!err) err = jxamd_event_record(s->ev1, s->stream);
    if (!err) err = jxamd_d2h(s->hsm + 4, s->err, sizeof(int), s->stream); /* to pinned memory: no wait of its own */
    if (!err) err = jxamd_stream_sync(s->stream);
    if (!err) memcpy(&eflag, (const void *)(s->hsm + 4), sizeof(int));
    if (st) {
        memset(st, 0, sizeof(*st));
        st->iterations = iter;
        st->converged = converged;
        st->num_precond = nprec;
        st->num_restarts = nrestart;
        st->num_false_conv = nfalse;
        st->diverged = diverged;
        st->b_norm = b_norm;
        st->final_rel_res = b_norm > 0.0 ? r_norm / b_norm : (r_norm_0 > 0.0 ? r_norm / r_norm_0 : 0.0);
        float ms = 0.f;
        if (!err) jxamd_event_elapsed(&ms, s->ev0, s->ev1);
        st->t_solve_ms = ms;
        st->t_precond_ms = s->t_prec_ms;
        if (!err) {
            /* the loop's last check already holds ||b - A x|| of this x when
             * it ended on convergence */
            double tr = true_norm;
            int e2 = 0;
            if (tr < 0.0) {
                e2 = bsr_apply(s, x, bvec, s->Rv, 1);
                if (!e2) e2 = host_norm(s, s->Rv, &tr);
            }
            st->true_rel_res = b_norm > 0.0 ? tr / b_norm : tr;
        }
    }
    if (!err || err == JXAMD_ERROR_GENERIC) {
        /* (a give-up can also keep a one-launch MGS step from announcing its
         * column: column_wait then ends the solve with JXAMD_ERROR_GENERIC) */
        int ef = eflag;
        if (ef < 0 && jxamd_solver_error_flag(s, &ef)) ef = 0;
        if (ef) {
            /* a bounded wait gave up during this solve: its result is not
             * trusted, and the synchronisation state is reset for the next */
            err = JXAMD_ERROR_GENERIC;
            reset_sync_state(s);
        }
    }
    for (int q = 0; q <= k_dim; q++) free(hh[q]);
    free(hh); free(rs); free(c); free(sn); free(hcol);
#undef TRY
#undef VEC
    return err;
}

int jxamd_solver_solve(jxamd_solver *s, const double *b, double *x, double *res_history, jxamd_stats *st)
{
    if (!s || !b || !x) return JXAMD_ERROR_ARG;
    CK(jxamd_h2d(s->Bd, b, sizeof(double) * s->N, s->stream));
    CK(jxamd_h2d(s->Xd, x, sizeof(double) * s->N, s->stream));
    int e = jxamd_solver_solve_device(s, s->Bd, s->Xd, res_history, st);
    int e2 = jxamd_d2h(x, s->Xd, sizeof(double) * s->N, s->stream);
    if (!e2) e2 = jxamd_stream_sync(s->stream);
    return e ? e : e2;
}
