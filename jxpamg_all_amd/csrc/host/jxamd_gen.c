/*
 * Synthetic reservoir Jacobians in block-CSR form.
 *
 * The reference ships no matrix generator (its inputs are SPE10 files read by
 * jxf_BSRMatrixRead_Binary, jxfpamg-0220/csrc/mat_vec_data/seq_bsr_matrix.c:1935),
 * so the systems solved here are built from seeded permeability fields:
 *
 *   - structured nx*ny*nz grid, natural cell order c = i + nx*(j + ny*k);
 *   - one block row per cell, b unknowns per cell (pressure first);
 *   - 7-point stencil; in every block row the diagonal block comes first and
 *     the neighbour blocks follow in ascending column order (the reference's
 *     relaxers read the diagonal at A_i[i], par_relax_bsr.c:551, par_relax_6.c:233);
 *   - face transmissibility T = harmonic mean of the two cell permeabilities,
 *     vertical faces scaled by kz/kx;
 *   - pressure row: -T off the diagonal, sum(T) plus a small compressibility
 *     shift on it (diagonally dominant M-matrix);
 *   - saturation / component rows: accumulation dominated, neighbour couplings
 *     drawn from U(0.01, 0.1) scaled by T;
 *   - right-hand side U(-1, 1).
 *
 * Every random number is a pure function of (seed, stream, index) through a
 * splitmix64 hash, so the output does not depend on loop order or threading.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "jxamd.h"

static inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* uniform in (0,1), 53-bit resolution, never exactly 0 */
static inline double urand(uint64_t seed, uint64_t stream, uint64_t idx)
{
    uint64_t h = mix64(mix64(seed ^ mix64(stream * 0x632BE59BD9B4E019ULL)) + idx);
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

/* standard normal by Box-Muller on two independent uniforms */
static inline double nrand(uint64_t seed, uint64_t stream, uint64_t idx)
{
    double u1 = urand(seed, stream, 2 * idx);
    double u2 = urand(seed, stream, 2 * idx + 1);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

enum { ST_LNK = 1, ST_CHAN = 2, ST_CELL = 3, ST_FACE = 4, ST_RHS = 5, ST_WELL = 6 };

/* In-place separable Gaussian filter (std = corr cells, truncated at 3 std,
 * zero-flux ends by renormalising the kernel), then rescale to unit variance. */
static int gauss_filter3(double *f, int nx, int ny, int nz, double corr, int filter_z)
{
    int r = (int)ceil(3.0 * corr);
    if (r < 1) return 0;
    double *w = (double *)malloc(sizeof(double) * (2 * r + 1));
    int nmax = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;
    double *line = (double *)malloc(sizeof(double) * nmax);
    double *out = (double *)malloc(sizeof(double) * nmax);
    if (!w || !line || !out) { free(w); free(line); free(out); return JXAMD_ERROR_MEMORY; }
    for (int t = -r; t <= r; t++) w[t + r] = exp(-0.5 * (double)t * t / (corr * corr));

    for (int axis = 0; axis < 3; axis++) {
        if (axis == 2 && !filter_z) break;
        int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
        long stride = axis == 0 ? 1 : (axis == 1 ? nx : (long)nx * ny);
        long nlines = (long)nx * ny * nz / n;
        for (long l = 0; l < nlines; l++) {
            long base;
            if (axis == 0) base = l * nx;
            else if (axis == 1) base = (l / nx) * (long)nx * ny + (l % nx);
            else base = l;
            for (int q = 0; q < n; q++) line[q] = f[base + q * stride];
            for (int q = 0; q < n; q++) {
                double s = 0.0, ws = 0.0;
                for (int t = -r; t <= r; t++) {
                    int p = q + t;
                    if (p < 0 || p >= n) continue;
                    s += w[t + r] * line[p];
                    ws += w[t + r];
                }
                out[q] = s / ws;
            }
            for (int q = 0; q < n; q++) f[base + q * stride] = out[q];
        }
    }
    long ncell = (long)nx * ny * nz;
    double mean = 0.0, var = 0.0;
    for (long c = 0; c < ncell; c++) mean += f[c];
    mean /= (double)ncell;
    for (long c = 0; c < ncell; c++) var += (f[c] - mean) * (f[c] - mean);
    var /= (double)ncell;
    double sd = var > 0.0 ? sqrt(var) : 1.0;
    for (long c = 0; c < ncell; c++) f[c] = (f[c] - mean) / sd;
    free(w); free(line); free(out);
    return 0;
}

/* ln k per cell, by field kind */
static int build_lnk(const jxamd_gen_params *p, double *lnk)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const long ncell = (long)nx * ny * nz;
    const uint64_t seed = p->seed;

    if (p->field == JXAMD_FIELD_IID) {
        for (long c = 0; c < ncell; c++) lnk[c] = p->sigma_lnk * nrand(seed, ST_LNK, (uint64_t)c);
        return 0;
    }
    if (p->field == JXAMD_FIELD_SMOOTH) {
        for (long c = 0; c < ncell; c++) lnk[c] = nrand(seed, ST_LNK, (uint64_t)c);
        int e = gauss_filter3(lnk, nx, ny, nz, p->corr_len, 1);
        if (e) return e;
        for (long c = 0; c < ncell; c++) lnk[c] *= p->sigma_lnk;
        return 0;
    }
    if (p->field == JXAMD_FIELD_LAYERED) {
        /* top n_smooth layers (k = 0 .. n_smooth-1): 2-D smooth lognormal per
         * layer block filtered in x,y and z; the rest: sinuous channels. */
        int ns = p->n_smooth_layers < nz ? p->n_smooth_layers : nz;
        if (ns > 0) {
            long nsc = (long)nx * ny * ns;
            for (long c = 0; c < nsc; c++) lnk[c] = nrand(seed, ST_LNK, (uint64_t)c);
            int e = gauss_filter3(lnk, nx, ny, ns, p->corr_len, 1);
            if (e) return e;
            for (long c = 0; c < nsc; c++) lnk[c] *= p->sigma_lnk;
        }
        const double lnratio = log(p->channel_ratio > 0.0 ? p->channel_ratio : 1.0);
        for (int k = ns; k < nz; k++) {
            /* channel set changes every 5 layers, meanders drift per layer */
            int grp = (k - ns) / 5;
            int nch = 2 + (int)(3.0 * urand(seed, ST_CHAN, (uint64_t)grp * 64));
            for (long c = (long)k * nx * ny; c < (long)(k + 1) * nx * ny; c++) lnk[c] = 0.0;
            for (int ch = 0; ch < nch; ch++) {
                uint64_t b = (uint64_t)grp * 64 + 1 + (uint64_t)ch * 8;
                double x0 = nx * urand(seed, ST_CHAN, b + 0);
                double amp = 3.0 + 7.0 * urand(seed, ST_CHAN, b + 1);
                double lam = 40.0 + 80.0 * urand(seed, ST_CHAN, b + 2);
                double phi = 6.283185307179586 * urand(seed, ST_CHAN, b + 3);
                double wid = 3.0 + 5.0 * urand(seed, ST_CHAN, b + 4);
                double drift = 0.15 * (double)((k - ns) % 5);
                for (int j = 0; j < ny; j++) {
                    double xc = x0 + amp * sin(6.283185307179586 * j / lam + phi + drift);
                    for (int i = 0; i < nx; i++) {
                        if (fabs((double)i - xc) <= 0.5 * wid)
                            lnk[(long)k * nx * ny + (long)j * nx + i] = lnratio;
                    }
                }
            }
        }
        return 0;
    }
    return JXAMD_ERROR_ARG;
}

int jxamd_gen_sizes(const jxamd_gen_params *p, int *nb, int *nnzb)
{
    if (!p || p->nx < 1 || p->ny < 1 || p->nz < 1) return JXAMD_ERROR_ARG;
    long nx = p->nx, ny = p->ny, nz = p->nz;
    long ncell = nx * ny * nz;
    long faces = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1);
    long nnz = ncell + 2 * faces;
    if (ncell > 0x7fffffffL || nnz > 0x7fffffffL) return JXAMD_ERROR_ARG;
    *nb = (int)ncell;
    *nnzb = (int)nnz;
    return 0;
}

int jxamd_gen_reservoir(const jxamd_gen_params *p, int *ia, int *ja, double *a, double *rhs)
{
    int nb, nnzb;
    int e = jxamd_gen_sizes(p, &nb, &nnzb);
    if (e) return e;
    const int b = p->block_size;
    if (b < 2 || b > JXAMD_MAX_BLOCK) return JXAMD_ERROR_ARG;
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const long nxy = (long)nx * ny;
    const int b2 = b * b;
    const uint64_t seed = p->seed;

    double *lnk = (double *)malloc(sizeof(double) * (size_t)nb);
    if (!lnk) return JXAMD_ERROR_MEMORY;
    e = build_lnk(p, lnk);
    if (e) { free(lnk); return e; }
    for (int c = 0; c < nb; c++) lnk[c] = exp(lnk[c]); /* now k */

    /* mean face transmissibility: scale for the accumulation and well terms */
    double tsum = 0.0;
    long tcount = 0;
    for (int c = 0; c < nb; c++) {
        int i = c % nx;
        int j = (c / nx) % ny;
        int k = c / (int)nxy;
        if (i + 1 < nx) { tsum += 2.0 * lnk[c] * lnk[c + 1] / (lnk[c] + lnk[c + 1]); tcount++; }
        if (j + 1 < ny) { tsum += 2.0 * lnk[c] * lnk[c + nx] / (lnk[c] + lnk[c + nx]); tcount++; }
        if (k + 1 < nz) { tsum += p->kz_ratio * 2.0 * lnk[c] * lnk[c + nxy] / (lnk[c] + lnk[c + nxy]); tcount++; }
    }
    const double tmean = tcount ? tsum / (double)tcount : 1.0;

    /* well columns: boost the pressure diagonal over the full height */
    int nw = p->n_wells < 16 ? p->n_wells : 16;
    int wi[16], wj[16];
    for (int w = 0; w < nw; w++) {
        wi[w] = (int)(nx * urand(seed, ST_WELL, 2 * (uint64_t)w));
        wj[w] = (int)(ny * urand(seed, ST_WELL, 2 * (uint64_t)w + 1));
    }

    int pos = 0;
    ia[0] = 0;
    for (int c = 0; c < nb; c++) {
        int i = c % nx;
        int j = (c / nx) % ny;
        int k = c / (int)nxy;
        int nbr[6];
        double tr[6];
        int nn = 0;
        /* ascending column order: -z, -y, -x, +x, +y, +z */
        if (k > 0)      { nbr[nn] = (int)(c - nxy); tr[nn++] = p->kz_ratio; }
        if (j > 0)      { nbr[nn] = c - nx;         tr[nn++] = 1.0; }
        if (i > 0)      { nbr[nn] = c - 1;          tr[nn++] = 1.0; }
        if (i + 1 < nx) { nbr[nn] = c + 1;          tr[nn++] = 1.0; }
        if (j + 1 < ny) { nbr[nn] = c + nx;         tr[nn++] = 1.0; }
        if (k + 1 < nz) { nbr[nn] = (int)(c + nxy); tr[nn++] = p->kz_ratio; }
        for (int q = 0; q < nn; q++) {
            int d = nbr[q];
            tr[q] *= 2.0 * lnk[c] * lnk[d] / (lnk[c] + lnk[d]);
        }

        /* per-cell mobility fractions of the secondary equations */
        double frac[JXAMD_MAX_BLOCK];
        for (int r = 1; r < b; r++) frac[r] = 0.2 + 0.4 * urand(seed, ST_CELL, (uint64_t)c * 16 + r);

        double *D = a + (size_t)pos * b2;
        memset(D, 0, sizeof(double) * b2);
        ja[pos++] = c;
        double tsum_c = 0.0;
        for (int q = 0; q < nn; q++) {
            int d = nbr[q];
            double T = tr[q];
            tsum_c += T;
            double *B = a + (size_t)pos * b2;
            ja[pos++] = d;
            /* face-local couplings, indexed by the directed face (c -> d) */
            uint64_t fidx = ((uint64_t)c * 6 + (uint64_t)q) * (uint64_t)b2;
            for (int r = 0; r < b; r++) {
                for (int s = 0; s < b; s++) {
                    double v;
                    if (r == 0 && s == 0) v = -T;
                    else if (s == 0) v = -T * frac[r];
                    else v = -T * (0.01 + 0.09 * urand(seed, ST_FACE, fidx + (uint64_t)(r * b + s)));
                    B[r * b + s] = v;
                    /* the diagonal block carries the flux balance of the same terms */
                    D[r * b + s] -= (r == 0 || s == 0 || r == s) ? v : 0.5 * v;
                }
            }
        }
        /* pressure: compressibility shift; others: accumulation dominates */
        D[0] += 1e-3 * (tsum_c > 0.0 ? tsum_c : tmean);
        for (int r = 1; r < b; r++)
            D[r * b + r] += (1.0 + urand(seed, ST_CELL, (uint64_t)c * 16 + 8 + r)) * (tsum_c + tmean);
        for (int w = 0; w < nw; w++)
            if (i == wi[w] && j == wj[w]) D[0] += p->well_boost * tmean;
        ia[c + 1] = pos;
    }
    free(lnk);

    if (rhs) {
        long n = (long)nb * b;
        for (long r = 0; r < n; r++) rhs[r] = 2.0 * urand(seed, ST_RHS, (uint64_t)r) - 1.0;
    }
    return 0;
}
