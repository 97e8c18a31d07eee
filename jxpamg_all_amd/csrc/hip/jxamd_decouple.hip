// jxamd_decouple.hip -- decoupling and pressure extraction on device arrays
// in the reference BSR layout (ia[nb+1], ja[nnzb], row-major b x b blocks,
// diagonal block first), bit for bit with the reference's
// jxf_BSRMatrixDecouple (jxfpamg-0220/csrc/cpr/jxf_bsr_decoup.c:1273-1348)
// and jxf_BSRMatrixGetSubmatrix (csrc/mat_vec_data/seq_bsr_matrix.c:214-250).
//
// Every decoupling type is a per-block-row transformation: a b x b matrix M_i
// (or a few coefficients) derived from block row i is applied to every block
// of the row and to the row's right-hand side.  k_decouple_apply runs one
// wavefront over 64 consecutive block rows: one lane per row forms M_i and
// updates the rhs, then the wavefront streams the rows' blocks through its
// own LDS tile of 64 blocks (coalesced loads and stores of the whole value
// range, one lane per block for the arithmetic), so the matrix is read and
// written once.  True-IMPES needs the column sums of A first
// (k_timpes_coef, one thread per block column, gathering the column's blocks
// through the transpose pattern of a structurally symmetric matrix) and
// stores its coefficients per row in a work buffer.
//
// The arithmetic follows the reference expression by expression (closed-form
// inverses, product orders, the in-place right-hand-side products whose
// output aliases their input); no multiply-add is contracted, so results are
// the reference's own roundings.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

extern "C" {
#include "jxamd_dev.h"
}

#pragma clang fp contract(off)

#define WAVE_DC 64    // lanes per wavefront: block rows per wavefront and blocks per LDS tile
#define DC_ROWS 256   // block rows per workgroup (four wavefronts)
#define TC_MAXD 16    // rows of at most this many blocks: True-IMPES column sums from registers

namespace {

// HIP status -> error flag; with JXAMD_DEBUG set, failures are named on stderr
int dc_chk(hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    if (getenv("JXAMD_DEBUG")) fprintf(stderr, "jxamd decouple: %s: %s\n", what, hipGetErrorName(e));
    return JXAMD_ERROR_DEVICE;
}

// closed-form inverses, jxf_bsr_decoup.c:53-183 (the 4x4 one is the
// reference's: diagonal cofactors right, off-diagonal ones transposed with
// the wrong sign pattern, a determinant pairing a12..a14 with minors that
// contain the first row; it is what the reference decouples with)
__device__ void inv1(double *a) { a[0] = fabs(a[0]) > 1e-15 ? 1.0 / a[0] : 1.0; }

__device__ void inv2(double *a)
{
    const double a11 = a[0], a12 = a[1], a21 = a[2], a22 = a[3];
    const double det = a11 * a22 - a12 * a21;
    if (fabs(det) < 1e-15) {
        a[0] = 1.0; a[1] = 0.0; a[2] = 0.0; a[3] = 1.0;
    } else {
        const double di = 1.0 / det;
        a[0] = a22 * di; a[1] = -a12 * di; a[2] = -a21 * di; a[3] = a11 * di;
    }
}

__device__ void inv3(double *a)
{
    const double a11 = a[0], a12 = a[1], a13 = a[2];
    const double a21 = a[3], a22 = a[4], a23 = a[5];
    const double a31 = a[6], a32 = a[7], a33 = a[8];
    const double M11 = a22 * a33 - a23 * a32, M12 = a21 * a33 - a23 * a31, M13 = a21 * a32 - a22 * a31;
    const double M21 = a12 * a33 - a13 * a32, M22 = a11 * a33 - a13 * a31, M23 = a11 * a32 - a12 * a31;
    const double M31 = a12 * a23 - a13 * a22, M32 = a11 * a23 - a13 * a21, M33 = a11 * a22 - a12 * a21;
    const double det = a11 * M11 - a12 * M12 + a13 * M13;
    if (fabs(det) < 1e-15) {
        for (int i = 0; i < 9; i++) a[i] = 0.0;
        a[0] = a[4] = a[8] = 1.0;
        return;
    }
    const double di = 1.0 / det;
    a[0] = M11 * di; a[1] = -M21 * di; a[2] = M31 * di;
    a[3] = -M12 * di; a[4] = M22 * di; a[5] = -M32 * di;
    a[6] = M13 * di; a[7] = -M23 * di; a[8] = M33 * di;
}

__device__ void inv4(double *a)
{
    const double a11 = a[0], a12 = a[1], a13 = a[2], a14 = a[3];
    const double a21 = a[4], a22 = a[5], a23 = a[6], a24 = a[7];
    const double a31 = a[8], a32 = a[9], a33 = a[10], a34 = a[11];
    const double a41 = a[12], a42 = a[13], a43 = a[14], a44 = a[15];
    const double M11 = a22 * a33 * a44 + a23 * a34 * a42 + a24 * a32 * a43 - a22 * a34 * a43 - a23 * a32 * a44 - a24 * a33 * a42;
    const double M12 = a12 * a34 * a43 + a13 * a32 * a44 + a14 * a33 * a42 - a12 * a33 * a44 - a13 * a34 * a42 - a14 * a32 * a43;
    const double M13 = a12 * a23 * a44 + a13 * a24 * a42 + a14 * a22 * a43 - a12 * a24 * a43 - a13 * a22 * a44 - a14 * a23 * a42;
    const double M14 = a12 * a24 * a33 + a13 * a22 * a34 + a14 * a23 * a32 - a12 * a23 * a34 - a13 * a24 * a32 - a14 * a22 * a33;
    const double det = a11 * M11 - a12 * M12 + a13 * M13 - a14 * M14;
    if (fabs(det) < 1e-15) {
        for (int i = 0; i < 16; i++) a[i] = 0.0;
        a[0] = a[5] = a[10] = a[15] = 1.0;
        return;
    }
    const double di = 1.0 / det;
    const double M21 = a21 * a34 * a43 + a23 * a31 * a44 + a24 * a33 * a41 - a21 * a33 * a44 - a23 * a34 * a41 - a24 * a31 * a43;
    const double M22 = a11 * a33 * a44 + a13 * a34 * a41 + a14 * a31 * a43 - a11 * a34 * a43 - a13 * a31 * a44 - a14 * a33 * a41;
    const double M23 = a11 * a24 * a43 + a13 * a21 * a44 + a14 * a23 * a41 - a11 * a23 * a44 - a13 * a24 * a41 - a14 * a21 * a43;
    const double M24 = a11 * a23 * a34 + a13 * a24 * a31 + a14 * a21 * a33 - a11 * a24 * a33 - a13 * a21 * a34 - a14 * a23 * a31;
    const double M31 = a21 * a32 * a44 + a22 * a34 * a41 + a24 * a31 * a42 - a21 * a34 * a42 - a22 * a31 * a44 - a24 * a32 * a41;
    const double M32 = a11 * a34 * a42 + a12 * a31 * a44 + a14 * a32 * a41 - a11 * a32 * a44 - a12 * a34 * a41 - a14 * a31 * a42;
    const double M33 = a11 * a22 * a44 + a12 * a24 * a41 + a14 * a21 * a42 - a11 * a24 * a42 - a12 * a21 * a44 - a14 * a22 * a41;
    const double M34 = a11 * a24 * a32 + a12 * a21 * a34 + a14 * a22 * a31 - a11 * a22 * a34 - a12 * a24 * a31 - a14 * a21 * a32;
    const double M41 = a21 * a33 * a42 + a22 * a31 * a43 + a23 * a32 * a41 - a21 * a32 * a43 - a22 * a33 * a41 - a23 * a31 * a42;
    const double M42 = a11 * a32 * a43 + a12 * a33 * a41 + a13 * a31 * a42 - a11 * a33 * a42 - a12 * a31 * a43 - a13 * a32 * a41;
    const double M43 = a11 * a23 * a42 + a12 * a21 * a43 + a13 * a22 * a41 - a11 * a22 * a43 - a12 * a23 * a41 - a13 * a21 * a42;
    const double M44 = a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a11 * a23 * a32 - a12 * a21 * a33 - a13 * a22 * a31;
    a[0] = M11 * di;  a[1] = -M21 * di; a[2] = M31 * di;   a[3] = -M41 * di;
    a[4] = -M12 * di; a[5] = M22 * di;  a[6] = -M32 * di;  a[7] = M42 * di;
    a[8] = M13 * di;  a[9] = -M23 * di; a[10] = M33 * di;  a[11] = -M43 * di;
    a[12] = -M14 * di; a[13] = M24 * di; a[14] = -M34 * di; a[15] = M44 * di;
}

// jxf_smat_inv_nc for b <= 4 (jxf_bsr_decoup.c:186-198)
template <int B> __device__ void inv_nc(double *a)
{
    if (B == 1) inv1(a);
    else if (B == 2) inv2(a);
    else if (B == 3) inv3(a);
    else inv4(a);
}

// blk <- M blk, jxf_smat_mul (jxf_bsr_decoup.c:286-296): every entry summed from zero
template <int B> __device__ void mul_generic(const double *M, double *blk)
{
    double t[B * B];
#pragma unroll
    for (int r = 0; r < B; r++)
#pragma unroll
        for (int c = 0; c < B; c++) {
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < B; q++) acc += M[r * B + q] * blk[q * B + c];
            t[r * B + c] = acc;
        }
#pragma unroll
    for (int e = 0; e < B * B; e++) blk[e] = t[e];
}

// v <- M v with output aliasing input, jxf_smat_vec_mul(M, v, v)
// (jxf_bsr_decoup.c:335-343): each component is cleared, then accumulated
// over the current (partly updated) vector
template <int B> __device__ void vec_inplace(const double *M, double *v)
{
#pragma unroll
    for (int r = 0; r < B; r++) {
        v[r] = 0.0;
#pragma unroll
        for (int q = 0; q < B; q++) v[r] += M[r * B + q] * v[q];
    }
}

// The row transformation of block row i (ABF, ANL, SEM, QI) from its
// diagonal block D; True-IMPES rows take their coefficients from coef.
template <int B, int T>
__device__ void row_matrix(const double *D, int thermal, const double *coef, double *M)
{
    constexpr int BB = B * B;
    if (T == JXAMD_DECOUP_ABF) {
#pragma unroll
        for (int e = 0; e < BB; e++) M[e] = D[e];
        inv_nc<B>(M);
    } else if (T == JXAMD_DECOUP_SEM) {
        // jxf_bsr_decoup.c:606-632
#pragma unroll
        for (int e = 0; e < BB; e++) M[e] = D[e];
        inv_nc<B>(M);
        if (thermal) {
            M[0] = 1.0;
            M[B - 1] = 0.0;
            M[(B - 1) * B] = 0.0;
            M[BB - 1] = 1.0;
            for (int l = 0; l < B - 2; l++) {
                M[1 + l] = -D[1 + l];
                M[(B - 1) * B + 1 + l] = -D[(B - 1) * B + 1 + l];
            }
        } else {
            M[0] = 1.0;
            for (int l = 0; l < B - 1; l++) M[1 + l] = -D[1 + l];
        }
    } else if (T == JXAMD_DECOUP_ANL || T == JXAMD_DECOUP_QI) {
        // jxf_bsr_decoup.c:541-563 (ANL), 672-691 (QI, no guard on the divisor)
#pragma unroll
        for (int r = 0; r < B; r++)
#pragma unroll
            for (int c = 0; c < B; c++) M[r * B + c] = r == c ? 1.0 : 0.0;
        const int nl = thermal ? B - 2 : B - 1;
        for (int l = 0; l < nl; l++) {
            if (T == JXAMD_DECOUP_ANL) {
                M[1 + l] = -D[1 + l];
                if (thermal) M[(B - 1) * B + 1 + l] = -D[(B - 1) * B + 1 + l];
            } else {
                M[1 + l] = -D[1 + l] / D[(l + 1) * B + l + 1];
                if (thermal) M[(B - 1) * B + 1 + l] = -D[(B - 1) * B + 1 + l] / D[(l + 1) * B + l + 1];
            }
        }
    } else {
        // True-IMPES: the row's coefficients
#pragma unroll
        for (int q = 0; q < B; q++) M[q] = coef[q];
    }
}

// one block of row i under the row's transformation M (diag: the row's first block)
template <int B, int T> __device__ void apply_block(const double *M, bool diag, double *blk)
{
    constexpr int BB = B * B;
    if (T == JXAMD_DECOUP_ABF) {
        if (diag) {
#pragma unroll
            for (int r = 0; r < B; r++)
#pragma unroll
                for (int c = 0; c < B; c++) blk[r * B + c] = r == c ? 1.0 : 0.0;
        } else if (B == 2) { // jxf_smat_mul_2x2
            double t[4];
            t[0] = M[0] * blk[0] + M[1] * blk[2];
            t[1] = M[0] * blk[1] + M[1] * blk[3];
            t[2] = M[2] * blk[0] + M[3] * blk[2];
            t[3] = M[2] * blk[1] + M[3] * blk[3];
#pragma unroll
            for (int e = 0; e < 4; e++) blk[e] = t[e];
        } else if (B == 4) { // jxf_smat_mul_4x4: one left-to-right expression per entry
            double t[16];
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int c = 0; c < 4; c++)
                    t[r * 4 + c] = M[r * 4 + 0] * blk[0 * 4 + c] + M[r * 4 + 1] * blk[1 * 4 + c] +
                                   M[r * 4 + 2] * blk[2 * 4 + c] + M[r * 4 + 3] * blk[3 * 4 + c];
#pragma unroll
            for (int e = 0; e < 16; e++) blk[e] = t[e];
        } else {
            mul_generic<B>(M, blk);
        }
    } else if (T == JXAMD_DECOUP_ANL || T == JXAMD_DECOUP_SEM || T == JXAMD_DECOUP_QI) {
        mul_generic<B>(M, blk);
    } else if (T == JXAMD_DECOUP_TIMPES) {
        if (B == 2) { // jxf_bsr_decoup.c:791-798
            blk[0] = blk[0] + M[0] * blk[2];
            blk[1] = blk[1] + M[0] * blk[3];
        } else if (B == 3) { // :962-975, the third coefficient unused
            const double n0 = M[0] * blk[0] + M[1] * blk[3];
            const double n1 = M[0] * blk[1] + M[1] * blk[4];
            const double n2 = M[0] * blk[2] + M[1] * blk[5];
            blk[0] = n0; blk[1] = n1; blk[2] = n2;
        } else { // :1051-1064
            double n[B];
#pragma unroll
            for (int c = 0; c < B; c++) {
                n[c] = 0.0;
#pragma unroll
                for (int q = 0; q < B; q++) n[c] += M[q] * blk[q * B + c];
            }
#pragma unroll
            for (int c = 0; c < B; c++) blk[c] = n[c];
        }
    } else { // True-IMPES, second form
        if (B == 2) { // :1157-1163
            blk[0] = blk[0] + M[0] * blk[2];
            blk[1] = blk[1] + M[0] * blk[3];
        } else if (B == 3) { // :1197-1203
            blk[0] = blk[0] + M[0] * blk[3] + M[1] * blk[6];
            blk[1] = blk[1] + M[0] * blk[4] + M[1] * blk[7];
            blk[2] = blk[2] + M[0] * blk[5] + M[1] * blk[8];
        } else { // :1246-1253
#pragma unroll
            for (int c = 0; c < B; c++) {
                double sum = blk[c];
#pragma unroll
                for (int q = 0; q < B - 1; q++) sum += M[q] * blk[(q + 1) * B + c];
                blk[c] = sum;
            }
        }
    }
    (void)BB;
}

// the row's right-hand side under the row's transformation
template <int B, int T> __device__ void apply_rhs(const double *M, double *v)
{
    if (T == JXAMD_DECOUP_ABF) {
        if (B == 2) { // jxf_smat_vec_mul_2 in place
            v[0] = M[0] * v[0] + M[1] * v[1];
            v[1] = M[2] * v[0] + M[3] * v[1];
        } else if (B == 4) { // jxf_smat_vec_mul_4 in place
            v[0] = M[0] * v[0] + M[1] * v[1] + M[2] * v[2] + M[3] * v[3];
            v[1] = M[4] * v[0] + M[5] * v[1] + M[6] * v[2] + M[7] * v[3];
            v[2] = M[8] * v[0] + M[9] * v[1] + M[10] * v[2] + M[11] * v[3];
            v[3] = M[12] * v[0] + M[13] * v[1] + M[14] * v[2] + M[15] * v[3];
        } else {
            vec_inplace<B>(M, v);
        }
    } else if (T == JXAMD_DECOUP_ANL || T == JXAMD_DECOUP_SEM || T == JXAMD_DECOUP_QI) {
        vec_inplace<B>(M, v);
    } else if (T == JXAMD_DECOUP_TIMPES) {
        if (B == 2) v[0] = v[0] + M[0] * v[1];
        else if (B == 3) v[0] = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
        else {
            double n0 = 0.0;
#pragma unroll
            for (int q = 0; q < B; q++) n0 += M[q] * v[q];
            v[0] = n0;
        }
    } else {
        if (B == 2) v[0] = v[0] + M[0] * v[1];
        else if (B == 3) v[0] = v[0] + M[0] * v[1] + M[1] * v[2];
        // other sizes: the second form leaves the rhs unchanged
    }
}

// LDS exchange between the lanes of one wavefront: the fences keep the
// compiler from moving LDS accesses across the point; a wavefront's LDS
// operations complete in order
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wavefront per 64 consecutive block rows, the wavefronts of a workgroup
// independent (no workgroup barrier).  Lane l forms M of row r0 + l and
// updates the row's rhs.  The wavefront then streams its rows' blocks
// [ia[r0], ia[r0 + 64]) through its own LDS tile of 64 blocks: each lane
// loads B*B values of the tile (coalesced; the next tile's loads are issued
// before the current tile is transformed), transforms one block in LDS with
// M of the block's row (read from the lane that formed it), and the tile is
// stored back coalesced.  The matrix is read and written once.
template <int B, int T>
__global__ void __launch_bounds__(256) k_decouple_apply(int nb, const int *__restrict__ ia, double *a, double *rhs,
                                                        int thermal, const double *__restrict__ coef, int pidx,
                                                        double *__restrict__ p_a)
{
    constexpr int BB = B * B;
    constexpr int WPB = 4; // wavefronts per workgroup
    __shared__ double sv[WPB][WAVE_DC * BB];
    __shared__ int sia[WPB][WAVE_DC + 1];
    const int w = threadIdx.x / WAVE_DC, lane = threadIdx.x % WAVE_DC;
    const int r0 = (blockIdx.x * WPB + w) * WAVE_DC;
    if (r0 >= nb) return;
    const int nr = min(WAVE_DC, nb - r0);
    sia[w][lane] = ia[r0 + min(lane, nr)];
    if (lane == 0) sia[w][WAVE_DC] = ia[r0 + nr];
    wave_sync();
    double M[BB];
#pragma unroll
    for (int e = 0; e < BB; e++) M[e] = 0.0;
    if (lane < nr) {
        const int i = r0 + lane;
        double D[BB];
        if (T == JXAMD_DECOUP_TIMPES || T == JXAMD_DECOUP_TIMPES2) {
            row_matrix<B, T>(D, thermal, coef + (size_t)i * B, M);
        } else {
            const double *d = a + (size_t)sia[w][lane] * BB;
#pragma unroll
            for (int e = 0; e < BB; e++) D[e] = d[e];
            row_matrix<B, T>(D, thermal, nullptr, M);
        }
        if (rhs) {
            double v[B];
#pragma unroll
            for (int q = 0; q < B; q++) v[q] = rhs[(size_t)i * B + q];
            apply_rhs<B, T>(M, v);
#pragma unroll
            for (int q = 0; q < B; q++) rhs[(size_t)i * B + q] = v[q];
        }
    }
    const int kb = sia[w][0], ke = sia[w][nr];
    const size_t vend = (size_t)ke * BB;
    double buf[BB];
#pragma unroll
    for (int e = 0; e < BB; e++) {
        const size_t x = (size_t)kb * BB + (size_t)e * WAVE_DC + lane;
        buf[e] = x < vend ? a[x] : 0.0;
    }
    for (int k0 = kb; k0 < ke; k0 += WAVE_DC) {
        const int nk = min(WAVE_DC, ke - k0);
#pragma unroll
        for (int e = 0; e < BB; e++) sv[w][e * WAVE_DC + lane] = buf[e];
        wave_sync();
        if (k0 + WAVE_DC < ke) {
#pragma unroll
            for (int e = 0; e < BB; e++) {
                const size_t x = (size_t)(k0 + WAVE_DC) * BB + (size_t)e * WAVE_DC + lane;
                buf[e] = x < vend ? a[x] : 0.0;
            }
        }
        // the row holding block k: the last of the wave's rows whose start is <= k
        const int k = k0 + lane;
        int lo = 0, hi = nr - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (sia[w][mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        double Mr[BB];
#pragma unroll
        for (int e = 0; e < BB; e++) Mr[e] = __shfl(M[e], lo, WAVE_DC);
        if (lane < nk) {
            double blk[BB];
#pragma unroll
            for (int e = 0; e < BB; e++) blk[e] = sv[w][lane * BB + e];
            apply_block<B, T>(Mr, k == sia[w][lo], blk);
#pragma unroll
            for (int e = 0; e < BB; e++) sv[w][lane * BB + e] = blk[e];
            if (p_a) p_a[k] = blk[pidx * B + pidx];
        }
        wave_sync();
        const size_t tend = (size_t)(k0 + nk) * BB;
#pragma unroll
        for (int e = 0; e < BB; e++) {
            const size_t x = (size_t)k0 * BB + (size_t)e * WAVE_DC + lane;
            if (x < tend) a[x] = sv[w][e * WAVE_DC + lane];
        }
        wave_sync();
    }
}

// True-IMPES coefficients of row c from the column sums Ac_c = sum_i A_ic
// (rows in ascending order, from zero; jxf_bsr_decoup.c:727-746, 917-935):
// with a structurally symmetric pattern the rows holding column c are the
// columns of row c.  A missing transpose entry sets *err and leaves coef.
template <int B, int SECOND>
__global__ void __launch_bounds__(256) k_timpes_coef(int nb, const int *__restrict__ ia, const int *__restrict__ ja,
                                                     const double *__restrict__ a, double *__restrict__ coef,
                                                     int *err)
{
    constexpr int BB = B * B;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nb; c += gridDim.x * blockDim.x) {
        double ac[BB];
#pragma unroll
        for (int e = 0; e < BB; e++) ac[e] = 0.0;
        const int cb = ia[c], ce = ia[c + 1];
        bool ok = true;
        if (ce - cb <= TC_MAXD) {
            // short rows: the row's distinct columns sorted in registers, then
            // every transpose search advanced one entry per step for all
            // columns at once (independent loads), then the sum in ascending
            // row order
            int cols[TC_MAXD], rb[TC_MAXD], re[TC_MAXD], pos[TC_MAXD];
            const int deg = ce - cb;
#pragma unroll
            for (int n = 0; n < TC_MAXD; n++) cols[n] = n < deg ? ja[cb + n] : 0x7fffffff;
            // insertion sort, repeated columns dropped
#pragma unroll
            for (int n = 1; n < TC_MAXD; n++) {
                const int v = cols[n];
                int q = n - 1;
                while (q >= 0 && cols[q] > v) {
                    cols[q + 1] = cols[q];
                    q--;
                }
                cols[q + 1] = v;
            }
            int nd = 0;
#pragma unroll
            for (int n = 0; n < TC_MAXD; n++)
                if (cols[n] != 0x7fffffff && (nd == 0 || cols[n] != cols[nd - 1])) cols[nd++] = cols[n];
            int maxlen = 0;
            for (int n = 0; n < nd; n++) {
                if (cols[n] < 0 || cols[n] >= nb) ok = false;
                rb[n] = ok ? ia[cols[n]] : 0;
                re[n] = ok ? ia[cols[n] + 1] : 0;
                pos[n] = -1;
                if (re[n] - rb[n] > maxlen) maxlen = re[n] - rb[n];
            }
            for (int t = 0; t < maxlen && ok; t++)
                for (int n = 0; n < nd; n++)
                    if (pos[n] < 0 && rb[n] + t < re[n] && ja[rb[n] + t] == c) pos[n] = rb[n] + t;
            for (int n = 0; n < nd && ok; n++) {
                if (pos[n] < 0) { ok = false; break; }
                const double *blk = a + (size_t)pos[n] * BB;
#pragma unroll
                for (int e = 0; e < BB; e++) ac[e] += blk[e];
            }
        }
        int prev = -1;
        for (int n = 0; ce - cb > TC_MAXD && n < ce - cb && ok; n++) {
            // the next row in ascending order among row c's columns
            int i = 0x7fffffff;
            for (int k = cb; k < ce; k++) {
                const int j = ja[k];
                if (j > prev && j < i) i = j;
            }
            if (i == 0x7fffffff) break; // repeated columns: every distinct row visited
            prev = i;
            if (i < 0 || i >= nb) { ok = false; break; }
            int pos = -1;
            for (int k = ia[i]; k < ia[i + 1]; k++)
                if (ja[k] == c) { pos = k; break; }
            if (pos < 0) { ok = false; break; }
            const double *blk = a + (size_t)pos * BB;
#pragma unroll
            for (int e = 0; e < BB; e++) ac[e] += blk[e];
        }
        if (!ok) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            continue;
        }
        double cf[B];
#pragma unroll
        for (int q = 0; q < B; q++) cf[q] = 0.0;
        if (B == 2) {
            const double a11 = ac[0], a12 = ac[1], a21 = ac[2], a22 = ac[3];
            const double det = a11 * a22 - a12 * a21;
            if (fabs(det) < 1e-15) cf[0] = 0.0;
            else if (SECOND) cf[0] = -a12 / a22;
            else cf[0] = fabs(a22) > 1e-15 ? -a12 / a22 : 0.0;
        } else {
            if (B == 3) inv3(ac);
            else inv_nc<B>(ac);
            if (!SECOND) {
#pragma unroll
                for (int q = 0; q < B; q++) cf[q] = ac[q];
            } else if (B == 3) {
                cf[0] = ac[0];
                cf[1] = ac[1];
            } else {
#pragma unroll
                for (int q = 0; q < B - 1; q++) cf[q] = ac[q + 1];
            }
        }
#pragma unroll
        for (int q = 0; q < B; q++) coef[(size_t)c * B + q] = cf[q];
    }
}

__global__ void k_pressure_values(long long nnzb, int bb, int pos, const double *__restrict__ a,
                                  double *__restrict__ p_a)
{
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nnzb;
         k += (long long)gridDim.x * blockDim.x)
        p_a[k] = a[(size_t)k * bb + pos];
}

template <int B>
int launch_apply(int type, int nb, const int *ia, double *a, double *rhs, int thermal, const double *coef, int pidx,
                 double *p_a, hipStream_t st)
{
    const int grid = (nb + DC_ROWS - 1) / DC_ROWS;
    switch (type) {
    case JXAMD_DECOUP_ABF: k_decouple_apply<B, JXAMD_DECOUP_ABF><<<grid, DC_ROWS, 0, st>>>(nb, ia, a, rhs, thermal, coef, pidx, p_a); break;
    case JXAMD_DECOUP_ANL: k_decouple_apply<B, JXAMD_DECOUP_ANL><<<grid, DC_ROWS, 0, st>>>(nb, ia, a, rhs, thermal, coef, pidx, p_a); break;
    case JXAMD_DECOUP_SEM: k_decouple_apply<B, JXAMD_DECOUP_SEM><<<grid, DC_ROWS, 0, st>>>(nb, ia, a, rhs, thermal, coef, pidx, p_a); break;
    case JXAMD_DECOUP_QI: k_decouple_apply<B, JXAMD_DECOUP_QI><<<grid, DC_ROWS, 0, st>>>(nb, ia, a, rhs, thermal, coef, pidx, p_a); break;
    case JXAMD_DECOUP_TIMPES: k_decouple_apply<B, JXAMD_DECOUP_TIMPES><<<grid, DC_ROWS, 0, st>>>(nb, ia, a, rhs, thermal, coef, pidx, p_a); break;
    default: k_decouple_apply<B, JXAMD_DECOUP_TIMPES2><<<grid, DC_ROWS, 0, st>>>(nb, ia, a, rhs, thermal, coef, pidx, p_a); break;
    }
    return dc_chk(hipGetLastError(), "launch");
}

template <int B>
int launch_coef(int second, int nb, const int *ia, const int *ja, const double *a, double *coef, int *err,
                hipStream_t st)
{
    int grid = (nb + 255) / 256;
    if (grid > 8192) grid = 8192;
    if (second) k_timpes_coef<B, 1><<<grid, 256, 0, st>>>(nb, ia, ja, a, coef, err);
    else k_timpes_coef<B, 0><<<grid, 256, 0, st>>>(nb, ia, ja, a, coef, err);
    return dc_chk(hipGetLastError(), "launch");
}

} // namespace

extern "C" {

int jxamd_bsr_decouple_device(int nb, int b, const int *d_ia, const int *d_ja, double *d_a, double *d_rhs, int type,
                              int is_thermal, int pidx, double *d_p_a, void *stream)
{
    if (nb < 0 || b < 2 || b > JXAMD_MAX_BLOCK || type < JXAMD_DECOUP_NONE || type > JXAMD_DECOUP_TIMPES2)
        return JXAMD_ERROR_ARG;
    if (d_p_a && (pidx < 0 || pidx >= b)) return JXAMD_ERROR_ARG;
    if (nb > 0 && (!d_ia || !d_ja || !d_a)) return JXAMD_ERROR_ARG;
    if (nb == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    const int thermal = is_thermal ? 1 : 0;
    if (type == JXAMD_DECOUP_NONE) {
        if (!d_p_a) return 0;
        int nnzb = 0;
        if (hipMemcpyAsync(&nnzb, d_ia + nb, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return JXAMD_ERROR_DEVICE;
        return jxamd_bsr_extract_pressure_device(nb, b, nnzb, d_ia, d_ja, d_a, pidx, nullptr, nullptr, d_p_a, stream);
    }
    double *coef = nullptr;
    int *err = nullptr;
    int e = 0;
    if (type == JXAMD_DECOUP_TIMPES || type == JXAMD_DECOUP_TIMPES2) {
        // coefficients first (they need the column sums of the undecoupled matrix)
        if (hipMalloc(&coef, sizeof(double) * (size_t)nb * b + sizeof(int) * 64) != hipSuccess)
            return JXAMD_ERROR_MEMORY;
        err = (int *)(coef + (size_t)nb * b);
        if (hipMemsetAsync(err, 0, sizeof(int), st) != hipSuccess) e = JXAMD_ERROR_DEVICE;
        const int second = type == JXAMD_DECOUP_TIMPES2;
        if (!e) {
            switch (b) {
            case 2: e = launch_coef<2>(second, nb, d_ia, d_ja, d_a, coef, err, st); break;
            case 3: e = launch_coef<3>(second, nb, d_ia, d_ja, d_a, coef, err, st); break;
            default: e = launch_coef<4>(second, nb, d_ia, d_ja, d_a, coef, err, st); break;
            }
        }
        int herr = 0;
        if (!e && (hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
                   hipStreamSynchronize(st) != hipSuccess))
            e = JXAMD_ERROR_DEVICE;
        if (!e && herr) e = JXAMD_ERROR_ARG; // not structurally symmetric: nothing modified
    }
    if (!e) {
        switch (b) {
        case 2: e = launch_apply<2>(type, nb, d_ia, d_a, d_rhs, thermal, coef, pidx, d_p_a, st); break;
        case 3: e = launch_apply<3>(type, nb, d_ia, d_a, d_rhs, thermal, coef, pidx, d_p_a, st); break;
        default: e = launch_apply<4>(type, nb, d_ia, d_a, d_rhs, thermal, coef, pidx, d_p_a, st); break;
        }
    }
    const int es = dc_chk(hipStreamSynchronize(st), "synchronise");
    if (es && !e) e = es;
    if (coef && hipFree(coef) != hipSuccess && !e) e = JXAMD_ERROR_DEVICE;
    return e;
}

int jxamd_bsr_extract_pressure_device(int nb, int b, int nnzb, const int *d_ia, const int *d_ja, const double *d_a,
                                      int pidx, int *d_p_ia, int *d_p_ja, double *d_p_a, void *stream)
{
    if (nb < 0 || nnzb < 0 || b < 1 || b > JXAMD_MAX_BLOCK || pidx < 0 || pidx >= b || !d_p_a) return JXAMD_ERROR_ARG;
    if (nnzb > 0 && (!d_a || (d_p_ja && !d_ja))) return JXAMD_ERROR_ARG;
    if (d_p_ia && !d_ia) return JXAMD_ERROR_ARG;
    hipStream_t st = (hipStream_t)stream;
    if (d_p_ia && hipMemcpyAsync(d_p_ia, d_ia, sizeof(int) * ((size_t)nb + 1), hipMemcpyDeviceToDevice, st) != hipSuccess)
        return JXAMD_ERROR_DEVICE;
    if (d_p_ja && nnzb &&
        hipMemcpyAsync(d_p_ja, d_ja, sizeof(int) * (size_t)nnzb, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return JXAMD_ERROR_DEVICE;
    if (nnzb == 0) return 0;
    long long g = ((long long)nnzb + 255) / 256;
    if (g > 16384) g = 16384;
    k_pressure_values<<<(int)g, 256, 0, st>>>(nnzb, b * b, pidx * b + pidx, d_a, d_p_a);
    return dc_chk(hipGetLastError(), "launch");
}

} // extern "C"
