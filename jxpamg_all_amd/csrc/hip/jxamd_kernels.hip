// jxamd_kernels.hip -- CDNA4 (gfx950) kernels of the CPR-GMRES solve phase and
// their extern "C" launchers.  Host C code drives them through jxamd_dev.h.
//
// Device layouts (built on the host by jxamd_solver.c):
//   SELL-64: rows are grouped in slices of 64 (one wavefront); slice s owns
//   slot rows [sptr[s], sptr[s+1]); slot row k holds the k-th stored entry of
//   each of the 64 rows, lane-contiguous, so a wavefront's loads of one slot
//   are one 512-byte (fp64) transaction.  Short rows are padded with the
//   row's own column and a zero value.  Block (BSR) values keep b*b planes
//   per slot row: val[(slot*b*b + e)*64 + lane].
//   Smoother matrices are stored in sweep order: position p holds row perm[p];
//   positions are sorted by dependency level so a wavefront's rows become
//   ready together.  The diagonal is kept apart (diag[p]).
//
// Sync-free sweeps (k_gs_sweep, k_bhgs_sweep) run persistent wavefronts over
// level-pure slices with per-slice completion flags; every wait is bounded
// and an overflow is recorded in err[0] (see the comment at the kernels).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#define WAVE 64
#define SPIN_LIMIT (1 << 24)
// entries whose loads one lane issues before it consumes any of them in the
// level kernels (SELL slots, row-group strides): the gathers behind a column
// index are a second dependent load, and a lane with one chain in flight
// leaves HBM latency-bound
#define SELL_U 8
#define GRP_U 8
#define DOT_U 4

extern "C" {
#include "jxamd_dev.h"
}

namespace {

__device__ __forceinline__ int ld_flag(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Publishing a slice.  Every value another wave reads back in the same
// launch is written with a device-scope (agent) atomic store, which is
// written through to the device coherence point, and read with a device-scope
// atomic load, which is served from it; so no L2 write-back or invalidation
// is needed, only completion order: wait until this wave's stores are
// acknowledged (vmcnt 0 covers stores on gfx9-family CUs), then raise the
// flag.  The workgroup-scope fence keeps the compiler from moving memory
// operations across this point and emits no cache maintenance.
__device__ __forceinline__ void st_flag_release(int *p, int v)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_coherent(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coherent(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool flag_ready(int f, int epoch) { return (int)(f - epoch) >= 0; }
__device__ __forceinline__ float ld_coherent(const float *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coherent(float *p, float v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Level matrices of the V-cycle are stored as V (double, float, or __half
// with a power-of-two scale per stored row: a_ij = rsc[p] * h_ij); vectors
// and accumulation are X (double with double, float otherwise).
__device__ __forceinline__ double cvt(double v) { return v; }
__device__ __forceinline__ float cvt(float v) { return v; }
__device__ __forceinline__ float cvt(__half v) { return __half2float(v); }
template <typename V> struct row_scaled { static constexpr bool value = false; };
template <typename V> struct constexpr_is_half { static constexpr bool value = false; };
template <> struct constexpr_is_half<__half> { static constexpr bool value = true; };
// non-temporal load for data streamed once per launch (the fine BSR matrix;
// the AMG level matrices are re-read several times per V-cycle and gain from
// staying cached, so their kernels use ordinary loads)
template <typename T> __device__ __forceinline__ T ldnt(const T *p) { return __builtin_nontemporal_load(p); }
template <> __device__ __forceinline__ __half ldnt<__half>(const __half *p) { return *p; }
template <> struct row_scaled<__half> { static constexpr bool value = true; };
// LDS type for matrix blocks staged by the stage-2 kernel: values stored
// narrower than double are exact in float
template <typename V> struct lds_block { typedef float T; };
template <> struct lds_block<double> { typedef double T; };

// 16-byte vector of X
template <typename X> struct vec16;
template <> struct vec16<double> { typedef double __attribute__((ext_vector_type(2))) T; };
template <> struct vec16<float> { typedef float __attribute__((ext_vector_type(4))) T; };

// XCD-aware tile order for streaming kernels: consecutive tiles land on the
// same XCD (dispatch is round-robin over the 8 XCDs), so the x entries a
// stencil row shares with its neighbours stay in one L2.
__device__ __forceinline__ int xcd_tile(int bid, int nblocks)
{
    const int nx = 8;
    const int base = nblocks / nx, rem = nblocks % nx;
    const int xcd = bid % nx, k = bid / nx; // k-th workgroup dispatched to this XCD
    // XCD x owns a contiguous run of base (+1 for the first rem XCDs) tiles
    return xcd < rem ? xcd * (base + 1) + k : rem * (base + 1) + (xcd - rem) * base + k;
}

// ---------------------------------------------------------------------------
// BSR SpMV, SELL-64, one lane per block row.
//   mode 0: y = A x;  mode 1: y = rhs - A x;  mode 2: y = alpha*A x + beta*y
// ---------------------------------------------------------------------------
template <int B, int SU, bool UNIT>
__global__ void __launch_bounds__(256) k_bsr_spmv(int nb, const int *__restrict__ sptr, const int *__restrict__ col,
                                                  const double *__restrict__ val, const double *__restrict__ x,
                                                  const double *__restrict__ rhs, double *__restrict__ y, int mode,
                                                  double alpha, double beta)
{
    const int nblk = gridDim.x;
    const int tile = xcd_tile(blockIdx.x, nblk);
    const int s = tile * (blockDim.x / WAVE) + threadIdx.x / WAVE; // slice
    const int lane = threadIdx.x % WAVE;
    const int row = s * WAVE + lane;
    const int nsl = (nb + WAVE - 1) / WAVE;
    if (s >= nsl) return;
    const int k0 = sptr[s], k1 = sptr[s + 1];
    double acc[B];
    // UNIT: the identity diagonal block is not stored; the row starts from
    // its product, x_i, exactly what the first slot would have added
#pragma unroll
    for (int r = 0; r < B; r++) acc[r] = (UNIT && row < nb) ? x[(size_t)row * B + r] : 0.0;
    // matrix entries are streamed once: non-temporal loads keep them from
    // displacing the reused x entries in the caches; SU slots per batch
    for (int k = k0; k < k1; k += SU) {
        int c[SU];
#pragma unroll
        for (int u = 0; u < SU; u++)
            if (k + u < k1) c[u] = __builtin_nontemporal_load(col + (size_t)(k + u) * WAVE + lane);
        double xv[SU][B];
#pragma unroll
        for (int u = 0; u < SU; u++)
            if (k + u < k1)
#pragma unroll
                for (int t = 0; t < B; t++) xv[u][t] = x[(size_t)c[u] * B + t];
#pragma unroll
        for (int u = 0; u < SU; u++) {
            if (k + u >= k1) continue;
            const double *v = val + (size_t)(k + u) * B * B * WAVE + lane;
#pragma unroll
            for (int r = 0; r < B; r++) {
#pragma unroll
                for (int t = 0; t < B; t++) acc[r] += __builtin_nontemporal_load(v + (r * B + t) * WAVE) * xv[u][t];
            }
        }
    }
    if (row >= nb) return;
#pragma unroll
    for (int r = 0; r < B; r++) {
        const size_t o = (size_t)row * B + r;
        double out;
        if (mode == 0) out = acc[r];
        else if (mode == 1) out = rhs[o] - acc[r];
        else out = alpha * acc[r] + (beta == 0.0 ? 0.0 : beta * y[o]);
        y[o] = out;
    }
}

// Sum of val[e] * x[j(e) * xs] over slot rows [k0, k1) of a SELL layout,
// e = slot * 64 + lane, in slot order.  k1 is the same for the whole
// wavefront, so each batch of SELL_U slots is predicated uniformly: the lane
// issues all of a batch's index and value loads, then all of its gathers,
// before it adds (a lane with one load chain in flight at a time leaves HBM
// latency-bound).  Column indices: col[e] (PK false) or, packed, two int16
// offsets j - row per int word, word w0 + (slot - k0) / 2 of the lane (PK
// true: half the index bytes; the words are loaded whole, beyond k1 too, so
// the array carries a batch of padding words at its end).
template <typename V, typename X, bool PK>
__device__ __forceinline__ X sell_dot(int k0, int k1, int w0, int lane, int row, const int *__restrict__ col,
                                      const V *__restrict__ val, const X *__restrict__ x, int xs)
{
    X acc = 0;
    for (int k = k0; k < k1; k += SELL_U) {
        int j[SELL_U];
        V v[SELL_U];
        if (PK) {
            int wv[SELL_U / 2];
            const int wb = w0 + (k - k0) / 2;
#pragma unroll
            for (int u = 0; u < SELL_U / 2; u++) wv[u] = col[(size_t)(wb + u) * WAVE + lane];
#pragma unroll
            for (int u = 0; u < SELL_U; u += 2) {
                j[u] = row + (int)(short)(wv[u / 2] & 0xffff);
                j[u + 1] = row + (wv[u / 2] >> 16);
            }
            if constexpr (constexpr_is_half<V>::value) {
                // FP16 values paired like the offsets: whole words
                const __half2 *v2 = reinterpret_cast<const __half2 *>(val);
#pragma unroll
                for (int u = 0; u < SELL_U; u += 2) {
                    const __half2 h = v2[(size_t)(wb + u / 2) * WAVE + lane];
                    v[u] = __low2half(h);
                    v[u + 1] = __high2half(h);
                }
            } else {
#pragma unroll
                for (int u = 0; u < SELL_U; u++)
                    if (k + u < k1) v[u] = val[(size_t)(k + u) * WAVE + lane];
            }
        } else {
#pragma unroll
            for (int u = 0; u < SELL_U; u++) {
                if (k + u < k1) {
                    const size_t e = (size_t)(k + u) * WAVE + lane;
                    j[u] = col[e];
                    v[u] = val[e];
                }
            }
        }
        X xv[SELL_U];
#pragma unroll
        for (int u = 0; u < SELL_U; u++)
            if (k + u < k1) xv[u] = x[(size_t)j[u] * xs];
#pragma unroll
        for (int u = 0; u < SELL_U; u++)
            if (k + u < k1) acc += (X)cvt(v[u]) * xv[u];
    }
    return acc;
}

// Phase A of the chunked smoothers for one row i (chunk rows [ns, ns+len))
// held by one lane of a SELL-64 layout: over slot rows [k, k1) at lane offset
// lo, entries of in-chunk columns other than i go to the row's LDS block
// Brow, and the others fold into acc -= a_ij uo_j.  Batches of U slots,
// predicated per lane (a block's rows may lie in two slices).
template <int U, typename V, typename X, bool PK>
__device__ __forceinline__ void chunk_row_gather(int k0, int k1, int w0, size_t lo, int i, int ns, int len, X sc,
                                                 const int *__restrict__ col, const V *__restrict__ val,
                                                 const X *__restrict__ uo, X *Brow, X &acc)
{
    for (int k = k0; k < k1; k += U) {
        int j[U];
        V v[U];
        if (PK) {
            // packed int16 offsets, as in sell_dot (whole words, padded array)
            int wv[U / 2];
            const int wb = w0 + (k - k0) / 2;
#pragma unroll
            for (int u = 0; u < U / 2; u++) wv[u] = col[(size_t)(wb + u) * WAVE + lo];
#pragma unroll
            for (int u = 0; u < U; u += 2) {
                j[u] = i + (int)(short)(wv[u / 2] & 0xffff);
                j[u + 1] = i + (wv[u / 2] >> 16);
            }
            if constexpr (constexpr_is_half<V>::value) {
                // FP16 values paired like the offsets: whole words
                const __half2 *v2 = reinterpret_cast<const __half2 *>(val);
#pragma unroll
                for (int u = 0; u < U; u += 2) {
                    const __half2 h = v2[(size_t)(wb + u / 2) * WAVE + lo];
                    v[u] = __low2half(h);
                    v[u + 1] = __high2half(h);
                }
            } else {
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (k + u < k1) v[u] = val[(size_t)(k + u) * WAVE + lo];
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (k + u < k1) {
                    const size_t e = (size_t)(k + u) * WAVE + lo;
                    j[u] = col[e];
                    v[u] = val[e];
                }
            }
        }
        X xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const bool out = k + u < k1 && !(j[u] >= ns && j[u] < ns + len);
            xv[u] = (out && uo) ? uo[j[u]] : (X)0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (k + u >= k1) continue;
            const X vv = sc * (X)cvt(v[u]);
            if (j[u] >= ns && j[u] < ns + len) {
                if (j[u] != i) Brow[j[u] - ns] += vv;
            } else if (uo) {
                acc -= vv * xv[u];
            }
        }
    }
}

// The same for one CSR row shared by a group of G lanes (LDS adds are
// atomic): entries k, k+G, ... below ke, U of them in flight per batch.
template <int U, int G, typename V, typename X>
__device__ __forceinline__ void csr_row_gather(int k, int ke, int i, int ns, int len, X sc, const int *__restrict__ ja,
                                               const V *__restrict__ val, const X *__restrict__ uo, X *Brow, X &acc)
{
    for (; k < ke; k += U * G) {
        int j[U];
        V v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int kk = k + u * G;
            j[u] = kk < ke ? ja[kk] : -1;
            v[u] = kk < ke ? val[kk] : (V)0;
        }
        X xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const bool out = j[u] >= 0 && !(j[u] >= ns && j[u] < ns + len);
            xv[u] = (out && uo) ? uo[j[u]] : (X)0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (j[u] < 0) continue;
            const X vv = sc * (X)cvt(v[u]);
            if (j[u] >= ns && j[u] < ns + len) {
                if (j[u] != i) atomicAdd(&Brow[j[u] - ns], vv);
            } else if (uo) {
                acc -= vv * xv[u];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Scalar CSR SpMV, SELL-64 (levels: A_l residual, R_l restriction, P_l
// prolongation).  mode 0: y = A x; 1: y = rhs - A x; 2: y += A x.
// x and rhs may be strided views (xs, rs) so the pressure block of a BSR
// vector is read in place.
// ---------------------------------------------------------------------------
// first polynomial step from a zero guess on one row whose right-hand side
// f is already in a register: u = 0 + w D^{-1} (f - 0)
template <typename X>
__device__ __forceinline__ void poly_first_row(int row, const X *__restrict__ dinv, X f, X *__restrict__ u, X w)
{
    const X r = f - (X)0;
    u[row] = (X)0 + w * dinv[row] * r;
}

template <typename V, typename X, bool PK>
__global__ void __launch_bounds__(256) k_csr_spmv(int n, const int *__restrict__ sptr, const int *__restrict__ col,
                                                  const V *__restrict__ val, const float *__restrict__ rsc,
                                                  const X *__restrict__ x, int xs, const X *__restrict__ rhs, int rs,
                                                  X *__restrict__ y, int mode, const X *__restrict__ d2,
                                                  X *__restrict__ u2, X w2)
{
    const int s = xcd_tile(blockIdx.x, gridDim.x) * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int lane = threadIdx.x % WAVE;
    const int nsl = (n + WAVE - 1) / WAVE;
    if (s >= nsl) return;
    const int row = s * WAVE + lane;
    X acc = sell_dot<V, X, PK>(sptr[s], sptr[s + 1], PK ? sptr[nsl + 1 + s] : 0, lane, row, col, val, x, xs);
    if (row >= n) return;
    if (row_scaled<V>::value) acc *= (X)rsc[row];
    if (mode == 0) y[row] = acc;
    else if (mode == 1) y[row] = rhs[(size_t)row * rs] - acc;
    else y[row] += acc;
    if (u2) poly_first_row(row, d2, acc, u2, w2); // mode 0 only (launcher checks)
}

// One step of the polynomial smoother on a level (SELL-64, full rows):
//   un_i = uo_i + omega * dinv_i * (f_i - sum_j a_ij uo_j)
// uo == nullptr stands for zeros.  dinv_i = 0 keeps a row at its old value.
template <typename V, typename X, bool PK>
__global__ void __launch_bounds__(256) k_poly_step(int n, const int *__restrict__ sptr, const int *__restrict__ col,
                                                   const V *__restrict__ val, const float *__restrict__ rsc,
                                                   const X *__restrict__ dinv, const X *__restrict__ f, int fs,
                                                   const X *__restrict__ uo, X *__restrict__ un, X omega)
{
    const int s = xcd_tile(blockIdx.x, gridDim.x) * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int lane = threadIdx.x % WAVE;
    const int nsl = (n + WAVE - 1) / WAVE;
    if (s >= nsl) return;
    const int row = s * WAVE + lane;
    X acc = uo ? sell_dot<V, X, PK>(sptr[s], sptr[s + 1], PK ? sptr[nsl + 1 + s] : 0, lane, row, col, val, uo, 1)
               : (X)0;
    if (row >= n) return;
    if (row_scaled<V>::value) acc *= (X)rsc[row];
    const X r = f[(size_t)row * fs] - acc;
    un[row] = (uo ? uo[row] : (X)0) + omega * dinv[row] * r;
}

// ---------------------------------------------------------------------------
// Sync-free sweeps over level-pure wavefront slices.
//
// The host sorts rows by dependency level and pads every level to a multiple
// of 64 positions, so a slice (one wavefront's 64 positions, perm[p] < 0 for
// padding) never holds a row and one of its dependencies.  dep_ptr/dep_list
// give, per slice, the earlier slices holding its rows' dependencies.  Waves
// are persistent: each claims slices in order from the ticket counter (so a
// claimed slice's dependencies belong to slices claimed before it, all held
// by running waves), loads its row data, waits on its dependency slices'
// flags (bounded), takes an acquire fence, computes, stores with device-
// coherent stores and publishes sflag[slice] = epoch with release semantics.
// Every wave ends with exactly one failed claim, so a launch consumes
// nsl + (number of waves) tickets.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int claim_slice(unsigned long long *ticket, unsigned long long base, int lane)
{
    int s = 0;
    if (lane == 0) s = (int)(__hip_atomic_fetch_add(ticket, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - base);
    return __shfl(s, 0);
}

__device__ __forceinline__ void wait_slices(const int *__restrict__ dep_ptr, const int *__restrict__ dep_list,
                                            const int *sflag, int s, int epoch, int lane, int *err)
{
    const int d0 = dep_ptr[s], d1 = dep_ptr[s + 1];
    for (int q = d0 + lane; q < d1; q += WAVE) {
        const int w = dep_list[q];
        int spins = 0, nap = 0;
        while (!flag_ready(ld_flag(sflag + w), epoch)) {
            if (++spins > SPIN_LIMIT) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            if (nap < 4) { __builtin_amdgcn_s_sleep(1); nap++; }
            else __builtin_amdgcn_s_sleep(4);
        }
    }
    // The flag loads have returned (the loop waited on them); the dependent
    // data is read afterwards with device-scope atomic loads (see
    // st_flag_release), so only a compiler barrier is needed here.
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// One Gauss-Seidel sweep of hybrid SSOR (relax 6, weights 1).
//   dir > 0: forward (dependencies j < i), dir < 0: backward (j > i).
//   Old values from uo (zeros if uo == nullptr), new values into un.
//   f may be strided (fs).  A row with a zero diagonal keeps its old value.
template <typename V, typename X>
__global__ void __launch_bounds__(256) k_gs_sweep(int nsl, const int *__restrict__ perm, const int *__restrict__ sptr,
                                                  const int *__restrict__ col, const V *__restrict__ val,
                                                  const float *__restrict__ rsc, const X *__restrict__ diag,
                                                  const X *__restrict__ f, int fs, const X *uo, X *un,
                                                  const int *__restrict__ dep_ptr, const int *__restrict__ dep_list,
                                                  int *sflag, int epoch, int dir, unsigned long long *ticket,
                                                  unsigned long long base, int *err)
{
    const int lane = threadIdx.x % WAVE;
    for (;;) {
        const int s = claim_slice(ticket, base, lane);
        if (s < 0 || s >= nsl) break;
        const int p = s * WAVE + lane;
        const int i = perm[p];
        const int k0 = sptr[s], k1 = sptr[s + 1];
        X res = 0, d = 1, uold_i = 0, sc = 1;
        if (i >= 0) {
            d = diag[p];
            res = f[(size_t)i * fs];
            if (uo) uold_i = uo[i];
            if (row_scaled<V>::value) sc = (X)rsc[p];
            // terms on old values need no wait
            for (int k = k0; k < k1; k++) {
                const size_t e = (size_t)k * WAVE + lane;
                const int j = col[e];
                const bool dep = dir > 0 ? (j < i) : (j > i);
                if (!dep && j != i && uo) res -= sc * (X)cvt(val[e]) * uo[j];
            }
        }
        wait_slices(dep_ptr, dep_list, sflag, s, epoch, lane, err);
        if (i >= 0) {
            for (int k = k0; k < k1; k++) {
                const size_t e = (size_t)k * WAVE + lane;
                const int j = col[e];
                const bool dep = dir > 0 ? (j < i) : (j > i);
                if (dep) res -= sc * (X)cvt(val[e]) * ld_coherent(un + j);
            }
            st_coherent(un + i, d != (X)0 ? res / d : uold_i);
        }
        st_flag_release(sflag + s, epoch);
    }
}

// K forward block hybrid SOR sweeps on the BSR matrix in one launch
// (par_relax_bsr.c semantics, one rank), sweep k from x^(k-1) to x^(k):
//   tmp = omega*b_i - sum_{j<i} A_ij x^(k)_j - sum_{j>i} A_ij x^(k-1)_j
//         + (1-omega) sum_{j!=i} A_ij x^(k-1)_j
//   x^(k)_i = (1 - w*omega) x^(k-1)_i + w * Dinv_i tmp
// Every x^(k) lives in slice order: the B values of the row at position p
// (perm[p] = i) sit at xs + k*xstride + p*B; x^(-1) is xsl0 in the same
// order.  A slice's 64 rows fill exactly 12 (B = 3) or 8/16 (B = 2/4) whole
// 128-byte lines, so a line is read only after the flag of the one slice that
// writes it: ordinary cached loads then see final values, and only the
// stores (written through at device scope) and the flags need coherence.
// Columns of the layout are positions, bit-inverted (~pos) for lower
// neighbours (j < i); the final sweep also stores x in row order into xout.
// Work items are virtual slices v = k*nsl + s taken in the order vorder (the
// identity if nullptr), a topological order of
//   (k, s) after (k, s') for s' in dlo(s)   (slices of lower neighbours)
//   (k, s) after (k-1, s'') for s'' in dall(s) (slices of all neighbours, s)
// so sweep k+1 starts on a slice as soon as sweep k has passed its
// neighbourhood: the K sweeps overlap instead of running back to back.
// Flags: sflag[k*nsl + s].
template <int B>
__global__ void __launch_bounds__(256)
k_bhgs_multi(int nsl, int nvs, const int *__restrict__ vorder, const int *__restrict__ perm,
             const int *__restrict__ sptr, const int *__restrict__ col, const double *__restrict__ val,
             const double *__restrict__ dinv, const double *__restrict__ rhs, const double *__restrict__ xsl0,
             double *xs, long long xstride, double *__restrict__ xout, const int *__restrict__ dlo_ptr,
             const int *__restrict__ dlo_list, const int *__restrict__ dall_ptr, const int *__restrict__ dall_list,
             int *sflag, int epoch, double w, double omega, unsigned long long *ticket, unsigned long long base,
             int *err)
{
    const int lane = threadIdx.x % WAVE;
    const int K = nvs / nsl;
    for (;;) {
        const int t = claim_slice(ticket, base, lane);
        if (t < 0 || t >= nvs) break;
        const int v = vorder ? vorder[t] : t;
        const int k = v / nsl, s = v - k * nsl;
        double *xn = xs + (size_t)k * xstride;
        const double *xo = k == 0 ? xsl0 : xs + (size_t)(k - 1) * xstride;
        if (k > 0) wait_slices(dall_ptr, dall_list, sflag + (size_t)(k - 1) * nsl, s, epoch, lane, err);
        const int p = s * WAVE + lane;
        const int i = perm[p];
        const int k0 = sptr[s], k1 = sptr[s + 1];
        double res0[B], res2[B];
#pragma unroll
        for (int r = 0; r < B; r++) { res0[r] = 0.0; res2[r] = 0.0; }
        if (i >= 0) {
            // all old-value terms, and the current-value terms of upper neighbours
            for (int kk = k0; kk < k1; kk++) {
                const int c = col[(size_t)kk * WAVE + lane];
                const bool lower = c < 0;
                const int jp = lower ? ~c : c;
                if (jp == p) continue; // padding slot
                double xold[B];
#pragma unroll
                for (int q = 0; q < B; q++) xold[q] = xo[(size_t)jp * B + q];
                const double *vv = val + (size_t)kk * B * B * WAVE + lane;
#pragma unroll
                for (int r = 0; r < B; r++) {
#pragma unroll
                    for (int q = 0; q < B; q++) {
                        const double a = vv[(r * B + q) * WAVE];
                        res2[r] += a * xold[q];
                        if (!lower) res0[r] -= a * xold[q];
                    }
                }
            }
        }
        double xi[B], rb[B], D[B * B];
        if (i >= 0) {
#pragma unroll
            for (int r = 0; r < B; r++) {
                xi[r] = xo[(size_t)p * B + r];
                rb[r] = rhs[(size_t)i * B + r];
            }
#pragma unroll
            for (int q = 0; q < B * B; q++) D[q] = dinv[(size_t)i * B * B + q];
        }
        wait_slices(dlo_ptr, dlo_list, sflag + (size_t)k * nsl, s, epoch, lane, err);
        if (i >= 0) {
            for (int kk = k0; kk < k1; kk++) {
                const int c = col[(size_t)kk * WAVE + lane];
                if (c >= 0) continue;
                const int jp = ~c;
                double xc[B];
#pragma unroll
                for (int q = 0; q < B; q++) xc[q] = xn[(size_t)jp * B + q];
                const double *vv = val + (size_t)kk * B * B * WAVE + lane;
#pragma unroll
                for (int r = 0; r < B; r++) {
#pragma unroll
                    for (int q = 0; q < B; q++) res0[r] -= vv[(r * B + q) * WAVE] * xc[q];
                }
            }
            double tmp[B];
#pragma unroll
            for (int r = 0; r < B; r++) tmp[r] = omega * rb[r] + res0[r] + (1.0 - omega) * res2[r];
            const double prod = 1.0 - w * omega;
#pragma unroll
            for (int r = 0; r < B; r++) {
                double o = 0.0;
#pragma unroll
                for (int q = 0; q < B; q++) o += D[r * B + q] * tmp[q];
                const double xv = xi[r] * prod + w * o;
                st_coherent(xn + (size_t)p * B + r, xv);
                if (k == K - 1) xout[(size_t)i * B + r] = xv;
            }
        }
        st_flag_release(sflag + (size_t)k * nsl + s, epoch);
    }
}

// One forward block hybrid SOR sweep as run on nparts ranks (contiguous
// block-row partitions of at most LMAX rows, first nb % nparts one row
// longer; par_relax_bsr.c:296-330 and :523-557 with the other ranks'
// neighbours at entry values): per row i of partition P
//   res  = b_i - sum_{j not in P} A_ij x_old_j
//   res0 = -sum_{j in P, j<i} A_ij x_new_j - sum_{j in P, j>i} A_ij x_old_j
//   res2 = sum_{j in P, j!=i} A_ij x_old_j
//   x_i  = (1 - w omega) x_old_i + w Dinv_i (omega res + res0 + (1-omega) res2)
// Everything but the x_new terms is row-parallel (phase A, one lane per row
// over the SpMV layout: SELL-64, all blocks, natural order); the in-partition
// lower blocks go to LDS and one lane per partition runs the short forward
// recurrence (phase B).  x_old: xo (B values per row) or, when xo ==
// nullptr, the prolongated pressure xp (slot pidx, zeros elsewhere; XP
// double, or float for a stage-1 result in FP32, widened as it is read).
// Blocks are stored as V (double, float, or __half scaled per block row:
// A_ij = rsc[i] * H_ij), inverted diagonal blocks as DV (dinv == nullptr:
// every diagonal block is the identity, as after ABF decoupling, and none
// is read); arithmetic in double.  PU slots per batch.
template <int B, int LMAX, typename V, typename DV, int PU, typename XP>
__global__ void __launch_bounds__(256) k_bhgs_part(int nb, int nparts, int cpb, const int *__restrict__ sptr,
                                                   const int *__restrict__ pptr,
                                                   const int *__restrict__ col, const V *__restrict__ val,
                                                   const float *__restrict__ rsc, const DV *__restrict__ dinv,
                                                   const double *__restrict__ rhs,
                                                   const double *__restrict__ xo, const XP *__restrict__ xp,
                                                   int pidx, double *__restrict__ xn, double w, double omega)
{
    constexpr int NL = LMAX * (LMAX - 1) / 2; // strictly lower pairs of a partition
    constexpr int LST = NL * B * B + 1;       // odd stride between partitions
    // in-partition lower blocks in LDS as LT: float for narrow block storage
    // (exact), which halves the footprint and doubles the resident workgroups
    typedef typename lds_block<V>::T LT;
    extern __shared__ __align__(16) unsigned char smraw[];
    double *sm = reinterpret_cast<double *>(smraw);
    double *h = sm, *xold = sm + 256 * B;
    LT *Lb = reinterpret_cast<LT *>(sm + 2 * 256 * B);
    const int size = nb / nparts, rest = nb - size * nparts;
    const int c0 = xcd_tile(blockIdx.x, gridDim.x) * cpb;
    const int c1 = min(c0 + cpb, nparts);
    const int row0 = c0 < rest ? c0 * (size + 1) : c0 * size + rest;
    const int row1 = c1 < rest ? c1 * (size + 1) : c1 * size + rest;
    // pairs without a coupling stay zero
    for (int e = threadIdx.x; e < cpb * LST; e += blockDim.x) Lb[e] = (LT)0;
    __syncthreads();
    const int r = threadIdx.x;
    const int i = row0 + r;
    if (i < row1) {
        const int big = rest * (size + 1);
        const int c = i < big ? i / (size + 1) : rest + (i - big) / size;
        const int ns = c < rest ? c * (size + 1) : c * size + rest;
        const int len = c < rest ? size + 1 : size;
        const int ti = i - ns;
        double res[B], res0[B], res2[B];
#pragma unroll
        for (int q = 0; q < B; q++) {
            res[q] = rhs[(size_t)i * B + q];
            res0[q] = 0.0;
            res2[q] = 0.0;
        }
        const int s = i / WAVE, lane = i % WAVE;
        const double sc = row_scaled<V>::value ? (double)rsc[i] : 1.0;
        const int ks = sptr[s], k1 = sptr[s + 1];
        for (int k = ks; k < k1; k += PU) {
            // a batch of PU slots: indices, then the neighbours' x, then the
            // blocks (their loads do not wait on the gathers)
            int jj[PU];
            double xj[PU][B];
            // FP16 blocks: two slots' values per 32-bit word (slot pairs
            // from the slice's first slot, pair base pptr[s]), loaded whole
            float hv[constexpr_is_half<V>::value ? PU : 1][B * B];
            if constexpr (constexpr_is_half<V>::value) {
                const __half2 *v2 = reinterpret_cast<const __half2 *>(val);
#pragma unroll
                for (int u = 0; u < PU; u += 2) {
                    if (k + u < k1) {
                        const __half2 *pv = v2 + ((size_t)(pptr[s] + (k - ks + u) / 2) * B * B) * WAVE + lane;
#pragma unroll
                        for (int e = 0; e < B * B; e++) {
                            const __half2 h = pv[e * WAVE];
                            hv[u][e] = __low2float(h);
                            hv[u + 1][e] = __high2float(h);
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < PU; u++) jj[u] = k + u < k1 ? col[(size_t)(k + u) * WAVE + lane] : i;
#pragma unroll
            for (int u = 0; u < PU; u++) {
                const int j = jj[u];
#pragma unroll
                for (int q = 0; q < B; q++)
                    xj[u][q] = j == i ? 0.0 : (xo ? xo[(size_t)j * B + q] : (q == pidx ? (double)xp[j] : 0.0));
            }
#pragma unroll
            for (int u = 0; u < PU; u++) {
                const int j = jj[u];
                if (j == i) continue; // the diagonal block, padding, or past the slice
                const V *vv = val + (size_t)(k + u) * B * B * WAVE + lane;
                auto blk = [&](int e) -> double {
                    if constexpr (constexpr_is_half<V>::value) return (double)hv[u][e];
                    else return (double)cvt(ldnt(vv + e * WAVE));
                };
                const bool in = j >= ns && j < ns + len;
                if (in && j < i) {
                    LT *L = Lb + (size_t)(c - c0) * LST + ((ti * (ti - 1)) / 2 + (j - ns)) * B * B;
#pragma unroll
                    for (int e = 0; e < B * B; e++) L[e] = (LT)(sc * blk(e));
                }
#pragma unroll
                for (int rr = 0; rr < B; rr++) {
                    double acc = 0.0;
#pragma unroll
                    for (int q = 0; q < B; q++) acc += blk(rr * B + q) * xj[u][q];
                    acc *= sc;
                    if (!in) res[rr] -= acc;
                    else {
                        res2[rr] += acc;
                        if (j > i) res0[rr] -= acc;
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < B; q++) {
            h[r * B + q] = omega * res[q] + res0[q] + (1.0 - omega) * res2[q];
            xold[r * B + q] = xo ? xo[(size_t)i * B + q] : (q == pidx ? (double)xp[i] : 0.0);
        }
    }
    __syncthreads();
    const int c = c0 + (int)threadIdx.x;
    if (c < c1) {
        const int ns = c < rest ? c * (size + 1) : c * size + rest;
        const int len = c < rest ? size + 1 : size;
        const int rl = ns - row0;
        const LT *Lc = Lb + (size_t)(c - c0) * LST;
        const double prod = 1.0 - w * omega;
        double xnew[LMAX][B];
#pragma unroll
        for (int t = 0; t < LMAX; t++) {
            if (t < len) {
                double tmp[B];
#pragma unroll
                for (int q = 0; q < B; q++) tmp[q] = h[(rl + t) * B + q];
#pragma unroll
                for (int u = 0; u < LMAX - 1; u++) {
                    if (u < t) {
                        const LT *L = Lc + ((t * (t - 1)) / 2 + u) * B * B;
#pragma unroll
                        for (int rr = 0; rr < B; rr++)
#pragma unroll
                            for (int q = 0; q < B; q++) tmp[rr] -= (double)L[rr * B + q] * xnew[u][q];
                    }
                }
                const DV *D = dinv ? dinv + (size_t)(ns + t) * B * B : nullptr;
#pragma unroll
                for (int rr = 0; rr < B; rr++) {
                    double o = 0.0;
                    if (D) {
#pragma unroll
                        for (int q = 0; q < B; q++) o += (double)D[rr * B + q] * tmp[q];
                    } else {
                        o = tmp[rr]; // identity diagonal blocks (ABF-decoupled rows)
                    }
                    xnew[t][rr] = prod * xold[(rl + t) * B + rr] + w * o;
                }
            } else {
#pragma unroll
                for (int q = 0; q < B; q++) xnew[t][q] = 0.0;
            }
        }
#pragma unroll
        for (int t = 0; t < LMAX; t++)
            if (t < len)
#pragma unroll
                for (int q = 0; q < B; q++) h[(rl + t) * B + q] = xnew[t][q];
    }
    __syncthreads();
    if (i < row1)
#pragma unroll
        for (int q = 0; q < B; q++) xn[(size_t)i * B + q] = h[r * B + q];
}

// x^(-1) of the stage-2 sweeps in slice order: out[p*B + q] = x0[i*B + q],
// or, if x0 == nullptr, the prolongated pressure (q == pidx ? xp[i] : 0),
// i = perm[p]; padding positions get zeros
__global__ void k_bhgs_gather(int npos, int B, const int *__restrict__ perm, const double *__restrict__ x0,
                              const double *__restrict__ xp, int pidx, double *__restrict__ out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= npos * B) return;
    const int p = e / B, q = e - p * B;
    const int i = perm[p];
    double v = 0.0;
    if (i >= 0) v = x0 ? x0[(size_t)i * B + q] : (q == pidx ? xp[i] : 0.0);
    out[e] = v;
}

// One block triangular solve of BILU(0), sync-free over level-pure slices:
//   out_i = Dinv_i (rhs_i - sum_j M_ij out_j)   (Dinv_i = I when dinv == nullptr)
// where the slice layout holds only the blocks M_ij of one triangle (the
// dependencies: lower blocks for the forward L solve, upper blocks for the
// backward U solve), stored as V (double, float, or __half scaled by the
// power of two rsc[p] per block row: M_ij = rsc[p] * stored; FP16 elements
// stored in pairs, slot k's word w at (k * ceil(B*B/2) + w) * 64 + lane), Dinv as DV,
// vectors and arithmetic double.  A lane loads the column indices of BT_U
// blocks, then their values and the dependency values, before it uses any,
// so one level step costs two dependent round trips rather than two per
// block.  The subtractions keep the order of the blocks in the row.
#define BT_U 4
template <int B, typename V, typename DV>
__global__ void __launch_bounds__(256) k_btri(int nsl, const int *__restrict__ perm, const int *__restrict__ sptr,
                                              const int *__restrict__ col, const V *__restrict__ val,
                                              const float *__restrict__ rsc, const DV *__restrict__ dinv,
                                              const double *__restrict__ rhs, double *out,
                                              const int *__restrict__ dep_ptr, const int *__restrict__ dep_list,
                                              int *sflag, int epoch, unsigned long long *ticket,
                                              unsigned long long base, int *err)
{
    const int lane = threadIdx.x % WAVE;
    for (;;) {
        const int s = claim_slice(ticket, base, lane);
        if (s < 0 || s >= nsl) break;
        const int p = s * WAVE + lane;
        const int i = perm[p];
        const int k0 = sptr[s], k1 = sptr[s + 1];
        double acc[B], D[B * B], sc = 1.0;
        if (i >= 0) {
#pragma unroll
            for (int r = 0; r < B; r++) acc[r] = rhs[(size_t)i * B + r];
            if (dinv) {
#pragma unroll
                for (int q = 0; q < B * B; q++) D[q] = (double)dinv[(size_t)i * B * B + q];
            }
            if (rsc) sc = (double)rsc[p];
        }
        wait_slices(dep_ptr, dep_list, sflag, s, epoch, lane, err);
        if (i >= 0) {
            for (int k = k0; k < k1; k += BT_U) {
                int j[BT_U];
#pragma unroll
                for (int u = 0; u < BT_U; u++) j[u] = k + u < k1 ? col[(size_t)(k + u) * WAVE + lane] : i;
                double xc[BT_U][B];
                double m[BT_U][B * B];
#pragma unroll
                for (int u = 0; u < BT_U; u++) {
                    // the row itself marks a padding slot (and the batch's tail)
#pragma unroll
                    for (int q = 0; q < B; q++) xc[u][q] = j[u] != i ? ld_coherent(out + (size_t)j[u] * B + q) : 0.0;
                    if constexpr (constexpr_is_half<V>::value) {
                        // FP16: elements 2w and 2w + 1 of a block in one
                        // 32-bit word (a zero high half past the last), so
                        // the loads are whole words
                        constexpr int NW = (B * B + 1) / 2;
                        const __half2 *vw = reinterpret_cast<const __half2 *>(val) + (size_t)(k + u) * NW * WAVE + lane;
#pragma unroll
                        for (int w = 0; w < NW; w++) {
                            const __half2 h = j[u] != i ? vw[w * WAVE] : __half2{};
                            m[u][2 * w] = (double)__low2float(h);
                            if (2 * w + 1 < B * B) m[u][2 * w + 1] = (double)__high2float(h);
                        }
                    } else {
                        const V *vv = val + (size_t)(k + u) * B * B * WAVE + lane;
#pragma unroll
                        for (int q = 0; q < B * B; q++) m[u][q] = j[u] != i ? (double)cvt(vv[q * WAVE]) : 0.0;
                    }
                }
#pragma unroll
                for (int u = 0; u < BT_U; u++) {
                    if (j[u] == i) continue;
#pragma unroll
                    for (int r = 0; r < B; r++) {
#pragma unroll
                        for (int q = 0; q < B; q++) acc[r] -= (sc * m[u][r * B + q]) * xc[u][q];
                    }
                }
            }
            if (dinv) {
#pragma unroll
                for (int r = 0; r < B; r++) {
                    double o = 0.0;
#pragma unroll
                    for (int q = 0; q < B; q++) o += D[r * B + q] * acc[q];
                    st_coherent(out + (size_t)i * B + r, o);
                }
            } else {
#pragma unroll
                for (int r = 0; r < B; r++) st_coherent(out + (size_t)i * B + r, acc[r]);
            }
        }
        st_flag_release(sflag + s, epoch);
    }
}

// ---------------------------------------------------------------------------
// Hybrid SSOR with T threads (relax 6, weights 1; par_relax_6.c:132-223):
// rows split into T contiguous chunks (first n % T chunks one row longer);
// one lane per chunk runs a forward then a backward Gauss-Seidel sweep over
// its chunk, reading neighbours outside the chunk from the values at entry.
// Entry values come from uo (zeros if uo == nullptr); results go to un.
// Inside the forward sweep a neighbour j of row i in the same chunk is new
// (un) if j < i and still at its entry value (uo) if j > i; in the backward
// sweep every in-chunk neighbour has been written to un.
// Layout: wave w owns slices [wstep[w], wstep[w+1]); slice wstep[w] + t holds
// the t-th row of each of the wave's 64 chunks (SELL-64, off-diagonals, the
// diagonal in diag[p]).  No synchronisation between lanes.
// ---------------------------------------------------------------------------
template <typename V, typename X>
__global__ void __launch_bounds__(256) k_gs_chunk(int n, int nchunks, const int *__restrict__ wstep,
                                                  const int *__restrict__ sptr, const int *__restrict__ col,
                                                  const V *__restrict__ val, const float *__restrict__ rsc,
                                                  const X *__restrict__ diag, const X *__restrict__ f, int fs,
                                                  const X *__restrict__ uo, X *__restrict__ un)
{
    const int w = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int lane = threadIdx.x % WAVE;
    const int c = w * WAVE + lane;
    if (w * WAVE >= nchunks) return;
    const int size = n / nchunks, rest = n - size * nchunks;
    int ns = 0, ne = 0;
    if (c < nchunks) {
        ns = c < rest ? c * size + c : c * size + rest;
        ne = ns + size + (c < rest ? 1 : 0);
    }
    const int len = ne - ns;
    const int s0 = wstep[w], nsteps = wstep[w + 1] - s0;
    // forward sweep
    for (int t = 0; t < nsteps; t++) {
        if (t >= len) continue;
        const int i = ns + t;
        const int s = s0 + t;
        const size_t p = (size_t)s * WAVE + lane;
        const X d = diag[p];
        const X sc = row_scaled<V>::value ? (X)rsc[p] : (X)1;
        X res = f[(size_t)i * fs];
        for (int k = sptr[s]; k < sptr[s + 1]; k++) {
            const size_t e = (size_t)k * WAVE + lane;
            const int j = col[e];
            if (j == i) continue; // padding slot
            X xj;
            if (j >= ns && j < i) xj = un[j];
            else xj = uo ? uo[j] : (X)0;
            res -= sc * (X)cvt(val[e]) * xj;
        }
        un[i] = d != (X)0 ? res / d : (uo ? uo[i] : (X)0);
    }
    // backward sweep
    for (int t = nsteps - 1; t >= 0; t--) {
        if (t >= len) continue;
        const int i = ns + t;
        const int s = s0 + t;
        const size_t p = (size_t)s * WAVE + lane;
        const X d = diag[p];
        if (d == (X)0) continue;
        const X sc = row_scaled<V>::value ? (X)rsc[p] : (X)1;
        X res = f[(size_t)i * fs];
        for (int k = sptr[s]; k < sptr[s + 1]; k++) {
            const size_t e = (size_t)k * WAVE + lane;
            const int j = col[e];
            if (j == i) continue;
            X xj;
            if (j >= ns && j < ne) xj = un[j];
            else xj = uo ? uo[j] : (X)0;
            res -= sc * (X)cvt(val[e]) * xj;
        }
        un[i] = res / d;
    }
}

// ---------------------------------------------------------------------------
// The same hybrid SSOR for short chunks (at most LMAX rows), as one pass over
// the matrix.  Every coupling to a row outside the chunk is read at its entry
// value in both the forward and the backward sweep, so it folds into the
// right-hand side: g_i = f_i - sum_{j outside} a_ij uo_j.  What is left per
// chunk is a len x len block B of in-chunk couplings, and both sweeps are
// Gauss-Seidel on (D + B) u = g starting from the entry values (a later row
// of the chunk is still at its entry value when the forward sweep reads it).
//   phase A  one lane per row (SELL-64 rows in natural order, off-diagonals,
//            diagonal in diag[i]): g_i, d_i, entry value and the row of B
//            into LDS
//   phase B  one lane per chunk: forward then backward sweep, u in registers
//   phase C  one lane per row: un_i from LDS (coalesced store)
// Block b owns chunks [b*cpb, (b+1)*cpb), at most 256 rows.  Rows with a zero
// diagonal keep their entry value.
// ---------------------------------------------------------------------------
template <int LMAX, typename V, typename X, bool PK>
__global__ void __launch_bounds__(256) k_gs_chunk_lds(int n, int nchunks, int cpb, const int *__restrict__ sptr,
                                                      const int *__restrict__ col, const V *__restrict__ val,
                                                      const float *__restrict__ rsc, const X *__restrict__ diag,
                                                      const X *__restrict__ f, int fs, const X *__restrict__ uo,
                                                      X *__restrict__ un)
{
    constexpr int BST = LMAX * LMAX + 1; // odd stride between chunk blocks: spreads LDS banks
    extern __shared__ __align__(16) unsigned char smraw[];
    X *sm = reinterpret_cast<X *>(smraw);
    X *g = sm, *ux = sm + 256, *dd = sm + 512, *B = sm + 768;
    const int size = n / nchunks, rest = n - size * nchunks;
    const int c0 = xcd_tile(blockIdx.x, gridDim.x) * cpb;
    const int c1 = min(c0 + cpb, nchunks);
    const int row0 = c0 < rest ? c0 * (size + 1) : c0 * size + rest;
    const int row1 = c1 < rest ? c1 * (size + 1) : c1 * size + rest;
    const int r = threadIdx.x;
    const int i = row0 + r;
    if (i < row1) {
        const int big = rest * (size + 1);
        const int c = i < big ? i / (size + 1) : rest + (i - big) / size;
        const int ns = c < rest ? c * (size + 1) : c * size + rest;
        const int len = c < rest ? size + 1 : size;
        X *Brow = B + (size_t)(c - c0) * BST + (i - ns) * LMAX;
        for (int k = 0; k < len; k++) Brow[k] = 0;
        X acc = f[(size_t)i * fs];
        const X sc = row_scaled<V>::value ? (X)rsc[i] : (X)1;
        const int s = i / WAVE, lane = i % WAVE;
        const int nsl = (n + WAVE - 1) / WAVE;
        chunk_row_gather<SELL_U, V, X, PK>(sptr[s], sptr[s + 1], PK ? sptr[nsl + 1 + s] : 0, (size_t)lane, i, ns, len,
                                           sc, col, val, uo, Brow, acc);
        dd[r] = diag[i];
        g[r] = acc;
        ux[r] = uo ? uo[i] : (X)0;
    }
    __syncthreads();
    const int c = c0 + (int)threadIdx.x;
    if (c < c1) {
        const int ns = c < rest ? c * (size + 1) : c * size + rest;
        const int len = c < rest ? size + 1 : size;
        const int rl = ns - row0;
        const X *Bq = B + (size_t)(c - c0) * BST;
        X u[LMAX];
#pragma unroll
        for (int t = 0; t < LMAX; t++) u[t] = t < len ? ux[rl + t] : (X)0;
#pragma unroll
        for (int pass = 0; pass < 2; pass++) {
#pragma unroll
            for (int tt = 0; tt < LMAX; tt++) {
                const int t = pass == 0 ? tt : LMAX - 1 - tt;
                if (t < len) {
                    const X d = dd[rl + t];
                    if (d != (X)0) {
                        X x = g[rl + t];
#pragma unroll
                        for (int q = 0; q < LMAX; q++)
                            if (q != t && q < len) x -= Bq[t * LMAX + q] * u[q];
                        u[t] = x / d;
                    }
                }
            }
        }
#pragma unroll
        for (int t = 0; t < LMAX; t++)
            if (t < len) ux[rl + t] = u[t];
    }
    __syncthreads();
    if (i < row1) un[i] = ux[r];
}

// ---------------------------------------------------------------------------
// Row-group variants for levels with long rows (coarse AMG levels hold rows of
// tens to hundreds of entries), over the SELL-G layout: a slice holds
// R = 64/G consecutive rows, G lanes per row, and slot row u of the slice
// holds entries u*G .. u*G+G-1 of each row (row order, diagonal first; short
// rows padded with zero values), so lane g of a group sums entries g, g+G, ...
// and every slot load of a wavefront is one contiguous transaction.  Groups
// reduce with cross-lane adds.  rsc[i] scales row i for FP16 storage.
// ---------------------------------------------------------------------------
template <int G, typename X> __device__ __forceinline__ X group_sum(X v)
{
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
    return v;
}

//   mode 0: y = A x;  1: y = rhs - A x;  2: y += A x
template <int G, typename V, typename X, bool PK>
__global__ void __launch_bounds__(256) k_csr_spmv_grp(int n, const int *__restrict__ sptr, const int *__restrict__ col,
                                                      const V *__restrict__ val, const float *__restrict__ rsc,
                                                      const X *__restrict__ x, int xs, const X *__restrict__ rhs,
                                                      int rs, X *__restrict__ y, int mode, const X *__restrict__ d2,
                                                      X *__restrict__ u2, X w2)
{
    constexpr int R = WAVE / G;
    const int nsl = (n + R - 1) / R;
    const int s = xcd_tile(blockIdx.x, gridDim.x) * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (s >= nsl) return; // whole wavefront
    const int lane = threadIdx.x % WAVE, gl = lane % G;
    const int row = s * R + lane / G;
    X acc = sell_dot<V, X, PK>(sptr[s], sptr[s + 1], PK ? sptr[nsl + 1 + s] : 0, lane, row, col, val, x, xs);
    acc = group_sum<G>(acc);
    if (row >= n || gl != 0) return;
    if (row_scaled<V>::value) acc *= (X)rsc[row];
    if (mode == 0) y[row] = acc;
    else if (mode == 1) y[row] = rhs[(size_t)row * rs] - acc;
    else y[row] += acc;
    if (u2) poly_first_row(row, d2, acc, u2, w2); // mode 0 only (launcher checks)
}

template <int G, typename V, typename X, bool PK>
__global__ void __launch_bounds__(256) k_poly_step_grp(int n, const int *__restrict__ sptr, const int *__restrict__ col,
                                                       const V *__restrict__ val, const float *__restrict__ rsc,
                                                       const X *__restrict__ dinv, const X *__restrict__ f, int fs,
                                                       const X *__restrict__ uo, X *__restrict__ un, X omega)
{
    constexpr int R = WAVE / G;
    const int nsl = (n + R - 1) / R;
    const int s = xcd_tile(blockIdx.x, gridDim.x) * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (s >= nsl) return;
    const int lane = threadIdx.x % WAVE, gl = lane % G;
    const int row = s * R + lane / G;
    X acc = uo ? sell_dot<V, X, PK>(sptr[s], sptr[s + 1], PK ? sptr[nsl + 1 + s] : 0, lane, row, col, val, uo, 1)
               : (X)0;
    acc = group_sum<G>(acc);
    if (row >= n || gl != 0) return;
    if (row_scaled<V>::value) acc *= (X)rsc[row];
    const X r = f[(size_t)row * fs] - acc;
    un[row] = (uo ? uo[row] : (X)0) + omega * dinv[row] * r;
}

// chunked hybrid SSOR (as k_gs_chunk_lds) with phase A done by row groups
// over plain CSR (ia, ja, val; diagonal first, kept apart in diag): a block
// covers whole chunks, so its rows start anywhere in a slice of any layout;
// plain rows carry no padding
template <int LMAX, int G, typename V, typename X>
__global__ void __launch_bounds__(256) k_gs_chunk_grp(int n, int nchunks, int cpb, const int *__restrict__ ia,
                                                      const int *__restrict__ ja, const V *__restrict__ val,
                                                      const float *__restrict__ rsc, const X *__restrict__ diag,
                                                      const X *__restrict__ f, int fs, const X *__restrict__ uo,
                                                      X *__restrict__ un)
{
    constexpr int BST = LMAX * LMAX + 1;
    extern __shared__ __align__(16) unsigned char smraw[];
    X *sm = reinterpret_cast<X *>(smraw);
    X *g = sm, *ux = sm + 256, *dd = sm + 512, *B = sm + 768;
    const int size = n / nchunks, rest = n - size * nchunks;
    const int c0 = xcd_tile(blockIdx.x, gridDim.x) * cpb;
    const int c1 = min(c0 + cpb, nchunks);
    const int row0 = c0 < rest ? c0 * (size + 1) : c0 * size + rest;
    const int row1 = c1 < rest ? c1 * (size + 1) : c1 * size + rest;
    for (int e = threadIdx.x; e < cpb * BST; e += blockDim.x) B[e] = 0;
    __syncthreads();
    const int gl = threadIdx.x % G;
    const int big = rest * (size + 1);
    for (int r = threadIdx.x / G; r < row1 - row0; r += blockDim.x / G) {
        const int i = row0 + r;
        const int c = i < big ? i / (size + 1) : rest + (i - big) / size;
        const int ns = c < rest ? c * (size + 1) : c * size + rest;
        const int len = c < rest ? size + 1 : size;
        X *Brow = B + (size_t)(c - c0) * BST + (i - ns) * LMAX;
        const X sc = row_scaled<V>::value ? (X)rsc[i] : (X)1;
        X acc = 0;
        csr_row_gather<GRP_U, G, V, X>(ia[i] + 1 + gl, ia[i + 1], i, ns, len, sc, ja, val, uo, Brow, acc);
        acc = group_sum<G>(acc);
        if (gl == 0) {
            dd[r] = diag[i];
            g[r] = f[(size_t)i * fs] + acc;
            ux[r] = uo ? uo[i] : (X)0;
        }
    }
    __syncthreads();
    const int c = c0 + (int)threadIdx.x;
    if (c < c1) {
        const int ns = c < rest ? c * (size + 1) : c * size + rest;
        const int len = c < rest ? size + 1 : size;
        const int rl = ns - row0;
        const X *Bq = B + (size_t)(c - c0) * BST;
        X u[LMAX];
#pragma unroll
        for (int t = 0; t < LMAX; t++) u[t] = t < len ? ux[rl + t] : (X)0;
#pragma unroll
        for (int pass = 0; pass < 2; pass++) {
#pragma unroll
            for (int tt = 0; tt < LMAX; tt++) {
                const int t = pass == 0 ? tt : LMAX - 1 - tt;
                if (t < len) {
                    const X d = dd[rl + t];
                    if (d != (X)0) {
                        X x = g[rl + t];
#pragma unroll
                        for (int q = 0; q < LMAX; q++)
                            if (q != t && q < len) x -= Bq[t * LMAX + q] * u[q];
                        u[t] = x / d;
                    }
                }
            }
        }
#pragma unroll
        for (int t = 0; t < LMAX; t++)
            if (t < len) ux[rl + t] = u[t];
    }
    __syncthreads();
    for (int r = threadIdx.x; r < row1 - row0; r += blockDim.x) un[row0 + r] = ux[r];
}

// Part of one row's Gauss-Seidel sum for a lane of a row group: entries
// k, k + G, ... < ke of the plain CSR row i, only the dependent couplings
// (DEP: j < i forward, j > i backward; x = the new values, read coherently)
// or only the others (x = the entry values, ordinary loads).  U entries per
// lane are loaded before any is used, then U gathers of x.
template <int U, int G, typename V, typename X, bool DEP>
__device__ __forceinline__ X gs_row_part(int k, int ke, int i, int dir, const int *__restrict__ ja,
                                         const V *__restrict__ val, const X *x)
{
    X acc = 0;
    for (; k < ke; k += U * G) {
        int j[U];
        V v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int kk = k + u * G;
            j[u] = kk < ke ? ja[kk] : -1;
            v[u] = kk < ke ? val[kk] : (V)0;
        }
        X xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const bool dep = dir > 0 ? j[u] < i : j[u] > i;
            const bool use = j[u] >= 0 && j[u] != i && dep == DEP;
            if (DEP) xv[u] = use ? ld_coherent(x + j[u]) : (X)0;
            else xv[u] = use ? x[j[u]] : (X)0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) acc += (X)cvt(v[u]) * xv[u];
    }
    return acc;
}

// One Gauss-Seidel sweep as k_gs_sweep for levels with long rows: a slice
// holds R = 64/G rows of one dependency level (perm[p] = row, -1 for
// padding), each row summed by a group of G lanes over its plain CSR row
// (ia, ja, val; the diagonal first, kept apart in diag and rsc, both indexed
// by row).  The couplings on entry values are summed before the wait for the
// dependency slices, the dependent ones after it; one lane per group then
// writes the new value.  A row with a zero diagonal keeps its old value.
template <int G, typename V, typename X>
__global__ void __launch_bounds__(256) k_gs_sweep_grp(int nsl, const int *__restrict__ perm,
                                                      const int *__restrict__ ia, const int *__restrict__ ja,
                                                      const V *__restrict__ val, const float *__restrict__ rsc,
                                                      const X *__restrict__ diag, const X *__restrict__ f, int fs,
                                                      const X *uo, X *un, const int *__restrict__ dep_ptr,
                                                      const int *__restrict__ dep_list, int *sflag, int epoch,
                                                      int dir, unsigned long long *ticket, unsigned long long base,
                                                      int *err)
{
    constexpr int R = WAVE / G;
    const int lane = threadIdx.x % WAVE, gl = lane % G;
    for (;;) {
        const int s = claim_slice(ticket, base, lane);
        if (s < 0 || s >= nsl) break;
        const int i = perm[s * R + lane / G];
        const int k0 = i >= 0 ? ia[i] + 1 + gl : 0, k1 = i >= 0 ? ia[i + 1] : 0;
        X acc = uo ? gs_row_part<GRP_U, G, V, X, false>(k0, k1, i, dir, ja, val, uo) : (X)0;
        wait_slices(dep_ptr, dep_list, sflag, s, epoch, lane, err);
        acc += gs_row_part<GRP_U, G, V, X, true>(k0, k1, i, dir, ja, val, un);
        acc = group_sum<G>(acc);
        if (i >= 0 && gl == 0) {
            const X d = diag[i];
            const X sc = row_scaled<V>::value ? (X)rsc[i] : (X)1;
            const X res = f[(size_t)i * fs] - sc * acc;
            st_coherent(un + i, d != (X)0 ? res / d : (uo ? uo[i] : (X)0));
        }
        st_flag_release(sflag + s, epoch);
    }
}

// dense y = M x (row-major n x n), one wavefront per row
template <typename X>
__global__ void __launch_bounds__(256) k_dense_gemv(int n, int ld, const X *__restrict__ M, const X *__restrict__ x,
                                                    X *__restrict__ y)
{
    // one wavefront per row; rows padded to ld (a multiple of 16 bytes), so
    // every lane loads 16-byte vectors, four of them before it adds
    typedef typename vec16<X>::T V;
    constexpr int E = sizeof(V) / sizeof(X);
    const int row = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int lane = threadIdx.x % WAVE;
    if (row >= n) return;
    const V *m = reinterpret_cast<const V *>(M + (size_t)row * ld);
    const V *xv = reinterpret_cast<const V *>(x);
    const int nv = ld / E;
    X acc = 0;
    int j = lane;
    for (; j + 3 * WAVE < nv; j += 4 * WAVE) {
        V a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a[u] = __builtin_nontemporal_load(m + j + u * WAVE);
            b[u] = xv[j + u * WAVE];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int e = 0; e < E; e++) acc += a[u][e] * b[u][e];
    }
    for (; j < nv; j += WAVE) {
        const V a = __builtin_nontemporal_load(m + j), b = xv[j];
#pragma unroll
        for (int e = 0; e < E; e++) acc += a[e] * b[e];
    }
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) y[row] = acc;
}

// y[i] = (float) x[i*stride + off]: restriction of an FP64 residual into an FP32 level vector
__global__ void k_gather_cast(int n, const double *__restrict__ x, int stride, int off, float *__restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (float)x[(size_t)i * stride + off];
}

// first polynomial step from a zero guess: u = 0 + w D^{-1} (f - 0), one
// thread per row (the same operations as k_poly_step with uo = NULL)
template <typename X>
__global__ void k_poly_first(int n, const X *__restrict__ dinv, const X *__restrict__ f, int fs, X *__restrict__ un,
                             X omega)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const X r = f[(size_t)i * fs] - (X)0;
    un[i] = (X)0 + omega * dinv[i] * r;
}

// the cast of the restricted residual (k_gather_cast) and the first
// polynomial step of level 0 from a zero guess (k_poly_first) in one pass:
// y = (float) x[i*stride + off], un = 0 + omega dinv (y - 0)
__global__ void k_cast_first(int n, const double *__restrict__ x, int stride, int off, float *__restrict__ y,
                             const float *__restrict__ dinv, float *__restrict__ un, float omega)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float f = (float)x[(size_t)i * stride + off];
    y[i] = f;
    const float r = f - 0.0f;
    un[i] = 0.0f + omega * dinv[i] * r;
}

__global__ void k_widen(int n, const float *__restrict__ x, double *__restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (double)x[i];
}

__global__ void k_fill_f32(int n, float *__restrict__ y, float v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}

__global__ void k_gather_stride(int n, const double *__restrict__ x, int stride, int off, double *__restrict__ y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[(size_t)i * stride + off];
}

// x[i*b + j] = (j == pidx) ? xp[i] : 0   (CPR prolongation type 0)
__global__ void k_prolong(int nb, int b, int pidx, const double *__restrict__ xp, double *__restrict__ x)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    for (int j = 0; j < b; j++) x[(size_t)i * b + j] = (j == pidx) ? xp[i] : 0.0;
}

__global__ void k_fill(size_t n, double *__restrict__ y, double v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) y[i] = v;
}

__global__ void k_copy(size_t n, const double *__restrict__ x, double *__restrict__ y)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) y[i] = x[i];
}

// y = a*x + y with a read from device memory (*pa, negated if neg) or given
__global__ void k_axpy(size_t n, const double *__restrict__ pa, double a, int neg, const double *__restrict__ x,
                       double *__restrict__ y)
{
    double s = pa ? *pa : a;
    if (neg) s = -s;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) y[i] += s * x[i];
}

// y = s*y, s = 1/sqrt(*pnorm2) if pnorm2 (zero norm leaves y), else s = a
__global__ void k_scale(size_t n, const double *__restrict__ pnorm2, double a, double *__restrict__ y)
{
    double s = a;
    if (pnorm2) {
        double t = sqrt(*pnorm2);
        if (t == 0.0) return;
        s = 1.0 / t;
    }
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) y[i] *= s;
}

// Deterministic dot product with an optional fused MGS update:
//   if (v_prev): y -= (*h_prev) * v_prev   (h_prev from device memory)
//   out = <u, y>  (u == y gives the squared norm)
// Block partials land in part[]; the last block to finish sums them in a
// fixed order and writes *out (and, if hout, the same value to hout, a
// device address of mapped host memory, by a system-scope store the host
// may poll), then resets the counter.
__global__ void __launch_bounds__(256) k_dot_mgs(size_t n, const double *u, double *y,
                                                 const double *__restrict__ v_prev, const double *__restrict__ h_prev,
                                                 double *part, unsigned int *count, double *out, double *hout)
{
    __shared__ double sm[256 / WAVE];
    __shared__ bool last;
    const double h = v_prev ? *h_prev : 0.0;
    double acc = 0.0;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    // u may be y (a squared norm): then the product uses the updated y.
    // Entries i, i + stride, ... touch disjoint elements, so each thread
    // loads DOT_U of them before it stores any (the compiler cannot know
    // that u and y only meet at equal indices).
    const bool self = (u == y);
    if (v_prev) {
        for (; i + (DOT_U - 1) * stride < n; i += DOT_U * stride) {
            double yv[DOT_U], vv[DOT_U], uv[DOT_U];
#pragma unroll
            for (int q = 0; q < DOT_U; q++) {
                yv[q] = y[i + q * stride];
                vv[q] = v_prev[i + q * stride];
                uv[q] = self ? 0.0 : u[i + q * stride];
            }
#pragma unroll
            for (int q = 0; q < DOT_U; q++) {
                const double yi = yv[q] - h * vv[q];
                y[i + q * stride] = yi;
                acc += (self ? yi : uv[q]) * yi;
            }
        }
        for (; i < n; i += stride) {
            const double yi = y[i] - h * v_prev[i];
            const double ui = self ? yi : u[i];
            y[i] = yi;
            acc += ui * yi;
        }
    } else {
        for (; i + (DOT_U - 1) * stride < n; i += DOT_U * stride) {
            double yv[DOT_U], uv[DOT_U];
#pragma unroll
            for (int q = 0; q < DOT_U; q++) {
                yv[q] = y[i + q * stride];
                uv[q] = u[i + q * stride];
            }
#pragma unroll
            for (int q = 0; q < DOT_U; q++) acc += uv[q] * yv[q];
        }
        for (; i < n; i += stride) acc += u[i] * y[i];
    }
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (threadIdx.x % WAVE == 0) sm[threadIdx.x / WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int w = 0; w < (int)(blockDim.x / WAVE); w++) b += sm[w];
        // the partial is a device-scope store: wait for it to complete, then
        // count (no L2 write-back/invalidate, as in st_flag_release)
        __hip_atomic_store(part + blockIdx.x, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
        unsigned int prev = __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (threadIdx.x < WAVE) {
        double a2 = 0.0;
        for (unsigned int b = threadIdx.x; b < gridDim.x; b += WAVE)
            a2 += __hip_atomic_load(part + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int off = WAVE / 2; off > 0; off >>= 1) a2 += __shfl_xor(a2, off);
        if (threadIdx.x == 0) {
            *out = a2;
            __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (hout) __hip_atomic_store(hout, a2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// Sum of v over the workgroup (all threads call it), in a fixed order:
// wavefront shuffles, then the wavefront sums in order.  Every thread gets
// the total; red holds blockDim.x / 64 doubles and is free again on return.
__device__ __forceinline__ double block_sum(double v, double *red)
{
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = v;
    __syncthreads();
    double s = 0.0;
    for (int q = 0; q < (int)(blockDim.x / WAVE); q++) s += red[q];
    __syncthreads();
    return s;
}

// One whole modified Gram-Schmidt step of the Arnoldi process in a single
// cooperative launch (every workgroup resident).  Each thread keeps TM
// entries of w in registers through the step:
//   for k = 0..m-1:  h_k = <v_k, w>;  w -= h_k v_k
//   h_m = <w, w>;    w /= sqrt(h_m)   (w kept as is when h_m = 0)
// so a step reads every v_k once and w once and writes w once (the launch-
// per-projection form reads and writes w for every k).  Each reduction is
// deterministic: per-workgroup partials in part[k * gridDim.x + b], summed in
// block order by the last workgroup to arrive (counter cnt), which publishes
// h_k in hval[k] and raises *flag to base + k + 1; the others wait on the
// flag (bounded by spin_limit polls: an overflow sets err[0] and the step
// goes on with whatever hval[k] holds; a waiter that gave up can then arrive
// at the next reduction before the counter is reset, so after such a step
// the host must clear cnt and flag before the next launch).
#define MGS_BS 1024 // threads per workgroup of k_mgs_coop: few workgroups, few arrivals per reduction

template <int TM>
__global__ void __launch_bounds__(MGS_BS) k_mgs_coop(size_t n, const double *__restrict__ V, size_t ld, int m,
                                                     double *__restrict__ w, double *hval, double *part,
                                                     unsigned int *cnt, int *flag, int base, int *err,
                                                     double *hhost, int spin_limit, int hpos, double hseq)
{
    __shared__ double red[MGS_BS / WAVE];
    __shared__ double hb;
    __shared__ int last;
    // wavefront q owns the contiguous entries [q * 64 * TM, (q + 1) * 64 * TM):
    // entry t of a lane is wb + 64 t + lane, so every load is one 512-byte
    // row at a wave-uniform address plus the lane offset (no per-entry
    // address registers)
    const unsigned int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE);
    const size_t wb = (size_t)wave * WAVE * TM;
    const unsigned int lane = threadIdx.x % WAVE;
    double wr[TM];
#pragma unroll
    for (int t = 0; t < TM; t++) {
        const size_t e = wb + (size_t)t * WAVE + lane;
        wr[t] = e < n ? w[e] : 0.0;
    }
    for (int k = 0; k <= m; k++) {
        double vr[TM];
        double acc = 0.0;
        if (k < m) {
            const double *v = V + (size_t)k * ld + wb + lane;
#pragma unroll
            for (int t = 0; t < TM; t++) vr[t] = wb + (size_t)t * WAVE + lane < n ? v[t * WAVE] : 0.0;
#pragma unroll
            for (int t = 0; t < TM; t++) acc += vr[t] * wr[t];
        } else {
#pragma unroll
            for (int t = 0; t < TM; t++) acc += wr[t] * wr[t];
        }
        const double b = block_sum(acc, red);
        const int target = base + k + 1;
        double *pk = part + (size_t)k * gridDim.x;
        if (threadIdx.x == 0) {
            st_coherent(pk + blockIdx.x, b);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_s_waitcnt(0);
            const unsigned int prev = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            last = prev == gridDim.x - 1;
        }
        __syncthreads();
        if (last) {
            // the last workgroup to arrive sums the partials (every thread a
            // strided share, then the fixed-order block sum) and publishes h_k
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            double q = 0.0;
            for (unsigned int i = threadIdx.x; i < gridDim.x; i += blockDim.x) q += ld_coherent(pk + i);
            const double sum = block_sum(q, red);
            if (threadIdx.x == 0) {
                st_coherent(hval + k, sum);
                __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                st_flag_release(flag, target);
                hb = sum;
            }
            if (k == m && hhost) {
                // the whole column to pinned host memory (system-scope
                // stores), then, if hpos >= 0, the sequence number hseq at
                // hhost[hpos] once every lane's column stores are ordered
                // before it: the host may poll that word instead of waiting
                // for the launch to complete
                for (int q = threadIdx.x; q <= m; q += blockDim.x)
                    __hip_atomic_store(hhost + q, q == m ? sum : ld_coherent(hval + q), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_SYSTEM);
                if (hpos >= 0) {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
                    __syncthreads();
                    if (threadIdx.x == 0) __hip_atomic_store(hhost + hpos, hseq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
                }
            }
        } else if (threadIdx.x == 0) {
            int spins = 0;
            while (!flag_ready(ld_flag(flag), target)) {
                if (++spins > spin_limit) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            hb = ld_coherent(hval + k);
        }
        __syncthreads();
        const double h = hb;
        if (k < m) {
#pragma unroll
            for (int t = 0; t < TM; t++) wr[t] -= h * vr[t];
        } else {
            const double nrm = sqrt(h);
            const double sc = nrm != 0.0 ? 1.0 / nrm : 1.0;
#pragma unroll
            for (int t = 0; t < TM; t++) {
                const size_t e = wb + (size_t)t * WAVE + lane;
                if (e < n) w[e] = wr[t] * sc;
            }
        }
        __syncthreads(); // hb and last are rewritten by the next reduction
    }
}

// w = sum_j y[j] * V[j]   (j = 0..m-1, V rows of length n, ld apart), y on host
__global__ void k_lincomb(size_t n, int m, const double *__restrict__ V, size_t ld, const double *__restrict__ y,
                          double *__restrict__ w)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double acc = y[m - 1] * V[(size_t)(m - 1) * ld + i];
        for (int j = m - 2; j >= 0; j--) acc += y[j] * V[(size_t)j * ld + i];
        w[i] = acc;
    }
}

// the same with the coefficients passed by value (no copy to the device):
// w = sum_j y[j] V[j], or with ADD x += sum_j y[j] V[j] (x + the sum, as an
// axpy of the sum with factor 1 would give)
template <bool ADD>
__global__ void k_lincomb_args(size_t n, int m, const double *__restrict__ V, size_t ld, jxamd_coefs y,
                               double *__restrict__ w)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double acc = y.c[m - 1] * V[(size_t)(m - 1) * ld + i];
        for (int j = m - 2; j >= 0; j--) acc += y.c[j] * V[(size_t)j * ld + i];
        w[i] = ADD ? w[i] + acc : acc;
    }
}

inline int nblocks_stream(size_t n, int bs) {
    size_t b = (n + bs - 1) / bs;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return (int)b;
}

} // namespace

// ===========================================================================
// extern "C" launchers (device pointers; stream is a hipStream_t)
// ===========================================================================
extern "C" {

static int chk(hipError_t e) { return e == hipSuccess ? 0 : JXAMD_ERROR_DEVICE; }

int jxamd_dev_count(int *n)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return 0;
}
int jxamd_dev_malloc(void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 8;
    return chk(hipMalloc(p, bytes));
}
int jxamd_dev_free(void *p) { return p ? chk(hipFree(p)) : 0; }
int jxamd_h2d(void *dst, const void *src, size_t bytes, void *stream)
{
    if (!bytes) return 0;
    return chk(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
}
int jxamd_d2h(void *dst, const void *src, size_t bytes, void *stream)
{
    if (!bytes) return 0;
    return chk(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
}
int jxamd_d2d(void *dst, const void *src, size_t bytes, void *stream)
{
    if (!bytes) return 0;
    return chk(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
}
int jxamd_stream_create(void **s)
{
    hipStream_t st;
    int e = chk(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *s = (void *)st;
    return e;
}
int jxamd_stream_destroy(void *s) { return s ? chk(hipStreamDestroy((hipStream_t)s)) : 0; }
int jxamd_stream_sync(void *s) { return chk(hipStreamSynchronize((hipStream_t)s)); }
int jxamd_event_create_sync(void **e)
{
    hipEvent_t ev;
    int r = chk(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    *e = (void *)ev;
    return r;
}
int jxamd_event_sync(void *e) { return chk(hipEventSynchronize((hipEvent_t)e)); }
int jxamd_host_alloc(void **p, size_t bytes)
{
    return chk(hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocMapped | hipHostMallocCoherent));
}
int jxamd_host_dev_ptr(void **d, void *h) { return chk(hipHostGetDevicePointer(d, h, 0)); }
int jxamd_host_free(void *p) { return p ? chk(hipHostFree(p)) : 0; }
int jxamd_event_create(void **e)
{
    hipEvent_t ev;
    int r = chk(hipEventCreate(&ev));
    *e = (void *)ev;
    return r;
}
int jxamd_event_destroy(void *e) { return e ? chk(hipEventDestroy((hipEvent_t)e)) : 0; }
int jxamd_event_record(void *e, void *s) { return chk(hipEventRecord((hipEvent_t)e, (hipStream_t)s)); }
int jxamd_event_elapsed(float *ms, void *e0, void *e1)
{
    return chk(hipEventElapsedTime(ms, (hipEvent_t)e0, (hipEvent_t)e1));
}
int jxamd_last_error(void) { return chk(hipGetLastError()); }

int jxamd_k_bsr_spmv(int b, int nb, int unit, const int *sptr, const int *col, const double *val, const double *x,
                     const double *rhs, double *y, int mode, double alpha, double beta, void *stream)
{
    const int nsl = (nb + WAVE - 1) / WAVE;
    const int grid = (nsl + 3) / 4;
    if (grid == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    switch (b) {
    // slots per batch, measured back to back: 3x3 blocks on the 60x220x85
    // system 8 (a 7-point row in one batch; 117.8 us with one slot per step,
    // 103.3 us with eight, 103.6 with six, 104.5 with four); 4x4 blocks on
    // the 100^3 system 6 (167.1 us with eight: 130 VGPRs, three waves per
    // SIMD; 153.8 us with six, 154.7 with four)
#define SPMV_B(BS, SU)                                                                                                 \
    do {                                                                                                               \
        if (unit) k_bsr_spmv<BS, SU, true><<<grid, 256, 0, st>>>(nb, sptr, col, val, x, rhs, y, mode, alpha, beta);  \
        else k_bsr_spmv<BS, SU, false><<<grid, 256, 0, st>>>(nb, sptr, col, val, x, rhs, y, mode, alpha, beta);       \
    } while (0)
    case 2: SPMV_B(2, 8); break;
    case 3: SPMV_B(3, 8); break;
    case 4: SPMV_B(4, 6); break;
    default: return JXAMD_ERROR_ARG;
#undef SPMV_B
    }
    return chk(hipGetLastError());
}

// (vt, xt) dispatch: F64/F64, F32/F32, F16/F32
// level-matrix kernels over SELL layouts: vt may carry JXAMD_T_PK16 (packed
// int16 column offsets; slice word offsets after sptr)
#define DISPATCH_VXP(vt, xt, CALL)                                                                                     \
    do {                                                                                                               \
        if ((vt) & JXAMD_T_PK16) {                                                                                     \
            constexpr bool P_ = true;                                                                                  \
            DISPATCH_VX((vt) & ~JXAMD_T_PK16, xt, CALL);                                                               \
        } else {                                                                                                       \
            constexpr bool P_ = false;                                                                                 \
            DISPATCH_VX(vt, xt, CALL);                                                                                 \
        }                                                                                                              \
    } while (0)

#define DISPATCH_VX(vt, xt, CALL)                                                                                      \
    do {                                                                                                               \
        if ((vt) == JXAMD_T_F64 && (xt) == JXAMD_T_F64) { typedef double V_; typedef double X_; CALL; }                \
        else if ((vt) == JXAMD_T_F32 && (xt) == JXAMD_T_F32) { typedef float V_; typedef float X_; CALL; }             \
        else if ((vt) == JXAMD_T_F16 && (xt) == JXAMD_T_F32) { typedef __half V_; typedef float X_; CALL; }            \
        else return JXAMD_ERROR_ARG;                                                                                   \
    } while (0)

int jxamd_k_csr_spmv(int vt, int xt, int n, const int *sptr, const int *col, const void *val, const float *rsc,
                     const void *x, int xs, const void *rhs, int rs, void *y, int mode, void *stream)
{
    const int nsl = (n + WAVE - 1) / WAVE;
    const int grid = (nsl + 3) / 4;
    if (grid == 0) return 0;
    DISPATCH_VXP(vt, xt,
                 (k_csr_spmv<V_, X_, P_><<<grid, 256, 0, (hipStream_t)stream>>>(
                     n, sptr, col, (const V_ *)val, rsc, (const X_ *)x, xs, (const X_ *)rhs, rs, (X_ *)y, mode,
                     nullptr, nullptr, (X_)0)));
    return chk(hipGetLastError());
}

int jxamd_k_poly_step(int vt, int xt, int n, const int *sptr, const int *col, const void *val, const float *rsc,
                      const void *dinv, const void *f, int fs, const void *uo, void *un, double omega, void *stream)
{
    const int nsl = (n + WAVE - 1) / WAVE;
    const int grid = (nsl + 3) / 4;
    if (grid == 0) return 0;
    DISPATCH_VXP(vt, xt,
                 (k_poly_step<V_, X_, P_><<<grid, 256, 0, (hipStream_t)stream>>>(
                     n, sptr, col, (const V_ *)val, rsc, (const X_ *)dinv, (const X_ *)f, fs, (const X_ *)uo,
                     (X_ *)un, (X_)omega)));
    return chk(hipGetLastError());
}

int jxamd_k_poly_first(int xt, int n, const void *dinv, const void *f, int fs, void *un, double omega, void *stream)
{
    if (n == 0) return 0;
    const int grid = (n + 255) / 256;
    if (xt == JXAMD_T_F64)
        k_poly_first<double><<<grid, 256, 0, (hipStream_t)stream>>>(n, (const double *)dinv, (const double *)f, fs,
                                                                     (double *)un, omega);
    else if (xt == JXAMD_T_F32)
        k_poly_first<float><<<grid, 256, 0, (hipStream_t)stream>>>(n, (const float *)dinv, (const float *)f, fs,
                                                                    (float *)un, (float)omega);
    else
        return JXAMD_ERROR_ARG;
    return chk(hipGetLastError());
}

int jxamd_sweep_waves(int nsl)
{
    const int cap = 2048; /* persistent wavefronts: 8 per CU */
    return nsl < cap ? (nsl > 0 ? nsl : 1) : cap;
}

int jxamd_k_gs_sweep(int vt, int xt, int nsl, const int *perm, const int *sptr, const int *col, const void *val,
                     const float *rsc, const void *diag, const void *f, int fs, const void *uo, void *un,
                     const int *dep_ptr, const int *dep_list, int *sflag, int epoch, int dir,
                     unsigned long long *ticket, unsigned long long *base, int *err, void *stream)
{
    if (nsl == 0) return 0;
    const int waves = jxamd_sweep_waves(nsl);
    const int grid = (waves + 3) / 4;
    DISPATCH_VX(vt, xt,
                (k_gs_sweep<V_, X_><<<grid, 256, 0, (hipStream_t)stream>>>(
                    nsl, perm, sptr, col, (const V_ *)val, rsc, (const X_ *)diag, (const X_ *)f, fs, (const X_ *)uo,
                    (X_ *)un, dep_ptr, dep_list, sflag, epoch, dir, ticket, *base, err)));
    *base += (unsigned long long)nsl + (unsigned long long)grid * 4;
    return chk(hipGetLastError());
}

int jxamd_k_bhgs_gather(int npos, int b, const int *perm, const double *x0, const double *xp, int pidx, double *out,
                        void *stream)
{
    if (npos == 0) return 0;
    const long n = (long)npos * b;
    k_bhgs_gather<<<(int)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(npos, b, perm, x0, xp, pidx, out);
    return chk(hipGetLastError());
}

int jxamd_k_bhgs_multi(int b, int nsl, int K, const int *vorder, const int *perm, const int *sptr, const int *col,
                       const double *val, const double *dinv, const double *rhs, const double *xsl0, double *xs,
                       long long xstride, double *xout, const int *dlo_ptr, const int *dlo_list, const int *dall_ptr,
                       const int *dall_list, int *sflag, int epoch, double w, double omega,
                       unsigned long long *ticket, unsigned long long *base, int *err, void *stream)
{
    if (nsl == 0 || K < 1) return 0;
    if (!xsl0 || !xs || !xout || (K > 1 && (!vorder || !dall_ptr || !dall_list))) return JXAMD_ERROR_ARG;
    if ((long long)nsl * K > 0x7fffffffLL) return JXAMD_ERROR_ARG;
    const int nvs = nsl * K;
    const int waves = jxamd_sweep_waves(nvs);
    const int grid = (waves + 3) / 4;
    hipStream_t st = (hipStream_t)stream;
    const unsigned long long b0 = *base;
#define LAUNCH_BHGS(BS)                                                                                                \
    k_bhgs_multi<BS><<<grid, 256, 0, st>>>(nsl, nvs, vorder, perm, sptr, col, val, dinv, rhs, xsl0, xs, xstride,      \
                                           xout, dlo_ptr, dlo_list, dall_ptr, dall_list, sflag, epoch, w, omega,       \
                                           ticket, b0, err)
    switch (b) {
    case 2: LAUNCH_BHGS(2); break;
    case 3: LAUNCH_BHGS(3); break;
    case 4: LAUNCH_BHGS(4); break;
    default: return JXAMD_ERROR_ARG;
    }
#undef LAUNCH_BHGS
    *base += (unsigned long long)nvs + (unsigned long long)grid * 4;
    return chk(hipGetLastError());
}

int jxamd_k_bhgs_part(int b, int vt, int nb, int nparts, const int *sptr, const int *pptr, const int *col,
                      const void *val, const float *rsc, const void *dinv, const double *rhs, const double *xo,
                      const void *xp, int xp_f32, int pidx, double *xn, double w, double omega, void *stream)
{
    if (vt != JXAMD_T_F64 && vt != JXAMD_T_F32 && vt != JXAMD_T_F16) return JXAMD_ERROR_ARG;
    if (nb == 0) return 0;
    const int maxlen = jxamd_gs_chunk_lds_rows(nb, nparts);
    if (maxlen < 1 || maxlen > 8 || (!xo && !xp) || (vt == JXAMD_T_F16 && (!rsc || !pptr))) return JXAMD_ERROR_ARG;
    hipStream_t st = (hipStream_t)stream;
    // slots per batch, measured on the 60x220x85 system: FP64 blocks 6
    // (129 -> 124 us per sweep against 4; 8 needs 142 VGPRs and drops to
    // three waves per SIMD: 140 us; one slot per step was slower still),
    // FP32 4, FP16 2 (66 -> 52 us against 4: 75 instead of 109 VGPRs, six
    // waves per SIMD instead of four)
    constexpr int PU64 = 6, PU32 = 4, PU16 = 2;
#define LAUNCH_PART(BS, LM)                                                                                            \
    do {                                                                                                               \
        if (xp_f32) LAUNCH_PART_X(BS, LM, float);                                                                     \
        else LAUNCH_PART_X(BS, LM, double);                                                                            \
    } while (0)
#define LAUNCH_PART_X(BS, LM, XT)                                                                                      \
    do {                                                                                                               \
        const int cpb = 256 / LM;                                                                                      \
        const int grid = (nparts + cpb - 1) / cpb;                                                                     \
        const size_t lb = (size_t)cpb * ((LM * (LM - 1) / 2) * BS * BS + 1);                                          \
        const size_t lds = sizeof(double) * 2 * 256 * BS + (vt == JXAMD_T_F64 ? sizeof(double) : sizeof(float)) * lb; \
        if (vt == JXAMD_T_F64)                                                                                        \
            k_bhgs_part<BS, LM, double, double, PU64, XT><<<grid, 256, lds, st>>>(                                     \
                nb, nparts, cpb, sptr, pptr, col, (const double *)val, rsc, (const double *)dinv, rhs, xo,             \
                (const XT *)xp,                                                                                        \
                pidx, xn, w, omega);                                                                                   \
        else if (vt == JXAMD_T_F32)                                                                                    \
            k_bhgs_part<BS, LM, float, float, PU32, XT><<<grid, 256, lds, st>>>(                                       \
                nb, nparts, cpb, sptr, pptr, col, (const float *)val, rsc, (const float *)dinv, rhs, xo,               \
                (const XT *)xp,                                                                                        \
                pidx, xn, w, omega);                                                                                   \
        else                                                                                                           \
            k_bhgs_part<BS, LM, __half, float, PU16, XT><<<grid, 256, lds, st>>>(                                      \
                nb, nparts, cpb, sptr, pptr, col, (const __half *)val, rsc, (const float *)dinv, rhs, xo,              \
                (const XT *)xp,                                                                                        \
                pidx, xn, w, omega);                                                                                   \
    } while (0)
#define PART_B(BS)                                                                                                     \
    do {                                                                                                               \
        if (maxlen <= 2) LAUNCH_PART(BS, 2);                                                                           \
        else if (maxlen <= 4) LAUNCH_PART(BS, 4);                                                                      \
        else LAUNCH_PART(BS, 8);                                                                                       \
    } while (0)
    switch (b) {
    case 2: PART_B(2); break;
    case 3: PART_B(3); break;
    case 4: PART_B(4); break;
    default: return JXAMD_ERROR_ARG;
    }
#undef PART_B
#undef LAUNCH_PART
#undef LAUNCH_PART_X
    return chk(hipGetLastError());
}

int jxamd_k_btri(int b, int vt, int nsl, const int *perm, const int *sptr, const int *col, const void *val,
                 const float *rsc, const void *dinv, const double *rhs, double *out, const int *dep_ptr,
                 const int *dep_list, int *sflag, int epoch, unsigned long long *ticket, unsigned long long *base,
                 int *err, void *stream)
{
    if (nsl == 0) return 0;
    if (vt != JXAMD_T_F64 && vt != JXAMD_T_F32 && vt != JXAMD_T_F16) return JXAMD_ERROR_ARG;
    if ((vt == JXAMD_T_F16) != (rsc != nullptr)) return JXAMD_ERROR_ARG;
    if (b < 2 || b > 4) return JXAMD_ERROR_ARG;
    const int waves = jxamd_sweep_waves(nsl);
    const int grid = (waves + 3) / 4;
    hipStream_t st = (hipStream_t)stream;
    const unsigned long long b0 = *base;
#define LAUNCH_BTRI(VT, DT)                                                                                            \
    do {                                                                                                               \
        if (b == 2)                                                                                                    \
            k_btri<2, VT, DT><<<grid, 256, 0, st>>>(nsl, perm, sptr, col, (const VT *)val, rsc, (const DT *)dinv, rhs,  \
                                                    out, dep_ptr, dep_list, sflag, epoch, ticket, b0, err);            \
        else if (b == 3)                                                                                               \
            k_btri<3, VT, DT><<<grid, 256, 0, st>>>(nsl, perm, sptr, col, (const VT *)val, rsc, (const DT *)dinv, rhs,  \
                                                    out, dep_ptr, dep_list, sflag, epoch, ticket, b0, err);            \
        else                                                                                                           \
            k_btri<4, VT, DT><<<grid, 256, 0, st>>>(nsl, perm, sptr, col, (const VT *)val, rsc, (const DT *)dinv, rhs,  \
                                                    out, dep_ptr, dep_list, sflag, epoch, ticket, b0, err);            \
    } while (0)
    if (vt == JXAMD_T_F64) LAUNCH_BTRI(double, double);
    else if (vt == JXAMD_T_F32) LAUNCH_BTRI(float, float);
    else LAUNCH_BTRI(__half, float);
#undef LAUNCH_BTRI
    *base = b0 + (unsigned long long)nsl + (unsigned long long)grid * 4;
    return chk(hipGetLastError());
}

int jxamd_k_gs_chunk(int vt, int xt, int n, int nchunks, const int *wstep, const int *sptr, const int *col,
                     const void *val, const float *rsc, const void *diag, const void *f, int fs, const void *uo,
                     void *un, void *stream)
{
    if (n == 0 || nchunks < 1) return 0;
    const int waves = (nchunks + WAVE - 1) / WAVE;
    DISPATCH_VX(vt, xt,
                (k_gs_chunk<V_, X_><<<(waves + 3) / 4, 256, 0, (hipStream_t)stream>>>(
                    n, nchunks, wstep, sptr, col, (const V_ *)val, rsc, (const X_ *)diag, (const X_ *)f, fs,
                    (const X_ *)uo, (X_ *)un)));
    return chk(hipGetLastError());
}

int jxamd_gs_chunk_lds_rows(int n, int nchunks)
{
    if (nchunks < 1 || n < nchunks) return 0;
    const int size = n / nchunks;
    return size + (n % nchunks ? 1 : 0);
}

int jxamd_k_gs_chunk_lds(int vt, int xt, int n, int nchunks, const int *sptr, const int *col, const void *val,
                         const float *rsc, const void *diag, const void *f, int fs, const void *uo, void *un,
                         void *stream)
{
    if (n == 0) return 0;
    const int maxlen = jxamd_gs_chunk_lds_rows(n, nchunks);
    if (maxlen < 1 || maxlen > JXAMD_CHUNK_LDS_MAX) return JXAMD_ERROR_ARG;
    const int cpb = 256 / maxlen;
    const int grid = (nchunks + cpb - 1) / cpb;
    hipStream_t st = (hipStream_t)stream;
#define LAUNCH_LDS(LM)                                                                                                 \
    DISPATCH_VXP(vt, xt,                                                                                               \
                 (k_gs_chunk_lds<LM, V_, X_, P_><<<grid, 256, sizeof(X_) * (768 + (size_t)cpb * (LM * LM + 1)),      \
                                                   st>>>(n, nchunks, cpb, sptr, col, (const V_ *)val, rsc,             \
                                                         (const X_ *)diag, (const X_ *)f, fs, (const X_ *)uo,          \
                                                         (X_ *)un)))
    if (maxlen <= 4) LAUNCH_LDS(4);
    else if (maxlen <= 8) LAUNCH_LDS(8);
    else LAUNCH_LDS(16);
#undef LAUNCH_LDS
    return chk(hipGetLastError());
}

#define DISPATCH_G(G, CALL)                                                                                            \
    do {                                                                                                               \
        if ((G) == 4) { constexpr int G_ = 4; CALL; }                                                                  \
        else if ((G) == 8) { constexpr int G_ = 8; CALL; }                                                             \
        else if ((G) == 16) { constexpr int G_ = 16; CALL; }                                                           \
        else if ((G) == 32) { constexpr int G_ = 32; CALL; }                                                           \
        else return JXAMD_ERROR_ARG;                                                                                   \
    } while (0)

// grid of the SELL-G kernels: one wavefront per slice of 64/G rows
static int grp_grid(int n, int G)
{
    const int rows = WAVE / G;
    const int nsl = (n + rows - 1) / rows;
    return (nsl + 3) / 4;
}

int jxamd_k_csr_spmv_grp(int G, int vt, int xt, int n, const int *sptr, const int *col, const void *val,
                         const float *rsc, const void *x, int xs, const void *rhs, int rs, void *y, int mode,
                         void *stream)
{
    if (n == 0) return 0;
    if (G < 4 || G > 32) return JXAMD_ERROR_ARG;
    const int grid = grp_grid(n, G);
    DISPATCH_G(G, DISPATCH_VXP(vt, xt,
                               (k_csr_spmv_grp<G_, V_, X_, P_><<<grid, 256, 0, (hipStream_t)stream>>>(
                                   n, sptr, col, (const V_ *)val, rsc, (const X_ *)x, xs, (const X_ *)rhs, rs,
                                   (X_ *)y, mode, nullptr, nullptr, (X_)0))));
    return chk(hipGetLastError());
}

int jxamd_k_restrict_first(int G, int vt, int xt, int n, const int *sptr, const int *col, const void *val,
                           const float *rsc, const void *x, void *y, const void *dinv, void *u, double omega,
                           void *stream)
{
    if (n == 0) return 0;
    if (!dinv || !u || (G != 1 && (G < 4 || G > 32))) return JXAMD_ERROR_ARG;
    if (G == 1) {
        const int nsl = (n + WAVE - 1) / WAVE;
        const int grid = (nsl + 3) / 4;
        DISPATCH_VXP(vt, xt,
                     (k_csr_spmv<V_, X_, P_><<<grid, 256, 0, (hipStream_t)stream>>>(
                         n, sptr, col, (const V_ *)val, rsc, (const X_ *)x, 1, (const X_ *)nullptr, 1, (X_ *)y, 0,
                         (const X_ *)dinv, (X_ *)u, (X_)omega)));
    } else {
        const int grid = grp_grid(n, G);
        DISPATCH_G(G, DISPATCH_VXP(vt, xt,
                                   (k_csr_spmv_grp<G_, V_, X_, P_><<<grid, 256, 0, (hipStream_t)stream>>>(
                                       n, sptr, col, (const V_ *)val, rsc, (const X_ *)x, 1, (const X_ *)nullptr, 1,
                                       (X_ *)y, 0, (const X_ *)dinv, (X_ *)u, (X_)omega))));
    }
    return chk(hipGetLastError());
}

int jxamd_k_poly_step_grp(int G, int vt, int xt, int n, const int *sptr, const int *col, const void *val,
                          const float *rsc, const void *dinv, const void *f, int fs, const void *uo, void *un,
                          double omega, void *stream)
{
    if (n == 0) return 0;
    if (G < 4 || G > 32) return JXAMD_ERROR_ARG;
    const int grid = grp_grid(n, G);
    DISPATCH_G(G, DISPATCH_VXP(vt, xt,
                               (k_poly_step_grp<G_, V_, X_, P_><<<grid, 256, 0, (hipStream_t)stream>>>(
                                   n, sptr, col, (const V_ *)val, rsc, (const X_ *)dinv, (const X_ *)f, fs,
                                   (const X_ *)uo, (X_ *)un, (X_)omega))));
    return chk(hipGetLastError());
}

int jxamd_k_gs_sweep_grp(int G, int vt, int xt, int nsl, const int *perm, const int *ia, const int *ja,
                         const void *val, const float *rsc, const void *diag, const void *f, int fs, const void *uo,
                         void *un, const int *dep_ptr, const int *dep_list, int *sflag, int epoch, int dir,
                         unsigned long long *ticket, unsigned long long *base, int *err, void *stream)
{
    if (nsl == 0) return 0;
    if (G < 4 || G > 32) return JXAMD_ERROR_ARG;
    const int waves = jxamd_sweep_waves(nsl);
    const int grid = (waves + 3) / 4;
    hipStream_t st = (hipStream_t)stream;
    DISPATCH_G(G, DISPATCH_VX(vt, xt,
                              (k_gs_sweep_grp<G_, V_, X_><<<grid, 256, 0, st>>>(
                                  nsl, perm, ia, ja, (const V_ *)val, rsc, (const X_ *)diag, (const X_ *)f, fs,
                                  (const X_ *)uo, (X_ *)un, dep_ptr, dep_list, sflag, epoch, dir, ticket, *base,
                                  err))));
    *base += (unsigned long long)nsl + (unsigned long long)grid * 4;
    return chk(hipGetLastError());
}

int jxamd_k_gs_chunk_grp(int G, int vt, int xt, int n, int nchunks, const int *ia, const int *ja, const void *val,
                         const float *rsc, const void *diag, const void *f, int fs, const void *uo, void *un,
                         void *stream)
{
    if (n == 0) return 0;
    const int maxlen = jxamd_gs_chunk_lds_rows(n, nchunks);
    if (maxlen < 1 || maxlen > JXAMD_CHUNK_LDS_MAX) return JXAMD_ERROR_ARG;
    if (G < 1 || G > 32) return JXAMD_ERROR_ARG;
    // one row per lane group per block (two or four were slower on the
    // 60x220x85 hierarchy): long rows make few rows a block's worth of work,
    // and many blocks keep every CU busy
    const int cpb = max(1, (256 / G) / maxlen);
    const int grid = (nchunks + cpb - 1) / cpb;
    hipStream_t st = (hipStream_t)stream;
#define LAUNCH_GRP(LM)                                                                                                 \
    DISPATCH_G(G, DISPATCH_VX(vt, xt,                                                                                  \
                              (k_gs_chunk_grp<LM, G_, V_, X_>                                                         \
                               <<<grid, 256, sizeof(X_) * (768 + (size_t)cpb * (LM * LM + 1)), st>>>(                 \
                                   n, nchunks, cpb, ia, ja, (const V_ *)val, rsc, (const X_ *)diag, (const X_ *)f,    \
                                   fs, (const X_ *)uo, (X_ *)un))))
    if (maxlen <= 4) LAUNCH_GRP(4);
    else if (maxlen <= 8) LAUNCH_GRP(8);
    else LAUNCH_GRP(16);
#undef LAUNCH_GRP
    return chk(hipGetLastError());
}

int jxamd_k_dense_gemv(int xt, int n, int ld, const void *M, const void *x, void *y, void *stream)
{
    if (n == 0) return 0;
    if (xt == JXAMD_T_F64 && ld % 2 == 0 && ld >= n)
        k_dense_gemv<double><<<(n + 3) / 4, 256, 0, (hipStream_t)stream>>>(n, ld, (const double *)M,
                                                                            (const double *)x, (double *)y);
    else if (xt == JXAMD_T_F32 && ld % 4 == 0 && ld >= n)
        k_dense_gemv<float><<<(n + 3) / 4, 256, 0, (hipStream_t)stream>>>(n, ld, (const float *)M, (const float *)x,
                                                                           (float *)y);
    else return JXAMD_ERROR_ARG;
    return chk(hipGetLastError());
}

int jxamd_k_gather_cast(int n, const double *x, int stride, int off, float *y, void *stream)
{
    if (n == 0) return 0;
    k_gather_cast<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(n, x, stride, off, y);
    return chk(hipGetLastError());
}

int jxamd_k_cast_first(int n, const double *x, int stride, int off, float *y, const float *dinv, float *un,
                       double omega, void *stream)
{
    if (n == 0) return 0;
    k_cast_first<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(n, x, stride, off, y, dinv, un, (float)omega);
    return chk(hipGetLastError());
}
int jxamd_k_widen(int n, const float *x, double *y, void *stream)
{
    if (n == 0) return 0;
    k_widen<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(n, x, y);
    return chk(hipGetLastError());
}

int jxamd_k_fill_f32(int n, float *y, float v, void *stream)
{
    if (n == 0) return 0;
    k_fill_f32<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(n, y, v);
    return chk(hipGetLastError());
}

int jxamd_k_gather_stride(int n, const double *x, int stride, int off, double *y, void *stream)
{
    if (n == 0) return 0;
    k_gather_stride<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(n, x, stride, off, y);
    return chk(hipGetLastError());
}

int jxamd_k_prolong(int nb, int b, int pidx, const double *xp, double *x, void *stream)
{
    if (nb == 0) return 0;
    k_prolong<<<(nb + 255) / 256, 256, 0, (hipStream_t)stream>>>(nb, b, pidx, xp, x);
    return chk(hipGetLastError());
}

int jxamd_k_fill(size_t n, double *y, double v, void *stream)
{
    if (n == 0) return 0;
    k_fill<<<nblocks_stream(n, 256), 256, 0, (hipStream_t)stream>>>(n, y, v);
    return chk(hipGetLastError());
}

int jxamd_k_copy(size_t n, const double *x, double *y, void *stream)
{
    if (n == 0) return 0;
    k_copy<<<nblocks_stream(n, 256), 256, 0, (hipStream_t)stream>>>(n, x, y);
    return chk(hipGetLastError());
}

int jxamd_k_axpy(size_t n, const double *pa, double a, int neg, const double *x, double *y, void *stream)
{
    if (n == 0) return 0;
    k_axpy<<<nblocks_stream(n, 256), 256, 0, (hipStream_t)stream>>>(n, pa, a, neg, x, y);
    return chk(hipGetLastError());
}

int jxamd_k_scale(size_t n, const double *pnorm2, double a, double *y, void *stream)
{
    if (n == 0) return 0;
    k_scale<<<nblocks_stream(n, 256), 256, 0, (hipStream_t)stream>>>(n, pnorm2, a, y);
    return chk(hipGetLastError());
}

// Cooperative MGS plan: the largest resident grid of MGS_BS-thread
// workgroups for the smallest register tile TM in {4, 8, 16} that covers n;
// 0 if none does or the device has no cooperative launch.
int jxamd_mgs_coop_plan(size_t n, int coop_launch, int *grid, int *tm)
{
    *grid = 0;
    *tm = 0;
    int dev = 0, coop = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return JXAMD_ERROR_DEVICE;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess) return JXAMD_ERROR_DEVICE;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return JXAMD_ERROR_DEVICE;
    if ((coop_launch && !coop) || cus < 1) return 0;
    const void *fn[3] = {(const void *)k_mgs_coop<4>, (const void *)k_mgs_coop<8>, (const void *)k_mgs_coop<16>};
    const int tms[3] = {4, 8, 16};
    for (int q = 0; q < 3; q++) {
        int per = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn[q], MGS_BS, 0) != hipSuccess)
            return JXAMD_ERROR_DEVICE;
        if (per < 1) continue;
        const size_t g = (size_t)per * cus;
        if (g * MGS_BS * (size_t)tms[q] >= n) {
            // the fewest workgroups that still cover n with this tile
            size_t need = (n + MGS_BS * (size_t)tms[q] - 1) / (MGS_BS * (size_t)tms[q]);
            *grid = (int)(need < g ? need : g);
            *tm = tms[q];
            return 0;
        }
    }
    return 0;
}

int jxamd_k_mgs_coop(int grid, int tm, int coop_launch, size_t n, const double *V, size_t ld, int m, double *w,
                     double *hval, double *part, unsigned int *cnt, int *flag, int base, int *err, double *hhost,
                     int spin_limit, int hpos, double hseq, void *stream)
{
    if (grid < 1 || m < 0 || (size_t)grid * MGS_BS * (size_t)tm < n || spin_limit < 0) return JXAMD_ERROR_ARG;
    if (hpos >= 0 && (!hhost || hpos <= m)) return JXAMD_ERROR_ARG;
    void *args[] = {&n, &V, &ld, &m, &w, &hval, &part, &cnt, &flag, &base, &err, &hhost, &spin_limit, &hpos, &hseq};
    const void *fn;
    switch (tm) {
    case 4: fn = (const void *)k_mgs_coop<4>; break;
    case 8: fn = (const void *)k_mgs_coop<8>; break;
    case 16: fn = (const void *)k_mgs_coop<16>; break;
    default: return JXAMD_ERROR_ARG;
    }
    if (coop_launch) return chk(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(MGS_BS), args, 0, (hipStream_t)stream));
    // an ordinary launch of a grid no larger than the resident capacity (at
    // most one workgroup per CU at this size): the stream's previous kernel
    // has finished when it starts, so every workgroup is placed at once;
    // were one ever held back, the bounded waits end with err[0] set
    return chk(hipLaunchKernel(fn, dim3(grid), dim3(MGS_BS), args, 0, (hipStream_t)stream));
}

int jxamd_dot_grid(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 1024) b = 1024;
    return b < 1 ? 1 : (int)b;
}

int jxamd_k_dot_mgs(size_t n, const double *u, double *y, const double *v_prev, const double *h_prev, double *part,
                    unsigned int *count, double *out, double *hout, void *stream)
{
    k_dot_mgs<<<jxamd_dot_grid(n), 256, 0, (hipStream_t)stream>>>(n, u, y, v_prev, h_prev, part, count, out, hout);
    return chk(hipGetLastError());
}

int jxamd_k_lincomb(size_t n, int m, const double *V, size_t ld, const double *y, double *w, void *stream)
{
    if (n == 0 || m == 0) return 0;
    k_lincomb<<<nblocks_stream(n, 256), 256, 0, (hipStream_t)stream>>>(n, m, V, ld, y, w);
    return chk(hipGetLastError());
}

int jxamd_k_lincomb_host(size_t n, int m, const double *V, size_t ld, const double *y, int add, double *w,
                         void *stream)
{
    if (n == 0 || m == 0) return 0;
    if (m < 0 || m > JXAMD_COEFS_MAX || !y) return JXAMD_ERROR_ARG;
    jxamd_coefs c;
    for (int j = 0; j < JXAMD_COEFS_MAX; j++) c.c[j] = j < m ? y[j] : 0.0;
    if (add) k_lincomb_args<true><<<nblocks_stream(n, 256), 256, 0, (hipStream_t)stream>>>(n, m, V, ld, c, w);
    else k_lincomb_args<false><<<nblocks_stream(n, 256), 256, 0, (hipStream_t)stream>>>(n, m, V, ld, c, w);
    return chk(hipGetLastError());
}

} // extern "C"
