// jxamd_resnorm.hip -- true residual norm of a BSR system on the device.
//
// norms[0] = ||b - A x||_2, norms[1] = ||b||_2 for A in the reference's BSR
// layout (0-based ia[nb+1], ja[nnzb], row-major b x b blocks), in one launch:
// each thread owns scalar rows (i, r) on a grid-stride loop and accumulates
// the squares of r_(i,r) = b_(i,r) - sum_k A_k[r][:] x_(ja[k]), workgroup
// partials go to a scratch array, and the last workgroup to finish adds them
// in a fixed order, so the result does not depend on scheduling.
//
// This is the check GMRES makes on convergence (the reference copies b, runs
// its matvec with alpha = -1, beta = 1 and takes the square root of the inner
// product: jxfpamg-0220/csrc/krylov/pgmres.c:1505-1511; ||b|| as at
// pgmres.c:1062).
#include <hip/hip_runtime.h>

#include <mutex>

extern "C" {
#include "jxamd_dev.h"
}

#define RN_THREADS 256
#define RN_WAVE 64
// the grid never exceeds RN_MAX_GROUPS workgroups (grid-stride loop), so a
// work buffer of RN_SCRATCH_BYTES always holds every partial
#define RN_MAX_GROUPS 2048
// work buffer layout: double partials[2 * RN_MAX_GROUPS], then the arrival counter
#define RN_PART_BYTES (2 * RN_MAX_GROUPS * sizeof(double))
#define RN_SCRATCH_BYTES (RN_PART_BYTES + 256)

namespace {

template <int B>
__global__ void __launch_bounds__(RN_THREADS) k_bsr_resnorm(int nb, const int *__restrict__ ia,
                                                            const int *__restrict__ ja,
                                                            const double *__restrict__ a,
                                                            const double *__restrict__ x,
                                                            const double *__restrict__ bv, double *part,
                                                            unsigned int *count, double *norms)
{
    __shared__ double sr[RN_THREADS / RN_WAVE], sb[RN_THREADS / RN_WAVE];
    __shared__ bool last;
    const long long n = (long long)nb * B;
    double rr = 0.0, bb = 0.0;
    for (long long row = (long long)blockIdx.x * RN_THREADS + threadIdx.x; row < n;
         row += (long long)gridDim.x * RN_THREADS) {
        const int i = (int)(row / B), r = (int)(row % B);
        const double bi = bv[row];
        double s = bi;
        const int k1 = ia[i + 1];
        for (int k = ia[i]; k < k1; k++) {
            const double *blk = a + (size_t)k * (B * B) + r * B;
            const double *xj = x + (size_t)ja[k] * B;
#pragma unroll
            for (int c = 0; c < B; c++) s -= blk[c] * xj[c];
        }
        rr += s * s;
        bb += bi * bi;
    }
#pragma unroll
    for (int off = RN_WAVE / 2; off > 0; off >>= 1) {
        rr += __shfl_xor(rr, off);
        bb += __shfl_xor(bb, off);
    }
    if (threadIdx.x % RN_WAVE == 0) {
        sr[threadIdx.x / RN_WAVE] = rr;
        sb[threadIdx.x / RN_WAVE] = bb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r0 = 0.0, b0 = 0.0;
        for (int w = 0; w < RN_THREADS / RN_WAVE; w++) {
            r0 += sr[w];
            b0 += sb[w];
        }
        // device-scope stores of the partials, completed before the arrival
        // is counted; the last arrival reads them with device-scope loads
        __hip_atomic_store(part + 2 * blockIdx.x, r0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(part + 2 * blockIdx.x + 1, b0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
        const unsigned int prev = __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    // fixed-order sum over the partials: thread t adds partials t, t + 256, ...
    rr = 0.0;
    bb = 0.0;
    for (unsigned int g = threadIdx.x; g < gridDim.x; g += RN_THREADS) {
        rr += __hip_atomic_load(part + 2 * g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bb += __hip_atomic_load(part + 2 * g + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int off = RN_WAVE / 2; off > 0; off >>= 1) {
        rr += __shfl_xor(rr, off);
        bb += __shfl_xor(bb, off);
    }
    __syncthreads();
    if (threadIdx.x % RN_WAVE == 0) {
        sr[threadIdx.x / RN_WAVE] = rr;
        sb[threadIdx.x / RN_WAVE] = bb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r0 = 0.0, b0 = 0.0;
        for (int w = 0; w < RN_THREADS / RN_WAVE; w++) {
            r0 += sr[w];
            b0 += sb[w];
        }
        norms[0] = sqrt(r0);
        norms[1] = sqrt(b0);
        __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void k_rn_clear(unsigned int *count) { *count = 0u; }

int rn_check(int nb, int b, const int *ia, const int *ja, const double *a, const double *x, const double *bv,
             double *norms)
{
    if (nb < 0 || b < 1 || b > JXAMD_MAX_BLOCK || !norms) return JXAMD_ERROR_ARG;
    if (nb > 0 && (!ia || !ja || !a || !x || !bv)) return JXAMD_ERROR_ARG;
    // scalar rows are counted in 64-bit, block rows and block positions in int32
    if ((long long)nb * b > (long long)0x7fffffff) return JXAMD_ERROR_ARG;
    return 0;
}

int rn_grid(int nb, int b)
{
    const long long n = (long long)nb * b;
    long long g = (n + RN_THREADS - 1) / RN_THREADS;
    if (g < 1) g = 1;
    if (g > RN_MAX_GROUPS) g = RN_MAX_GROUPS;
    return (int)g;
}

int rn_launch(int nb, int b, const int *ia, const int *ja, const double *a, const double *x, const double *bv,
              double *norms, double *part, unsigned int *count, hipStream_t st)
{
    const int grid = rn_grid(nb, b);
    if (grid > RN_MAX_GROUPS) return JXAMD_ERROR_ARG;
    hipLaunchKernelGGL(k_rn_clear, dim3(1), dim3(1), 0, st, count);
    switch (b) {
    case 1: hipLaunchKernelGGL(k_bsr_resnorm<1>, dim3(grid), dim3(RN_THREADS), 0, st, nb, ia, ja, a, x, bv, part, count, norms); break;
    case 2: hipLaunchKernelGGL(k_bsr_resnorm<2>, dim3(grid), dim3(RN_THREADS), 0, st, nb, ia, ja, a, x, bv, part, count, norms); break;
    case 3: hipLaunchKernelGGL(k_bsr_resnorm<3>, dim3(grid), dim3(RN_THREADS), 0, st, nb, ia, ja, a, x, bv, part, count, norms); break;
    default: hipLaunchKernelGGL(k_bsr_resnorm<4>, dim3(grid), dim3(RN_THREADS), 0, st, nb, ia, ja, a, x, bv, part, count, norms); break;
    }
    return hipGetLastError() == hipSuccess ? 0 : JXAMD_ERROR_DEVICE;
}

// Scratch of the allocation-free entry point.  In host code `scratch` names
// a host-side stand-in, not device memory: only the address returned by
// hipGetSymbolAddress (one per device) may be handed to kernels.
static __device__ __align__(256) unsigned char scratch[RN_SCRATCH_BYTES];

#define RN_MAX_DEVICES 64
std::mutex rn_mutex;
void *rn_scratch_dev[RN_MAX_DEVICES];

// this device's copy of the scratch array, looked up once per device
int rn_scratch(void **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return JXAMD_ERROR_DEVICE;
    if (dev < 0 || dev >= RN_MAX_DEVICES) return JXAMD_ERROR_DEVICE;
    std::lock_guard<std::mutex> g(rn_mutex);
    if (!rn_scratch_dev[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return JXAMD_ERROR_DEVICE;
        rn_scratch_dev[dev] = p;
    }
    *out = rn_scratch_dev[dev];
    return 0;
}

} // namespace

extern "C" {

int jxamd_bsr_residual_norm(int nb, int b, const int *d_ia, const int *d_ja, const double *d_a, const double *d_x,
                            const double *d_b, double *d_norms, void *stream)
{
    int e = rn_check(nb, b, d_ia, d_ja, d_a, d_x, d_b, d_norms);
    if (e) return e;
    hipStream_t st = (hipStream_t)stream;
    void *work = nullptr;
    if (hipMalloc(&work, RN_SCRATCH_BYTES) != hipSuccess) return JXAMD_ERROR_MEMORY;
    e = rn_launch(nb, b, d_ia, d_ja, d_a, d_x, d_b, d_norms, (double *)work,
                  (unsigned int *)((unsigned char *)work + RN_PART_BYTES), st);
    // the launch must finish before its partials are released
    if (hipStreamSynchronize(st) != hipSuccess && !e) e = JXAMD_ERROR_DEVICE;
    if (hipFree(work) != hipSuccess && !e) e = JXAMD_ERROR_DEVICE;
    return e;
}

int jxamd_bsr_residual_norm_noalloc(int nb, int b, const int *d_ia, const int *d_ja, const double *d_a,
                                    const double *d_x, const double *d_b, double *d_norms, void *stream)
{
    int e = rn_check(nb, b, d_ia, d_ja, d_a, d_x, d_b, d_norms);
    if (e) return e;
    void *work = nullptr;
    e = rn_scratch(&work);
    if (e) return e;
    return rn_launch(nb, b, d_ia, d_ja, d_a, d_x, d_b, d_norms, (double *)work,
                     (unsigned int *)((unsigned char *)work + RN_PART_BYTES), (hipStream_t)stream);
}

} // extern "C"
