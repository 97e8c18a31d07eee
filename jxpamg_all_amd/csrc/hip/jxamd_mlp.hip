// jxamd_mlp.hip -- the learned part of the polynomial AMG smoother: a 3-layer
// MLP (16 level features -> 64 -> 64 -> 4 polynomial step sizes, tanh hidden
// units) with BF16 operands on the CDNA matrix cores (v_mfma_f32_16x16x16_bf16,
// FP32 accumulation), FP32 master weights and Adam.  No reference counterpart.
//
// Parameters (FP32, JXAMD_MLP_NPARAM floats): W1[64][16] b1[64] W2[64][64]
// b2[64] W3[4][64] b3[4], rows = output units.
//
// Both kernels run one 256-thread workgroup (4 wavefronts) per batch of 32
// samples; every matrix product is tiled into 16x16 output tiles, one
// wavefront per tile, K in steps of 16.  Operands are read as FP32 (LDS for
// activations, global memory for the small weight set) and rounded to BF16
// (round to nearest even) as they are packed for the MFMA.
#include <hip/hip_runtime.h>
#include <stdint.h>

extern "C" {
#include "jxamd_dev.h"
}

#define WAVE 64
#define MB 32 // samples per batch

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

enum { OW1 = 0, OB1 = 1024, OW2 = 1088, OB2 = 5184, OW3 = 5248, OB3 = 5504, NPAR = 5508 };

__device__ __forceinline__ short to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// C[m][n] = sum_k A[m][k] B[k][n] over the real extents (Mr, Nr, Kr); A(m, k)
// at A[m*sam + k*sak], B(k, n) at B[k*sbk + n*sbn], C(m, n) at C[m*scm + n*scn].
// Padded extents are multiples of 16; out-of-range operands read as zero and
// out-of-range results are not stored.  Ends with a block barrier.
__device__ void gemm_bf16(int Mr, int Nr, int Kr, const float *A, int sam, int sak, const float *B, int sbk,
                          int sbn, float *C, int scm, int scn)
{
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int Mt = (Mr + 15) / 16, Nt = (Nr + 15) / 16, Kp = (Kr + 15) / 16 * 16;
    const int r16 = lane % 16, q = lane / 16;
    for (int t = wave; t < Mt * Nt; t += blockDim.x / WAVE) {
        const int m = (t / Nt) * 16 + r16; // operand A row of this lane
        const int n = (t % Nt) * 16 + r16; // operand B column of this lane
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kb = 0; kb < Kp; kb += 16) {
            bf16x4 a, b;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int k = kb + 4 * q + i;
                a[i] = (m < Mr && k < Kr) ? to_bf16(A[m * sam + k * sak]) : (short)0;
                b[i] = (n < Nr && k < Kr) ? to_bf16(B[k * sbk + n * sbn]) : (short)0;
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int mo = (t / Nt) * 16 + 4 * q + i, no = (t % Nt) * 16 + r16;
            if (mo < Mr && no < Nr) C[mo * scm + no * scn] = acc[i];
        }
    }
    __syncthreads();
}

struct Acts {
    float X[MB * 16], H1[MB * 64], H2[MB * 64], Y[MB * 4], S[MB * 64], D[MB * 64];
};

// forward pass of nb (<= MB) samples already in a.X; results in a.H1, a.H2, a.Y
__device__ void forward(const float *P, Acts &a, int nb)
{
    gemm_bf16(nb, 64, 16, a.X, 16, 1, P + OW1, 1, 16, a.H1, 64, 1);
    for (int e = threadIdx.x; e < nb * 64; e += blockDim.x) a.H1[e] = tanhf(a.H1[e] + P[OB1 + e % 64]);
    __syncthreads();
    gemm_bf16(nb, 64, 64, a.H1, 64, 1, P + OW2, 1, 64, a.H2, 64, 1);
    for (int e = threadIdx.x; e < nb * 64; e += blockDim.x) a.H2[e] = tanhf(a.H2[e] + P[OB2 + e % 64]);
    __syncthreads();
    gemm_bf16(nb, 4, 64, a.H2, 64, 1, P + OW3, 1, 64, a.Y, 4, 1);
    for (int e = threadIdx.x; e < nb * 4; e += blockDim.x) a.Y[e] += P[OB3 + e % 4];
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_mlp_forward(const float *__restrict__ P, const float *__restrict__ x, int n,
                                                     float *__restrict__ y)
{
    __shared__ Acts a;
    const int b0 = blockIdx.x * MB;
    const int nb = min(MB, n - b0);
    for (int e = threadIdx.x; e < MB * 16; e += blockDim.x) a.X[e] = (e / 16 < nb) ? x[(size_t)b0 * 16 + e] : 0.f;
    __syncthreads();
    forward(P, a, nb);
    for (int e = threadIdx.x; e < nb * 4; e += blockDim.x) y[(size_t)b0 * 4 + e] = a.Y[e];
}

// steps Adam steps on the mean squared error of the first nout outputs over
// mini-batches of MB samples taken cyclically from (x, t[n][nout]); step s
// uses samples (t0 + s) * MB + r mod n.  The other outputs get no gradient.
// G receives the gradient of the last step, loss[s] the loss before step s.
__global__ void __launch_bounds__(256) k_mlp_train(float *__restrict__ P, float *__restrict__ M, float *__restrict__ V,
                                                   float *__restrict__ G, const float *__restrict__ x,
                                                   const float *__restrict__ t, int n, int nout, int steps, int t0,
                                                   float lr, float beta1, float beta2, float eps, int update,
                                                   float *__restrict__ loss)
{
    __shared__ Acts a;
    __shared__ float red[256];
    const int nb = min(MB, n);
    for (int s = 0; s < steps; s++) {
        const long base = (long)(t0 + s) * MB;
        for (int e = threadIdx.x; e < MB * 16; e += blockDim.x) {
            const int r = e / 16;
            a.X[e] = r < nb ? x[(size_t)((base + r) % n) * 16 + e % 16] : 0.f;
        }
        __syncthreads();
        forward(P, a, nb);
        // dY = 2 (y - t) / nb into a.S (nb x 4; zero beyond nout), loss
        float part = 0.f;
        for (int e = threadIdx.x; e < nb * 4; e += blockDim.x) {
            const int r = e / 4, o = e % 4;
            const float d = o < nout ? a.Y[e] - t[(size_t)((base + r) % n) * nout + o] : 0.f;
            part += d * d;
            a.S[e] = 2.f * d / (float)nb;
        }
        red[threadIdx.x] = part;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0 && loss) loss[s] = red[0] / (float)nb;
        // dW3[k][j] = sum_b dY[b][k] H2[b][j];  db3
        gemm_bf16(4, 64, nb, a.S, 1, 4, a.H2, 64, 1, G + OW3, 64, 1);
        for (int k = threadIdx.x; k < 4; k += blockDim.x) {
            float acc = 0.f;
            for (int b = 0; b < nb; b++) acc += a.S[b * 4 + k];
            G[OB3 + k] = acc;
        }
        // dH2 = dY W3 -> D;  dZ2 = dH2 (1 - H2^2) -> D
        gemm_bf16(nb, 64, 4, a.S, 4, 1, P + OW3, 64, 1, a.D, 64, 1);
        for (int e = threadIdx.x; e < nb * 64; e += blockDim.x) a.D[e] *= 1.f - a.H2[e] * a.H2[e];
        __syncthreads();
        // dW2[o][j] = sum_b dZ2[b][o] H1[b][j];  db2
        gemm_bf16(64, 64, nb, a.D, 1, 64, a.H1, 64, 1, G + OW2, 64, 1);
        for (int o = threadIdx.x; o < 64; o += blockDim.x) {
            float acc = 0.f;
            for (int b = 0; b < nb; b++) acc += a.D[b * 64 + o];
            G[OB2 + o] = acc;
        }
        // dH1 = dZ2 W2 -> S;  dZ1 = dH1 (1 - H1^2) -> S
        gemm_bf16(nb, 64, 64, a.D, 64, 1, P + OW2, 64, 1, a.S, 64, 1);
        for (int e = threadIdx.x; e < nb * 64; e += blockDim.x) a.S[e] *= 1.f - a.H1[e] * a.H1[e];
        __syncthreads();
        // dW1[o][i] = sum_b dZ1[b][o] X[b][i];  db1
        gemm_bf16(64, 16, nb, a.S, 1, 64, a.X, 16, 1, G + OW1, 16, 1);
        for (int o = threadIdx.x; o < 64; o += blockDim.x) {
            float acc = 0.f;
            for (int b = 0; b < nb; b++) acc += a.S[b * 64 + o];
            G[OB1 + o] = acc;
        }
        __syncthreads();
        if (update) {
            const int it = t0 + s + 1;
            const float c1 = 1.f / (1.f - powf(beta1, (float)it)), c2 = 1.f / (1.f - powf(beta2, (float)it));
            for (int e = threadIdx.x; e < NPAR; e += blockDim.x) {
                const float g = G[e];
                const float m = beta1 * M[e] + (1.f - beta1) * g;
                const float v = beta2 * V[e] + (1.f - beta2) * g * g;
                M[e] = m;
                V[e] = v;
                P[e] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
            }
        }
        __syncthreads();
    }
}

int chk(hipError_t e) { return e == hipSuccess ? 0 : JXAMD_ERROR_DEVICE; }

} // namespace

extern "C" {

int jxamd_mlp_forward(const float *d_params, const float *d_x, int n, float *d_y, void *stream)
{
    if (n < 0 || !d_params) return JXAMD_ERROR_ARG;
    if (n == 0) return 0;
    k_mlp_forward<<<(n + MB - 1) / MB, 256, 0, (hipStream_t)stream>>>(d_params, d_x, n, d_y);
    return chk(hipGetLastError());
}

int jxamd_mlp_train(float *d_params, float *d_m, float *d_v, float *d_grad, const float *d_x, const float *d_t, int n,
                    int nout, int steps, int t0, float lr, int update, float *d_loss, void *stream)
{
    if (n < 1 || nout < 1 || nout > JXAMD_MLP_NOUT || steps < 1 || t0 < 0 || !d_params || !d_grad ||
        (update && (!d_m || !d_v)))
        return JXAMD_ERROR_ARG;
    k_mlp_train<<<1, 256, 0, (hipStream_t)stream>>>(d_params, d_m, d_v, d_grad, d_x, d_t, n, nout, steps, t0, lr, 0.9f,
                                                    0.999f, 1e-8f, update, d_loss);
    return chk(hipGetLastError());
}

} // extern "C"
