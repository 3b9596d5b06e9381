// FP32 sizing predictor: MLP 12 -> 64 -> 64 -> 3 (ReLU), trained on (path features ->
// log1p(chord count), log1p(fill count), log1p(alpha tile count)) triples that the
// flatten and bin count passes emit.  It
// replaces the reference's "count, read back, grow to 2^k, retry" sizing
// (core/d3d11/renderer.cpp:537-577, 707-724, 834-845) for batched launches; an
// under-prediction costs a retry, never a wrong pixel.  Tiny matrices: plain FP32
// VALU kernels, no MFMA.
//
// Parameter layout (5187 floats): W1[12][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3]
// (row-major, W[in][out]).  Hidden activations per sample: h1[64], h2[64].
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pfh {

constexpr int NI = 12, NH = 64, NO = 3;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NI * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NO, N_PARAMS = OFF_B3 + NO;

int64_t mlp_param_count() { return N_PARAMS; }

__global__ void k_zero(float *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}


__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ P, const float *__restrict__ x,
                                                     int64_t n, float *__restrict__ hidden, float *__restrict__ out) {
    __shared__ float w[N_PARAMS];
    for (int i = threadIdx.x; i < N_PARAMS; i += blockDim.x) w[i] = P[i];
    __syncthreads();
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float xi[NI];
#pragma unroll
    for (int i = 0; i < NI; i++) xi[i] = x[s * NI + i];
    float h1[NH];
#pragma unroll 4
    for (int j = 0; j < NH; j++) {
        float a = w[OFF_B1 + j];
#pragma unroll
        for (int i = 0; i < NI; i++) a += xi[i] * w[OFF_W1 + i * NH + j];
        h1[j] = fmaxf(a, 0.0f);
    }
    float o[NO];
#pragma unroll
    for (int q = 0; q < NO; q++) o[q] = w[OFF_B3 + q];
    for (int j = 0; j < NH; j++) {
        float a = w[OFF_B2 + j];
#pragma unroll 8
        for (int i = 0; i < NH; i++) a += h1[i] * w[OFF_W2 + i * NH + j];
        float h2 = fmaxf(a, 0.0f);
        hidden[s * 2 * NH + NH + j] = h2;
#pragma unroll
        for (int q = 0; q < NO; q++) o[q] += h2 * w[OFF_W3 + j * NO + q];
    }
#pragma unroll 8
    for (int j = 0; j < NH; j++) hidden[s * 2 * NH + j] = h1[j];
#pragma unroll
    for (int q = 0; q < NO; q++) out[s * NO + q] = o[q];
}

// Per-sample output and hidden gradients of the mean squared error
// L = 1/(n*NO) * sum (out - target)^2; writes g[s] = (dout[NO], dh2[64], dh1[64]).
__global__ __launch_bounds__(256) void k_mlp_backward_act(const float *__restrict__ P, int64_t n,
                                                          const float *__restrict__ hidden,
                                                          const float *__restrict__ out,
                                                          const float *__restrict__ target, float *__restrict__ g,
                                                          float *__restrict__ loss) {
    __shared__ float w2[NH * NH];
    __shared__ float w3[NH * NO];
    __shared__ float red[4];
    for (int i = threadIdx.x; i < NH * NH; i += blockDim.x) w2[i] = P[OFF_W2 + i];
    for (int i = threadIdx.x; i < NH * NO; i += blockDim.x) w3[i] = P[OFF_W3 + i];
    __syncthreads();
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (s < n) {
        const float scale = 2.0f / (float)(n * NO);
        float d[NO];
        float *gs = g + s * (NO + 2 * NH);
#pragma unroll
        for (int q = 0; q < NO; q++) {
            float e = out[s * NO + q] - target[s * NO + q];
            l += e * e;
            d[q] = scale * e;
            gs[q] = d[q];
        }
        const float *h = hidden + s * 2 * NH;
        float dh2[NH];
#pragma unroll 8
        for (int j = 0; j < NH; j++) {
            float acc = 0.0f;
#pragma unroll
            for (int q = 0; q < NO; q++) acc += d[q] * w3[j * NO + q];
            float v = h[NH + j] > 0.0f ? acc : 0.0f;
            dh2[j] = v;
            gs[NO + j] = v;
        }
        for (int i = 0; i < NH; i++) {
            float a = 0.0f;
#pragma unroll 8
            for (int j = 0; j < NH; j++) a += w2[i * NH + j] * dh2[j];
            gs[NO + NH + i] = h[i] > 0.0f ? a : 0.0f;
        }
    }
    // block reduction of the loss
    for (int d = 32; d > 0; d >>= 1) l += __shfl_down(l, d, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (red[0] + red[1] + red[2] + red[3]) / (float)(n * NO));
}

// dW[ka][kb] += sum_s A[s][:ka]^T B[s][:kb]; db[kb] += sum_s B[s][:kb].  A block reduces
// OUTER_SAMPLES samples in LDS tiles of OUTER_TILE, keeps its outputs in registers and
// adds them to the global gradient once.
constexpr int OUTER_TILE = 64;
constexpr int OUTER_SAMPLES = 4096;
constexpr int OUTER_PER_THREAD = (NH * NH + NH + 255) / 256;  // 17
__global__ __launch_bounds__(256) void k_mlp_outer(const float *__restrict__ A, int lda, int ka,
                                                   const float *__restrict__ B, int ldb, int boff, int kb, int64_t n,
                                                   float *__restrict__ dW, float *__restrict__ db) {
    __shared__ float sa[OUTER_TILE][NH + 1];
    __shared__ float sb[OUTER_TILE][NH + 1];
    float acc[OUTER_PER_THREAD];
#pragma unroll
    for (int k = 0; k < OUTER_PER_THREAD; k++) acc[k] = 0.0f;
    const int total = ka * kb;
    int64_t first = (int64_t)blockIdx.x * OUTER_SAMPLES;
    int64_t last = min<int64_t>(n, first + OUTER_SAMPLES);
    for (int64_t s0 = first; s0 < last; s0 += OUTER_TILE) {
        int cnt = (int)min<int64_t>(OUTER_TILE, last - s0);
        __syncthreads();
        for (int i = threadIdx.x; i < cnt * ka; i += blockDim.x) sa[i / ka][i % ka] = A[(s0 + i / ka) * lda + i % ka];
        for (int i = threadIdx.x; i < cnt * kb; i += blockDim.x)
            sb[i / kb][i % kb] = B[(s0 + i / kb) * ldb + boff + i % kb];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < OUTER_PER_THREAD; k++) {
            int o = threadIdx.x + k * 256;
            if (o < total) {
                int i = o / kb, j = o % kb;
                float a = 0.0f;
                for (int s = 0; s < cnt; s++) a += sa[s][i] * sb[s][j];
                acc[k] += a;
            } else if (o < total + kb) {
                int j = o - total;
                float a = 0.0f;
                for (int s = 0; s < cnt; s++) a += sb[s][j];
                acc[k] += a;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < OUTER_PER_THREAD; k++) {
        int o = threadIdx.x + k * 256;
        if (o < total) atomicAdd(&dW[o], acc[k]);
        else if (o < total + kb) atomicAdd(&db[o - total], acc[k]);
    }
}


// Sizing inference: per path forward pass, expm1 of the (log1p) outputs clamped at 0,
// summed over the paths into sums[NO] (float atomics per block).
__global__ __launch_bounds__(256) void k_mlp_predict_sum(const float *__restrict__ P, const float *__restrict__ x,
                                                         int64_t n, float *__restrict__ sums) {
    __shared__ float w[N_PARAMS];
    __shared__ float red[4][NO];
    for (int i = threadIdx.x; i < N_PARAMS; i += blockDim.x) w[i] = P[i];
    __syncthreads();
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float o[NO];
#pragma unroll
    for (int q = 0; q < NO; q++) o[q] = 0.0f;
    if (s < n) {
        float xi[NI];
#pragma unroll
        for (int i = 0; i < NI; i++) xi[i] = x[s * NI + i];
        float h1[NH];
#pragma unroll 4
        for (int j = 0; j < NH; j++) {
            float a = w[OFF_B1 + j];
#pragma unroll
            for (int i = 0; i < NI; i++) a += xi[i] * w[OFF_W1 + i * NH + j];
            h1[j] = fmaxf(a, 0.0f);
        }
#pragma unroll
        for (int q = 0; q < NO; q++) o[q] = w[OFF_B3 + q];
        for (int j = 0; j < NH; j++) {
            float a = w[OFF_B2 + j];
#pragma unroll 8
            for (int i = 0; i < NH; i++) a += h1[i] * w[OFF_W2 + i * NH + j];
            const float h2 = fmaxf(a, 0.0f);
#pragma unroll
            for (int q = 0; q < NO; q++) o[q] += h2 * w[OFF_W3 + j * NO + q];
        }
#pragma unroll
        for (int q = 0; q < NO; q++) o[q] = expm1f(fmaxf(o[q], 0.0f));
    }
#pragma unroll
    for (int q = 0; q < NO; q++) {
        float v = o[q];
        for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < NO) atomicAdd(&sums[threadIdx.x], red[0][threadIdx.x] + red[1][threadIdx.x] +
                                                            red[2][threadIdx.x] + red[3][threadIdx.x]);
}

hipError_t mlp_predict_sum(const float *params, const float *x, int64_t n, float *sums, hipStream_t s) {
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(32), 0, s, sums, (int64_t)NO);
    if (n > 0)
        hipLaunchKernelGGL(k_mlp_predict_sum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, params, x, n, sums);
    return hipGetLastError();
}

hipError_t mlp_forward(const float *params, const float *x, int64_t n, float *hidden, float *out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, params, x, n, hidden, out);
    return hipGetLastError();
}

// grad: N_PARAMS floats (overwritten); loss: 1 float (overwritten); scratch: n * (NO + 2 NH) floats.
hipError_t mlp_backward(const float *params, const float *x, int64_t n, const float *hidden, const float *out,
                        const float *target, float *grad, float *loss, float *scratch, hipStream_t s) {
    hipLaunchKernelGGL(k_zero, dim3((N_PARAMS + 255) / 256), dim3(256), 0, s, grad, (int64_t)N_PARAMS);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(32), 0, s, loss, (int64_t)1);
    if (n == 0) return hipGetLastError();
    unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_mlp_backward_act, dim3(nb), dim3(256), 0, s, params, n, hidden, out, target, scratch, loss);
    unsigned nc = (unsigned)((n + OUTER_SAMPLES - 1) / OUTER_SAMPLES);
    const int LG = NO + 2 * NH;
    // layer 3: A = h2, B = dout
    hipLaunchKernelGGL(k_mlp_outer, dim3(nc), dim3(256), 0, s, hidden + NH, 2 * NH, NH, scratch, LG, 0, NO, n,
                       grad + OFF_W3, grad + OFF_B3);
    // layer 2: A = h1, B = dh2
    hipLaunchKernelGGL(k_mlp_outer, dim3(nc), dim3(256), 0, s, hidden, 2 * NH, NH, scratch, LG, NO, NH, n,
                       grad + OFF_W2, grad + OFF_B2);
    // layer 1: A = x, B = dh1
    hipLaunchKernelGGL(k_mlp_outer, dim3(nc), dim3(256), 0, s, x, NI, NI, scratch, LG, NO + NH, NH, n, grad + OFF_W1,
                       grad + OFF_B1);
    return hipGetLastError();
}

__global__ void k_adam(float *p, float *m, float *v, const float *g, int64_t n, float lr, float c1, float c2) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float gi = g[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

hipError_t mlp_adam(float *params, float *m, float *v, const float *grad, int64_t n, float lr, int step,
                    hipStream_t s) {
    float c1 = 1.0f - powf(0.9f, (float)step), c2 = 1.0f - powf(0.999f, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, params, m, v, grad, n, lr, c1, c2);
    return hipGetLastError();
}

}  // namespace pfh
