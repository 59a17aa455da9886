// The MLP of the learned block-Jacobi correction (no reference counterpart): per retained block, 26 FP32 features
// -> 64 -> 64 (ReLU) -> 1, trained and applied on the device.
//
// Parameters are one flat FP32 vector: W1 [NF][NH], b1 [NH], W2 [NH][NH], b2 [NH], W3 [NH], b3. Inputs are
// normalised by (x - mean) / std with statistics kept beside the weights.
//   k_mlp_forward    one thread per sample, weights staged in LDS; optionally keeps the activations and forms the
//                    backward pass of the loss L = mean over samples of (o - y)^2 / 2 (dL/do, dL/da2, dL/da1)
//   k_mlp_wgrad      one thread per parameter: the sum over samples of its gradient terms, in sample order
//                    (deterministic, no atomics)
//   k_mlp_adam       one thread per parameter: Adam (Kingma & Ba, 2015) with bias correction
#include <cmath>

#include "bls_internal.h"

using namespace bls;

namespace {

constexpr int NF = 26, NH = 64;
constexpr int O_W1 = 0, O_B1 = O_W1 + NF * NH, O_W2 = O_B1 + NH, O_B2 = O_W2 + NH * NH, O_W3 = O_B2 + NH,
              O_B3 = O_W3 + NH, NP = O_B3 + 1;
constexpr int WORK_PER_SAMPLE = 4 * NH + 1;   // h1, h2, d1, d2, dO

struct Work {
    float *h1, *h2, *d1, *d2, *dO;
};

__device__ inline Work work_of(float *w, int64_t n) {
    return Work{w, w + n * NH, w + 2 * n * NH, w + 3 * n * NH, w + 4 * n * NH};
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float *P, const float *fmean, const float *fstd,
                                                      const float *X, const float *y, int64_t n, float *out,
                                                      float *work) {
    __shared__ float w[NP];
    for (int q = threadIdx.x; q < NP; q += blockDim.x) w[q] = P[q];
    __syncthreads();
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float xn[NF], h1[NH], h2[NH];
#pragma unroll
    for (int f = 0; f < NF; f++) xn[f] = (X[s * NF + f] - fmean[f]) / fstd[f];
#pragma unroll
    for (int h = 0; h < NH; h++) {
        float acc = w[O_B1 + h];
#pragma unroll
        for (int f = 0; f < NF; f++) acc += xn[f] * w[O_W1 + f * NH + h];
        h1[h] = acc > 0.0f ? acc : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < NH; k++) {
        float acc = w[O_B2 + k];
#pragma unroll
        for (int h = 0; h < NH; h++) acc += h1[h] * w[O_W2 + h * NH + k];
        h2[k] = acc > 0.0f ? acc : 0.0f;
    }
    float o = w[O_B3];
#pragma unroll
    for (int k = 0; k < NH; k++) o += h2[k] * w[O_W3 + k];
    if (out) out[s] = o;
    if (!work) return;
    const Work W = work_of(work, n);
    const float dO = (o - y[s]) / (float)n;
    float d2[NH];
#pragma unroll
    for (int k = 0; k < NH; k++) {
        d2[k] = h2[k] > 0.0f ? dO * w[O_W3 + k] : 0.0f;
        W.h2[s * NH + k] = h2[k];
        W.d2[s * NH + k] = d2[k];
    }
#pragma unroll
    for (int h = 0; h < NH; h++) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < NH; k++) acc += w[O_W2 + h * NH + k] * d2[k];
        W.h1[s * NH + h] = h1[h];
        W.d1[s * NH + h] = h1[h] > 0.0f ? acc : 0.0f;
    }
    W.dO[s] = dO;
}

__global__ __launch_bounds__(256) void k_mlp_wgrad(const float *fmean, const float *fstd, const float *X, int64_t n,
                                                    const float *work, float *G) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NP) return;
    const Work W = work_of(const_cast<float *>(work), n);
    float g = 0.0f;
    if (q < O_B1) {
        const int f = q / NH, h = q % NH;
        const float mu = fmean[f], sd = fstd[f];
        for (int64_t s = 0; s < n; s++) g += ((X[s * NF + f] - mu) / sd) * W.d1[s * NH + h];
    } else if (q < O_W2) {
        const int h = q - O_B1;
        for (int64_t s = 0; s < n; s++) g += W.d1[s * NH + h];
    } else if (q < O_B2) {
        const int h = (q - O_W2) / NH, k = (q - O_W2) % NH;
        for (int64_t s = 0; s < n; s++) g += W.h1[s * NH + h] * W.d2[s * NH + k];
    } else if (q < O_W3) {
        const int k = q - O_B2;
        for (int64_t s = 0; s < n; s++) g += W.d2[s * NH + k];
    } else if (q < O_B3) {
        const int k = q - O_W3;
        for (int64_t s = 0; s < n; s++) g += W.h2[s * NH + k] * W.dO[s];
    } else {
        for (int64_t s = 0; s < n; s++) g += W.dO[s];
    }
    G[q] = g;
}

__global__ void k_mlp_adam(float *P, const float *G, float *m, float *v, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NP) return;
    const float g = G[q];
    const float mq = b1 * m[q] + (1.0f - b1) * g;
    const float vq = b2 * v[q] + (1.0f - b2) * g * g;
    m[q] = mq;
    v[q] = vq;
    P[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

}  // namespace

extern "C" int bls_mlp_dims(int *nfeatures, int *nhidden, int *nparams, int *work_per_sample) {
    if (nfeatures) *nfeatures = NF;
    if (nhidden) *nhidden = NH;
    if (nparams) *nparams = NP;
    if (work_per_sample) *work_per_sample = WORK_PER_SAMPLE;
    return BLS_OK;
}

extern "C" int bls_mlp_forward_f32(const float *params, const float *fmean, const float *fstd, const float *X,
                                   int64_t n, float *out, void *stream) {
    if (!params || !fmean || !fstd || (n > 0 && (!X || !out)) || n < 0)
        return fail(BLS_ERR_ARG, "bls_mlp_forward_f32: bad arguments");
    if (n == 0) return BLS_OK;
    k_mlp_forward<<<grid_for(n, 256), 256, 0, (hipStream_t)stream>>>(params, fmean, fstd, X, nullptr, n, out, nullptr);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

extern "C" int bls_mlp_grad_f32(const float *params, const float *fmean, const float *fstd, const float *X,
                                const float *y, int64_t n, float *out, float *grad, float *work, void *stream) {
    if (!params || !fmean || !fstd || !X || !y || !grad || !work || n <= 0)
        return fail(BLS_ERR_ARG, "bls_mlp_grad_f32: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    k_mlp_forward<<<grid_for(n, 256), 256, 0, s>>>(params, fmean, fstd, X, y, n, out, work);
    k_mlp_wgrad<<<grid_for(NP, 128), 128, 0, s>>>(fmean, fstd, X, n, work, grad);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

extern "C" int bls_mlp_adam_f32(float *params, const float *grad, float *m, float *v, int64_t step, float lr,
                                void *stream) {
    if (!params || !grad || !m || !v || step < 1) return fail(BLS_ERR_ARG, "bls_mlp_adam_f32: bad arguments");
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    k_mlp_adam<<<grid_for(NP, 256), 256, 0, (hipStream_t)stream>>>(params, grad, m, v, lr, b1, b2, 1e-8f, c1, c2);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}
