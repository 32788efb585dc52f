// hgap_mlp.hip -- the learned launch-shape predictor: a two-layer perceptron
//   h = relu(W1 x + b1),  y = W2 h + b2
// trained by masked mean-squared error on the autotuner's measured (relative log-time) costs,
// with hand-written forward, backward and Adam kernels (fp32).  Sizes are tiny (a few thousand
// samples, ~1000 parameters): each kernel is a handful of workgroups of plain FMA loops.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "hgap.h"

namespace hgap {

// Parameters are one flat fp32 array: W1 [H x I], b1 [H], W2 [O x H], b2 [O].
struct MlpDims {
    uint32_t in, hidden, out;
};

__device__ __forceinline__ const float *w1(const float *p, MlpDims) { return p; }
__device__ __forceinline__ const float *b1(const float *p, MlpDims d) { return p + d.hidden * d.in; }
__device__ __forceinline__ const float *w2(const float *p, MlpDims d) {
    return p + d.hidden * d.in + d.hidden;
}
__device__ __forceinline__ const float *b2(const float *p, MlpDims d) {
    return p + d.hidden * d.in + d.hidden + d.out * d.hidden;
}

// one thread per (sample, hidden unit) then per (sample, output)
__global__ void mlp_hidden_kernel(const float *__restrict__ x, uint32_t n, const float *params,
                                  MlpDims d, float *__restrict__ h) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n * d.hidden) return;
    const uint32_t s = (uint32_t)(i / d.hidden), u = (uint32_t)(i % d.hidden);
    const float *w = w1(params, d) + (uint64_t)u * d.in;
    float acc = b1(params, d)[u];
    for (uint32_t k = 0; k < d.in; k++) acc = fmaf(w[k], x[(uint64_t)s * d.in + k], acc);
    h[i] = acc > 0.f ? acc : 0.f;
}

__global__ void mlp_output_kernel(const float *__restrict__ h, uint32_t n, const float *params,
                                  MlpDims d, float *__restrict__ y) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n * d.out) return;
    const uint32_t s = (uint32_t)(i / d.out), o = (uint32_t)(i % d.out);
    const float *w = w2(params, d) + (uint64_t)o * d.hidden;
    float acc = b2(params, d)[o];
    for (uint32_t k = 0; k < d.hidden; k++) acc = fmaf(w[k], h[(uint64_t)s * d.hidden + k], acc);
    y[i] = acc;
}

// dy = 2 (y - t) m / sum(m)   (masked mean squared error), one thread per element
__global__ void mlp_dloss_kernel(const float *__restrict__ y, const float *__restrict__ t,
                                 const float *__restrict__ m, uint64_t count, float inv_norm,
                                 float *__restrict__ dy) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    dy[i] = 2.f * (y[i] - t[i]) * m[i] * inv_norm;
}

// dh = (W2^T dy) * [h > 0], one thread per (sample, hidden unit)
__global__ void mlp_dhidden_kernel(const float *__restrict__ dy, const float *__restrict__ h,
                                   uint32_t n, const float *params, MlpDims d,
                                   float *__restrict__ dh) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n * d.hidden) return;
    const uint32_t s = (uint32_t)(i / d.hidden), u = (uint32_t)(i % d.hidden);
    if (h[i] <= 0.f) {
        dh[i] = 0.f;
        return;
    }
    const float *w = w2(params, d);
    float acc = 0.f;
    for (uint32_t o = 0; o < d.out; o++) acc = fmaf(w[(uint64_t)o * d.hidden + u], dy[(uint64_t)s * d.out + o], acc);
    dh[i] = acc;
}

// parameter gradients: one thread per parameter, summing over the samples
__global__ void mlp_grad_kernel(const float *__restrict__ x, const float *__restrict__ h,
                                const float *__restrict__ dh, const float *__restrict__ dy,
                                uint32_t n, MlpDims d, float *__restrict__ g) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nw1 = d.hidden * d.in, nb1 = d.hidden, nw2 = d.out * d.hidden, nb2 = d.out;
    if (p >= nw1 + nb1 + nw2 + nb2) return;
    float acc = 0.f;
    if (p < nw1) {
        const uint32_t u = p / d.in, k = p % d.in;
        for (uint32_t s = 0; s < n; s++) acc = fmaf(dh[(uint64_t)s * d.hidden + u], x[(uint64_t)s * d.in + k], acc);
    } else if (p < nw1 + nb1) {
        const uint32_t u = p - nw1;
        for (uint32_t s = 0; s < n; s++) acc += dh[(uint64_t)s * d.hidden + u];
    } else if (p < nw1 + nb1 + nw2) {
        const uint32_t q = p - nw1 - nb1, o = q / d.hidden, u = q % d.hidden;
        for (uint32_t s = 0; s < n; s++) acc = fmaf(dy[(uint64_t)s * d.out + o], h[(uint64_t)s * d.hidden + u], acc);
    } else {
        const uint32_t o = p - nw1 - nb1 - nw2;
        for (uint32_t s = 0; s < n; s++) acc += dy[(uint64_t)s * d.out + o];
    }
    g[p] = acc;
}

// Adam (Kingma & Ba 2015): m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
// p -= lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps)
__global__ void mlp_adam_kernel(float *__restrict__ p, const float *__restrict__ g,
                                float *__restrict__ m, float *__restrict__ v, uint32_t count,
                                float lr, float beta1, float beta2, float eps, float c1,
                                float c2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = beta1 * m[i] + (1.f - beta1) * gi;
    const float vi = beta2 * v[i] + (1.f - beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

static unsigned blocks_for(uint64_t count) { return (unsigned)((count + 255) / 256); }

}  // namespace hgap

using namespace hgap;

extern "C" {

uint32_t hgap_mlp_param_count(uint32_t in, uint32_t hidden, uint32_t out) {
    return hidden * in + hidden + out * hidden + out;
}

hgap_status hgap_mlp_forward(const float *d_x, uint32_t n, uint32_t in, uint32_t hidden,
                             uint32_t out, const float *d_params, float *d_h, float *d_y,
                             hgap_stream_t stream) {
    if (n == 0) return HGAP_OK;
    if (!d_x || !d_params || !d_h || !d_y || in == 0 || hidden == 0 || out == 0)
        return HGAP_ERR_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const MlpDims d{in, hidden, out};
    (void)hipGetLastError();
    hipLaunchKernelGGL(mlp_hidden_kernel, dim3(blocks_for((uint64_t)n * hidden)), dim3(256), 0, s,
                       d_x, n, d_params, d, d_h);
    hipLaunchKernelGGL(mlp_output_kernel, dim3(blocks_for((uint64_t)n * out)), dim3(256), 0, s,
                       d_h, n, d_params, d, d_y);
    return hipGetLastError() == hipSuccess ? HGAP_OK : HGAP_ERR_HIP;
}

hgap_status hgap_mlp_backward(const float *d_x, const float *d_h, const float *d_y,
                              const float *d_target, const float *d_mask, uint32_t n,
                              uint32_t in, uint32_t hidden, uint32_t out, float mask_sum,
                              const float *d_params, float *d_dy, float *d_dh, float *d_grad,
                              hgap_stream_t stream) {
    if (n == 0) return HGAP_OK;
    if (!d_x || !d_h || !d_y || !d_target || !d_mask || !d_params || !d_dy || !d_dh || !d_grad ||
        mask_sum <= 0.f)
        return HGAP_ERR_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const MlpDims d{in, hidden, out};
    const uint64_t ny = (uint64_t)n * out;
    (void)hipGetLastError();
    hipLaunchKernelGGL(mlp_dloss_kernel, dim3(blocks_for(ny)), dim3(256), 0, s, d_y, d_target,
                       d_mask, ny, 1.f / mask_sum, d_dy);
    hipLaunchKernelGGL(mlp_dhidden_kernel, dim3(blocks_for((uint64_t)n * hidden)), dim3(256), 0, s,
                       d_dy, d_h, n, d_params, d, d_dh);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks_for(hgap_mlp_param_count(in, hidden, out))),
                       dim3(256), 0, s, d_x, d_h, d_dh, d_dy, n, d, d_grad);
    return hipGetLastError() == hipSuccess ? HGAP_OK : HGAP_ERR_HIP;
}

hgap_status hgap_mlp_adam_step(float *d_params, const float *d_grad, float *d_m, float *d_v,
                               uint32_t count, uint32_t step, float lr, float beta1, float beta2,
                               float eps, hgap_stream_t stream) {
    if (count == 0) return HGAP_OK;
    if (!d_params || !d_grad || !d_m || !d_v || step == 0) return HGAP_ERR_INVALID_ARG;
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    (void)hipGetLastError();
    hipLaunchKernelGGL(mlp_adam_kernel, dim3(blocks_for(count)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), d_params, d_grad, d_m, d_v, count, lr,
                       beta1, beta2, eps, c1, c2);
    return hipGetLastError() == hipSuccess ? HGAP_OK : HGAP_ERR_HIP;
}

}  // extern "C"
