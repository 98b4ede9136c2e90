// sched_mlp.hip -- the learned tile-cost model: a small MLP
// (65 inputs -> 64 -> 64 -> 1, ReLU) whose weights and inputs are bf16
// values, trained on the GPU by the kernels below.
//
// Arithmetic: weights (fp32 master copy) and inputs are rounded to bf16
// (nearest, ties to even) when they are read; sums and activations are fp32.
// Training takes the gradient with respect to the rounded weights and
// applies it to the master copy (straight-through), by full-batch Adam.
// Every reduction runs in a fixed order, so a training run is reproducible
// bit for bit.
//
// Parameter layout (kParams floats): W1[65][64], b1[64], W2[64][64], b2[64],
// W3[64], b3.  Each block boundary is a multiple of 64, so a wavefront of
// the gradient kernel (one parameter per lane) never straddles two blocks.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "sched_mlp.h"

namespace fav {
namespace mlp {
namespace {

constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kInputs * kHidden;
constexpr int kW2 = kB1 + kHidden;
constexpr int kB2 = kW2 + kHidden * kHidden;
constexpr int kW3 = kB2 + kHidden;
constexpr int kB3 = kW3 + kHidden;
static_assert(kB3 + 1 == kParams, "parameter layout");
static_assert(kB1 % 64 == 0 && kW2 % 64 == 0 && kB2 % 64 == 0 && kW3 % 64 == 0 && kB3 % 64 == 0,
              "parameter blocks start on wavefront boundaries");

__device__ __forceinline__ float bf16_round(float v) {
    uint32_t bits = __float_as_uint(v);
    if ((bits & 0x7f800000U) == 0x7f800000U) return v;  // inf / NaN unchanged
    bits += 0x7fffU + ((bits >> 16) & 1U);
    return __uint_as_float(bits & 0xffff0000U);
}

__device__ __forceinline__ void stage_weights(const float* w, float* sw) {
    for (int i = threadIdx.x; i < kParams; i += 256) sw[i] = bf16_round(w[i]);
    __syncthreads();
}

// One sample per thread.  y = W3 . h2 + b3, h2 = relu(W2^T h1 + b2),
// h1 = relu(W1^T x + b1).  `act` (when set) receives h1 and h2.
__global__ void __launch_bounds__(256) forward_kernel(const float* w, const float* x, int n, float* y, float* act) {
    __shared__ float sw[kParams];
    stage_weights(w, sw);
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const float* xs = x + static_cast<size_t>(s) * kInputs;
    float h1[kHidden];
    for (int j = 0; j < kHidden; ++j) h1[j] = sw[kB1 + j];
    for (int i = 0; i < kInputs; ++i) {
        const float xi = bf16_round(xs[i]);
        for (int j = 0; j < kHidden; ++j) h1[j] += xi * sw[kW1 + i * kHidden + j];
    }
    for (int j = 0; j < kHidden; ++j) h1[j] = h1[j] > 0.0F ? h1[j] : 0.0F;
    float* as = act ? act + static_cast<size_t>(s) * kActivations : nullptr;
    float out = sw[kB3];
    for (int k = 0; k < kHidden; ++k) {
        float a = sw[kB2 + k];
        for (int j = 0; j < kHidden; ++j) a += h1[j] * sw[kW2 + j * kHidden + k];
        a = a > 0.0F ? a : 0.0F;
        out += a * sw[kW3 + k];
        if (as) as[kHidden + k] = a;
    }
    if (as)
        for (int j = 0; j < kHidden; ++j) as[j] = h1[j];
    y[s] = out;
}

// Back-propagated errors of the loss mean((y - t)^2): per sample
// dh1 (64), d2 (64) and dy, plus the squared error.
__global__ void __launch_bounds__(256) delta_kernel(const float* w, const float* t, int n, const float* y,
                                                    const float* act, float* delta, float* err2) {
    __shared__ float sw[kParams];
    stage_weights(w, sw);
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const float* h1 = act + static_cast<size_t>(s) * kActivations;
    const float* h2 = h1 + kHidden;
    float* ds = delta + static_cast<size_t>(s) * kDeltas;
    const float err = y[s] - t[s];
    err2[s] = err * err;
    const float dy = 2.0F * err / static_cast<float>(n);
    float d2[kHidden];
    for (int k = 0; k < kHidden; ++k) {
        d2[k] = h2[k] > 0.0F ? dy * sw[kW3 + k] : 0.0F;
        ds[kHidden + k] = d2[k];
    }
    for (int j = 0; j < kHidden; ++j) {
        float acc = 0.0F;
        if (h1[j] > 0.0F)
            for (int k = 0; k < kHidden; ++k) acc += sw[kW2 + j * kHidden + k] * d2[k];
        ds[j] = acc;
    }
    ds[2 * kHidden] = dy;
}

// Partial gradients: one parameter per thread, one chunk of kChunk samples
// per blockIdx.y, summed in sample order.
__global__ void __launch_bounds__(256) gradient_kernel(const float* x, const float* act, const float* delta, int n,
                                                       float* partial) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kParams) return;
    const int first = blockIdx.y * kChunk;
    const int last = first + kChunk < n ? first + kChunk : n;
    float acc = 0.0F;
    for (int s = first; s < last; ++s) {
        const float* xs = x + static_cast<size_t>(s) * kInputs;
        const float* as = act + static_cast<size_t>(s) * kActivations;
        const float* ds = delta + static_cast<size_t>(s) * kDeltas;
        if (p < kB1) acc += bf16_round(xs[p / kHidden]) * ds[p % kHidden];     // W1[i][j]: x_i dh1_j
        else if (p < kW2) acc += ds[p - kB1];                                    // b1[j]: dh1_j
        else if (p < kB2) acc += as[(p - kW2) / kHidden] * ds[kHidden + (p - kW2) % kHidden];  // W2[j][k]
        else if (p < kW3) acc += ds[kHidden + p - kB2];                          // b2[k]: d2_k
        else if (p < kB3) acc += ds[2 * kHidden] * as[kHidden + p - kW3];        // W3[k]: dy h2_k
        else acc += ds[2 * kHidden];                                             // b3: dy
    }
    partial[static_cast<size_t>(blockIdx.y) * kParams + p] = acc;
}

__global__ void __launch_bounds__(256) reduce_kernel(const float* partial, int chunks, float* grad) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kParams) return;
    float acc = 0.0F;
    for (int c = 0; c < chunks; ++c) acc += partial[static_cast<size_t>(c) * kParams + p];
    grad[p] = acc;
}

// Mean squared error: strided per-thread sums, then a fixed LDS tree.
__global__ void __launch_bounds__(256) loss_kernel(const float* err2, int n, float* loss) {
    __shared__ float part[256];
    float acc = 0.0F;
    for (int s = threadIdx.x; s < n; s += 256) acc += err2[s];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int width = 128; width > 0; width >>= 1) {
        if (threadIdx.x < width) part[threadIdx.x] += part[threadIdx.x + width];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = part[0] / static_cast<float>(n);
}

// Adam step; c1 and c2 are the bias corrections 1 - beta1^t and 1 - beta2^t.
__global__ void __launch_bounds__(256) adam_kernel(float* master, const float* grad, float* m, float* v, float lr,
                                                   float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kParams) return;
    constexpr float beta1 = kBeta1, beta2 = kBeta2, eps = 1e-8F;
    const float g = grad[i];
    const float mi = beta1 * m[i] + (1.0F - beta1) * g;
    const float vi = beta2 * v[i] + (1.0F - beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    master[i] -= lr * (mi / c1) / (__builtin_sqrtf(vi / c2) + eps);
}

__global__ void __launch_bounds__(256) features_kernel(const int32_t* work, int pw, int ph, int tiles_x, int count,
                                                       float zoom_feature, float* out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const int tile = i / kInputs, e = i - tile * kInputs;
    if (e == kPatch * kPatch) {
        out[i] = zoom_feature;
        return;
    }
    const int tx = tile % tiles_x, ty = tile / tiles_x;
    const int x = min(tx * kPatch + e % kPatch, pw - 1);
    const int y = min(ty * kPatch + e / kPatch, ph - 1);
    const int v = work[static_cast<size_t>(y) * static_cast<size_t>(pw) + x];
    out[i] = log1pf(static_cast<float>(v > 0 ? v : 0)) * kFeatureScale;
}

constexpr unsigned blocks_for(int count) { return static_cast<unsigned>((count + 255) / 256); }

}  // namespace

hipError_t forward(const float* weights, const float* x, int n, float* y, float* act, hipStream_t stream) {
    forward_kernel<<<blocks_for(n), 256, 0, stream>>>(weights, x, n, y, act);
    return hipGetLastError();
}

hipError_t gradient(const float* weights, const float* x, const float* t, int n, const Workspace& ws, float* grad,
                    float* loss, hipStream_t stream) {
    forward_kernel<<<blocks_for(n), 256, 0, stream>>>(weights, x, n, ws.y, ws.act);
    delta_kernel<<<blocks_for(n), 256, 0, stream>>>(weights, t, n, ws.y, ws.act, ws.delta, ws.err2);
    const int chunks = chunk_count(n);
    gradient_kernel<<<dim3(blocks_for(kParams), static_cast<unsigned>(chunks)), 256, 0, stream>>>(
        x, ws.act, ws.delta, n, ws.partial);
    reduce_kernel<<<blocks_for(kParams), 256, 0, stream>>>(ws.partial, chunks, grad);
    loss_kernel<<<1, 256, 0, stream>>>(ws.err2, n, loss);
    return hipGetLastError();
}

hipError_t features(const int32_t* work, int pw, int ph, int tiles_x, int tiles_y, float zoom_feature, float* out,
                    hipStream_t stream) {
    const int count = tiles_x * tiles_y * kInputs;
    features_kernel<<<blocks_for(count), 256, 0, stream>>>(work, pw, ph, tiles_x, count, zoom_feature, out);
    return hipGetLastError();
}

hipError_t adam(float* weights, const float* grad, float* m, float* v, float lr, int step, hipStream_t stream) {
    const float c1 = static_cast<float>(1.0 - std::pow(static_cast<double>(kBeta1), step));
    const float c2 = static_cast<float>(1.0 - std::pow(static_cast<double>(kBeta2), step));
    adam_kernel<<<blocks_for(kParams), 256, 0, stream>>>(weights, grad, m, v, lr, c1, c2);
    return hipGetLastError();
}

}  // namespace mlp
}  // namespace fav
