// sched_mlp.h -- shapes and launchers of the learned tile-cost model
// (sched_mlp.hip): 65 inputs (an 8x8 preview patch of one 32x32 tile plus
// the zoom's log10 / 100) -> 64 -> 64 -> 1 with ReLU, bf16 weights and
// inputs, fp32 sums.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace fav {
namespace mlp {

constexpr int kInputs = 65;
constexpr int kHidden = 64;
constexpr int kParams = kInputs * kHidden + kHidden + kHidden * kHidden + kHidden + kHidden + 1;  // 8449
constexpr int kActivations = 2 * kHidden;   // h1, h2 per sample
constexpr int kDeltas = 2 * kHidden + 1;    // dh1, d2, dy per sample
constexpr int kChunk = 64;                  // samples per partial gradient
constexpr float kBeta1 = 0.9F, kBeta2 = 0.999F;  // Adam moments

constexpr int kTile = 32;                   // field pixels per tile side
constexpr int kPatch = 8;                   // preview pixels per tile side (quarter resolution)
constexpr float kFeatureScale = 0.125F;     // features are log(1 + preview work) * kFeatureScale

static_assert(kInputs == kPatch * kPatch + 1, "one input per preview pixel of a tile, plus the zoom");

inline int chunk_count(int n) { return (n + kChunk - 1) / kChunk; }

// Device scratch of one gradient evaluation over n samples.
struct Workspace {
    float* y;        // n
    float* act;      // n * kActivations
    float* delta;    // n * kDeltas
    float* err2;     // n
    float* partial;  // chunk_count(n) * kParams
};

// All pointers are device memory; shapes are checked by the caller.
hipError_t forward(const float* weights, const float* x, int n, float* y, float* act, hipStream_t stream);
hipError_t gradient(const float* weights, const float* x, const float* t, int n, const Workspace& ws, float* grad,
                    float* loss, hipStream_t stream);
// Tile features (tiles_x * tiles_y rows of kInputs floats, row-major tiles)
// of a preview work map of pw x ph pixels: patch element (r, c) of tile
// (tx, ty) is log1p(work) * kFeatureScale at preview pixel
// (min(8 tx + c, pw - 1), min(8 ty + r, ph - 1)); the last column is
// zoom_feature.
hipError_t features(const int32_t* work, int pw, int ph, int tiles_x, int tiles_y, float zoom_feature, float* out,
                    hipStream_t stream);
// Adam step `step` (counted from 1) on the fp32 master weights.
hipError_t adam(float* weights, const float* grad, float* m, float* v, float lr, int step, hipStream_t stream);

}  // namespace mlp
}  // namespace fav
