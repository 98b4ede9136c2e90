// fx_kernel.hip -- the FP32 floatexp perturbation path on gfx950 (per-pixel
// code in fx_core.h).  One pixel per lane; a 256-thread block covers a
// 16x16 tile as four 8x8 wavefronts.
#include <hip/hip_runtime.h>

#include "fx_core.h"

namespace fav {
namespace {

__device__ __forceinline__ void tile_pixel(int& px, int& py) {
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    px = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    py = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
}

__global__ void __launch_bounds__(256) fx_kernel(RenderParams p, fx::FxOut out) {
    int px, py;
    tile_pixel(px, py);
    if (px < p.width && py < p.height) fx::pixel(p, out, px, py);
}

__global__ void __launch_bounds__(256) fx_alt_kernel(AltRenderParams p, fx::FxOut out) {
    int px, py;
    tile_pixel(px, py);
    if (px < p.width && py < p.height) fx::alt_pixel(p, out, px, py);
}

}  // namespace

// Shapes validated by the caller: out.field / out.counts hold width*height
// entries; p carries the FP64 tables and offsets of the same render.
hipError_t launch_fx_kernel(const RenderParams& p, const fx::FxOut& out, hipStream_t stream) {
    fx_kernel<<<dim3((p.width + 15) / 16, (p.height + 15) / 16), dim3(256), 0, stream>>>(p, out);
    return hipGetLastError();
}

// The same for an alternate-formula reference (Julia, burning ship, tricorn).
hipError_t launch_fx_alt_kernel(const AltRenderParams& p, const fx::FxOut& out, hipStream_t stream) {
    fx_alt_kernel<<<dim3((p.width + 15) / 16, (p.height + 15) / 16), dim3(256), 0, stream>>>(p, out);
    return hipGetLastError();
}

}  // namespace fav
