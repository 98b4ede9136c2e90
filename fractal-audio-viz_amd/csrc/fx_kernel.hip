// fx_kernel.hip -- the FP32 floatexp perturbation path on gfx950 (per-pixel
// code in fx_core.h).  One pixel per lane; a 256-thread block covers a
// 16x16 tile as four 8x8 wavefronts.
#include <hip/hip_runtime.h>

#include "fx_core.h"

// Occupancy targets (waves per SIMD), measured on the 2560x1440
// near-nucleus fields: the Mandelbrot kernel is left to the compiler (95
// registers, 5 waves; 6 and 8 waves spill and run 3-29% slower); the
// alternate-formula kernel at 4 waves (128 registers, from 166 and 3 waves)
// renders the burning-ship fields 12-14% faster, at 5 waves 5-15%.
#ifndef FAV_FX_WAVES
#define FAV_FX_WAVES 1
#endif
#ifndef FAV_FX_ALT_WAVES
#define FAV_FX_ALT_WAVES 4
#endif

namespace fav {
namespace {

__device__ __forceinline__ void tile_pixel(int& px, int& py) {
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    px = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    py = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FAV_FX_WAVES, 8)))
fx_kernel(RenderParams p, fx::FxOut out) {
    int px, py;
    tile_pixel(px, py);
    if (px < p.width && py < p.height) fx::pixel(p, out, px, py);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FAV_FX_ALT_WAVES, 8)))
fx_alt_kernel(AltRenderParams p, fx::FxOut out) {
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
