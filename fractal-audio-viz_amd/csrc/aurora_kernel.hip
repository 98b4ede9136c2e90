// aurora_kernel.hip -- the ordinary-palette colouriser and the per-frame
// crop / atlas compositors (aurora_core.h): one output pixel per lane,
// 256-thread blocks of 64x4 pixels so each wavefront reads and writes one
// contiguous row segment.  HBM-bound byte and float work; the mode is a
// template parameter so each launch compiles to its own straight-line body.
#include <hip/hip_runtime.h>

#include "aurora_core.h"

namespace fav {
namespace {

template <int Mode>
__global__ void __launch_bounds__(256) aurora_kernel(aurora::Params p) {
    const int x = static_cast<int>(blockIdx.x) * 64 + static_cast<int>(threadIdx.x & 63);
    const int y = static_cast<int>(blockIdx.y) * 4 + static_cast<int>(threadIdx.x >> 6);
    if (x >= p.width || y >= p.height) return;
    const size_t i = static_cast<size_t>(y) * static_cast<size_t>(p.width) + static_cast<size_t>(x);
    if constexpr (Mode == aurora::kColour) {
        aurora::colour_pixel(p, p.parent_field[i], p.out_rgb + i * 3);
    } else if constexpr (Mode == aurora::kAccents) {
        aurora::accents_pixel(p, p.out_rgb + i * 3);
    } else if constexpr (Mode == aurora::kCropRgb) {
        aurora::sample_rgb(p.parent_rgb, p.parent_width, p.parent_x, p.parent_y, x, y, p.out_rgb + i * 3);
    } else if constexpr (Mode == aurora::kAtlasRgb) {
        aurora::atlas_rgb_pixel(p, x, y, p.out_rgb + i * 3);
    } else if constexpr (Mode == aurora::kCropField) {
        p.out_field[i] = aurora::sample_plain(p.parent_field, p.parent_width, p.parent_x, p.parent_y, x, y);
    } else if constexpr (Mode == aurora::kAtlasField) {
        p.out_field[i] = aurora::atlas_field_pixel(p, x, y);
    } else if constexpr (Mode == aurora::kCropColour) {
        aurora::crop_colour_pixel(p, x, y, p.out_rgb + i * 3);
    } else {
        aurora::atlas_colour_pixel(p, x, y, p.out_rgb + i * 3);
    }
}

}  // namespace

// Shapes validated by the caller: the sources, axes, seam weights and
// palette hold what the mode reads; the outputs hold width*height values.
hipError_t launch_aurora(const aurora::Params& p, hipStream_t stream) {
    const dim3 grid(static_cast<unsigned>((p.width + 63) / 64), static_cast<unsigned>((p.height + 3) / 4));
    switch (p.mode) {
        case aurora::kColour: aurora_kernel<aurora::kColour><<<grid, dim3(256), 0, stream>>>(p); break;
        case aurora::kAccents: aurora_kernel<aurora::kAccents><<<grid, dim3(256), 0, stream>>>(p); break;
        case aurora::kCropRgb: aurora_kernel<aurora::kCropRgb><<<grid, dim3(256), 0, stream>>>(p); break;
        case aurora::kAtlasRgb: aurora_kernel<aurora::kAtlasRgb><<<grid, dim3(256), 0, stream>>>(p); break;
        case aurora::kCropField: aurora_kernel<aurora::kCropField><<<grid, dim3(256), 0, stream>>>(p); break;
        case aurora::kAtlasField: aurora_kernel<aurora::kAtlasField><<<grid, dim3(256), 0, stream>>>(p); break;
        case aurora::kCropColour: aurora_kernel<aurora::kCropColour><<<grid, dim3(256), 0, stream>>>(p); break;
        case aurora::kAtlasColour: aurora_kernel<aurora::kAtlasColour><<<grid, dim3(256), 0, stream>>>(p); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace fav
