// atlas_kernel.hip -- the compose step of the raw KFP atlas colour pass
// (atlas_core.h): one output pixel per lane, 256-thread blocks of 64x4
// pixels so each wavefront writes one contiguous row segment.  The KFP
// colour kernel (kfp_kernel.hip) then colours the composed field.
#include <hip/hip_runtime.h>

#include "atlas_core.h"

namespace fav {
namespace {

__global__ void __launch_bounds__(256) atlas_compose_kernel(atlas::ComposeParams a) {
    const int x = static_cast<int>(blockIdx.x) * 64 + static_cast<int>(threadIdx.x & 63);
    const int y = static_cast<int>(blockIdx.y) * 4 + static_cast<int>(threadIdx.x >> 6);
    if (x >= a.width || y >= a.height) return;
    a.out[static_cast<size_t>(y) * static_cast<size_t>(a.width) + static_cast<size_t>(x)] =
        atlas::compose_pixel(a, x, y);
}

__global__ void __launch_bounds__(256) atlas_planes_kernel(atlas::PlanesCompose a) {
    const int x = static_cast<int>(blockIdx.x) * 64 + static_cast<int>(threadIdx.x & 63);
    const int y = static_cast<int>(blockIdx.y) * 4 + static_cast<int>(threadIdx.x >> 6);
    if (x >= a.width || y >= a.height) return;
    atlas::compose_planes_pixel(a, x, y);
}

}  // namespace

// The plane-aware compose; shapes validated by the caller as below.
hipError_t launch_atlas_planes(const atlas::PlanesCompose& a, hipStream_t stream) {
    const dim3 grid(static_cast<unsigned>((a.width + 63) / 64), static_cast<unsigned>((a.height + 3) / 4));
    atlas_planes_kernel<<<grid, dim3(256), 0, stream>>>(a);
    return hipGetLastError();
}

// Shapes validated by the caller: the axes hold width / height / sampled
// child entries and index inside their tiles.
hipError_t launch_atlas_compose(const atlas::ComposeParams& a, hipStream_t stream) {
    const dim3 grid(static_cast<unsigned>((a.width + 63) / 64), static_cast<unsigned>((a.height + 3) / 4));
    atlas_compose_kernel<<<grid, dim3(256), 0, stream>>>(a);
    return hipGetLastError();
}

}  // namespace fav
