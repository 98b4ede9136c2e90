// render_kernel.hip -- the FP64 deep-zoom field kernel for gfx950.
//
// Work layout: a 256-thread block covers a 16x16 pixel tile as four 8x8
// wavefronts.  Neighbouring pixels follow near-identical orbits, so an 8x8
// wave keeps its lanes on the same reference index and BLA level for most of
// the iteration, and the orbit / table loads are wave-uniform broadcasts.
// The per-pixel state machine lives in field_core.h.
#include <hip/hip_runtime.h>

#include "field_core.h"

namespace fav {
namespace {

template <bool Planes, bool Stats>
__global__ void __launch_bounds__(256) field_kernel(RenderParams p) {
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int px = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    if (px >= p.width || py >= p.height) return;
    core::render_one<Planes, Stats>(p, px, py);
}

__global__ void budget_start_kernel(unsigned long long* start, int* flag) {
    start[0] = wall_clock64();
    flag[0] = 0;
}

}  // namespace

// Launches the field kernel on `stream`.  The caller validated the shapes:
// width*height <= 1e8 and every pointer in `p` sized for width x height.
hipError_t launch_field_kernel(const RenderParams& p, bool planes, bool stats, hipStream_t stream) {
    if (p.budget_ticks != 0)
        budget_start_kernel<<<1, 1, 0, stream>>>(const_cast<unsigned long long*>(p.budget_start), p.budget_flag);
    const dim3 grid((p.width + 15) / 16, (p.height + 15) / 16);
    const dim3 block(256);
    if (planes) {
        if (stats) field_kernel<true, true><<<grid, block, 0, stream>>>(p);
        else field_kernel<true, false><<<grid, block, 0, stream>>>(p);
    } else {
        if (stats) field_kernel<false, true><<<grid, block, 0, stream>>>(p);
        else field_kernel<false, false><<<grid, block, 0, stream>>>(p);
    }
    return hipGetLastError();
}

}  // namespace fav
