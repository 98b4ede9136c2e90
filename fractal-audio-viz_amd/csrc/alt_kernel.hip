// alt_kernel.hip -- alternate deep formulas (Julia, burning ship, tricorn)
// on gfx950; the per-pixel loop is alt_core.h.
//
// One pixel per lane; a 256-thread block covers a 16x16 tile as four 8x8
// wavefronts so neighbouring pixels, which share their reference index and
// BLA level, run together.  FP64 VALU-bound like the Mandelbrot kernels.
#include <hip/hip_runtime.h>

#include "alt_core.h"

namespace fav {
namespace {

// 5 waves per SIMD (102 VGPRs, some spills) against the 3 that the
// unconstrained allocation (151 VGPRs) gives: ≈24% less time on 3840x2160
// Julia fields at zoom 1e100; 4 and 6 waves measured slower or equal.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) alt_kernel(AltRenderParams p) {
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    int bx = blockIdx.x, by = blockIdx.y;
    if (p.tile_order) {
        // blocks in launch order cover the 32x32 tiles of tile_order, four
        // 16x16 quarters each; the hardware dispatches blocks in index order
        const uint32_t tile = p.tile_order[blockIdx.x >> 2];
        const int tiles_x = (p.width + 31) / 32;
        const int quarter = blockIdx.x & 3;
        bx = static_cast<int>(tile % static_cast<uint32_t>(tiles_x)) * 2 + (quarter & 1);
        by = static_cast<int>(tile / static_cast<uint32_t>(tiles_x)) * 2 + (quarter >> 1);
    }
    const int px = bx * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = by * 16 + (wave >> 1) * 8 + (lane >> 3);
    if (px >= p.width || py >= p.height) return;
    if (p.point_re && static_cast<size_t>(py) * static_cast<size_t>(p.width) + static_cast<size_t>(px)
                          >= static_cast<size_t>(p.point_count))
        return;  // the unused end of a point render's last row
    alt::pixel(p, px, py);
}

__global__ void alt_prologue(unsigned long long* budget_start, int* budget_flag) {
    budget_start[0] = wall_clock64();
    budget_flag[0] = 0;
}

}  // namespace

// Shapes validated by the caller: offsets hold width / height entries and
// output (and work) width*height entries; tile_order, when set, is a
// permutation of the ceil(w/32)*ceil(h/32) tiles; the orbit, margins and BLA
// steps are the reference's device tables.
hipError_t launch_alt_kernel(const AltRenderParams& p, hipStream_t stream) {
    alt_prologue<<<1, 1, 0, stream>>>(const_cast<unsigned long long*>(p.budget_start), p.budget_flag);
    if (p.tile_order)
        alt_kernel<<<dim3(static_cast<unsigned>(p.tile_count) * 4U), dim3(256), 0, stream>>>(p);
    else
        alt_kernel<<<dim3((p.width + 15) / 16, (p.height + 15) / 16), dim3(256), 0, stream>>>(p);
    return hipGetLastError();
}

}  // namespace fav
