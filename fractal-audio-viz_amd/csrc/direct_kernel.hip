// direct_kernel.hip -- shallow direct renderer on gfx950 (per-pixel code in
// direct_core.h).
//
// A 256-thread block covers a 16x16 tile as four 8x8 wavefronts, so the
// lanes of a wave are neighbours with similar escape times.  The work is
// FP64 VALU-bound (two complex squarings per iteration, no table reads).
#include <hip/hip_runtime.h>

#include "direct_core.h"

namespace fav {
namespace {

template <int F, bool Planes, bool Derivative>
__global__ void __launch_bounds__(256) direct_kernel(direct::DirectParams p) {
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int px = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    if (px < p.width && py < p.height) direct::pixel<F, Planes, Derivative>(p, px, py);
}

template <int F, bool Planes, bool Derivative>
void launch(const direct::DirectParams& p, hipStream_t stream) {
    direct_kernel<F, Planes, Derivative>
        <<<dim3((p.width + 15) / 16, (p.height + 15) / 16), dim3(256), 0, stream>>>(p);
}

}  // namespace

// Shapes validated by the caller: x has width and y height entries, output
// (and the planes when present) width*height.  Planes with the analytic DE
// exist for Mandelbrot and Julia only.
hipError_t launch_direct_kernel(const direct::DirectParams& p, bool planes, hipStream_t stream) {
    const bool de = planes && (p.plane_flags & direct::kHintSkipAnalyticDe) == 0U;
    switch (p.formula) {
        case 0:
            if (!planes) launch<0, false, false>(p, stream);
            else if (de) launch<0, true, true>(p, stream);
            else launch<0, true, false>(p, stream);
            break;
        case 1:
            if (!planes) launch<1, false, false>(p, stream);
            else if (de) launch<1, true, true>(p, stream);
            else launch<1, true, false>(p, stream);
            break;
        case 2:
            if (!planes) launch<2, false, false>(p, stream);
            else launch<2, true, false>(p, stream);
            break;
        case 3:
            if (!planes) launch<3, false, false>(p, stream);
            else launch<3, true, false>(p, stream);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace fav
