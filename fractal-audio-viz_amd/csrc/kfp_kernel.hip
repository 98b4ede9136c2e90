// kfp_kernel.hip -- Kalles SetColor colour transfer of a scalar field on
// gfx950 (per-pixel code in kfp_core.h).
//
// The pass reads the float32 field (4 bytes per pixel; the stencil's
// neighbours come from the same cache lines) and writes interleaved RGB8,
// so it is HBM-bound.  A 256-thread block covers a 64x4 strip: each wave is
// one row segment of 64 pixels, so field reads and RGB writes are
// coalesced.  Colour method 4 first reduces the transfer range with one
// 64-bit atomic min/max per wave.
#include <hip/hip_runtime.h>

#include <climits>
#include <mutex>

#include "kfp_core.h"

namespace fav {
namespace {

// The parameters are read in place from the kernel-argument segment, where
// the only explicit argument sits at offset 0: a reference to the by-value
// argument would make the compiler copy all of it (3.4 KB with the period
// tables) into every lane's private memory.
__global__ void __launch_bounds__(256) kfp_colour_kernel(kfp::KfpParams) {
    const kfp::KfpParams& k = *(const kfp::KfpParams*)__builtin_amdgcn_kernarg_segment_ptr();
    const int x = blockIdx.x * 64 + static_cast<int>(threadIdx.x & 63);
    const int y = blockIdx.y * 4 + static_cast<int>(threadIdx.x >> 6);
    if (x >= k.width || y >= k.height) return;
    kfp::colour_pixel(k, x, y);
}

// bounds[0] / bounds[1] hold the running min / max as int64 (initialised to
// LLONG_MAX / LLONG_MIN by kfp_bounds_init).
__global__ void __launch_bounds__(256) kfp_bounds_kernel(kfp::KfpParams k, long long* bounds) {
    const size_t n = static_cast<size_t>(k.width) * static_cast<size_t>(k.height);
    long long lo = LLONG_MAX, hi = LLONG_MIN;
    for (size_t i = blockIdx.x * 256ULL + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256ULL) {
        long long v;
        if (kfp::transfer_bound_value(k, i, v)) {
            lo = v < lo ? v : lo;
            hi = v > hi ? v : hi;
        }
    }
    for (int offset = 32; offset > 0; offset >>= 1) {
        const long long olo = __shfl_down(lo, offset);
        const long long ohi = __shfl_down(hi, offset);
        lo = olo < lo ? olo : lo;
        hi = ohi > hi ? ohi : hi;
    }
    if ((threadIdx.x & 63) == 0) {
        if (lo != LLONG_MAX) atomicMin(bounds, lo);
        if (hi != LLONG_MIN) atomicMax(bounds + 1, hi);
    }
}

__global__ void kfp_bounds_init(long long* bounds) {
    bounds[0] = LLONG_MAX;
    bounds[1] = LLONG_MIN;
}

// Converts the int64 range to the doubles colour_pixel reads; an empty range
// keeps the reference's defaults 0 and 1 (renderer.cpp:7329-7333).
__global__ void kfp_bounds_finish(const long long* bounds, double* out) {
    const bool empty = bounds[0] == LLONG_MAX || bounds[1] == LLONG_MIN;
    out[0] = empty ? 0.0 : static_cast<double>(bounds[0]);
    out[1] = empty ? 1.0 : static_cast<double>(bounds[1]);
}

__global__ void __launch_bounds__(256) kfp_rect_atlas_kernel(const kfp::KfpParams* params, kfp::RectAtlas g) {
    const kfp::KfpParams& parent = params[0];
    const int x = blockIdx.x * 64 + static_cast<int>(threadIdx.x & 63);
    const int y = blockIdx.y * 4 + static_cast<int>(threadIdx.x >> 6);
    if (x >= parent.width || y >= parent.height) return;
    kfp::rect_atlas_pixel(parent, params[1], g, x, y);
}

void launch_bounds(const kfp::KfpParams& k, long long* bounds_i, double* bounds_d, int compute_units,
                   hipStream_t stream) {
    kfp_bounds_init<<<1, 1, 0, stream>>>(bounds_i);
    kfp_bounds_kernel<<<dim3(compute_units * 4), dim3(256), 0, stream>>>(k, bounds_i);
    kfp_bounds_finish<<<1, 1, 0, stream>>>(bounds_i, bounds_d);
}

}  // namespace

// The explicit-rectangle KFP atlas: `params` is device memory for two
// KfpParams (parent, child) that the host fills after this call sets their
// bounds; `scratch` holds 4 int64 + 4 doubles.  The caller copies
// `host_params` (bounds pointers set here) to `params` on `stream` before
// the colour launch, which this function issues after the copy.
hipError_t launch_kfp_rect_atlas(kfp::KfpParams* host_params, kfp::KfpParams* params, const kfp::RectAtlas& g,
                                 void* scratch, int compute_units, hipStream_t stream) {
    long long* bounds_i = static_cast<long long*>(scratch);
    double* bounds_d = reinterpret_cast<double*>(bounds_i + 4);
    for (int t = 0; t < 2; ++t) {
        host_params[t].bounds = bounds_d + 2 * t;
        if (host_params[t].color_method == 4 && (t == 0 || g.use_child))
            launch_bounds(host_params[t], bounds_i + 2 * t, bounds_d + 2 * t, compute_units, stream);
    }
    hipError_t e = hipMemcpyAsync(params, host_params, 2 * sizeof(kfp::KfpParams), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return e;
    const kfp::KfpParams& parent = host_params[0];
    kfp_rect_atlas_kernel<<<dim3((parent.width + 63) / 64, (parent.height + 3) / 4), dim3(256), 0, stream>>>(params, g);
    return hipGetLastError();
}

// ---- scratch of the allocation-free entry --------------------------------
// fractal_hip_colourise_kfp_device_noalloc keeps its scratch in this
// file-scope array instead of library-allocated buffers: bytes [0, 256) are
// the colour-method-4 range reduction (2 int64 + 2 doubles), then the
// palette LUT staged from the host (at most 65536 entries of 3 bytes).
// Every device the library runs on holds one copy until the process ends;
// every call on a device shares it.
constexpr size_t kNoAllocLutOffset = 256;
constexpr size_t kNoAllocScratchBytes = kNoAllocLutOffset + 65536 * 3;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// The device address of `scratch` on `device` (which must be current):
// looked up once per device with hipGetSymbolAddress and kept; host code
// uses only this address, never the array's name.
hipError_t kfp_noalloc_scratch(int device, unsigned char** out, size_t* lut_offset, size_t* bytes) {
    static std::mutex mutex;
    static unsigned char* addresses[64] = {};
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mutex);
    if (!addresses[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addresses[device] = static_cast<unsigned char*>(ptr);
    }
    *out = addresses[device];
    *lut_offset = kNoAllocLutOffset;
    *bytes = kNoAllocScratchBytes;
    return hipSuccess;
}

// `scratch` holds 2 int64 + 2 doubles.  Shapes validated by the caller.
hipError_t launch_kfp_kernel(const kfp::KfpParams& params, void* scratch, int compute_units, hipStream_t stream) {
    kfp::KfpParams k = params;
    long long* bounds_i = static_cast<long long*>(scratch);
    double* bounds_d = reinterpret_cast<double*>(bounds_i + 2);
    if (k.color_method == 4) {
        kfp_bounds_init<<<1, 1, 0, stream>>>(bounds_i);
        kfp_bounds_kernel<<<dim3(compute_units * 4), dim3(256), 0, stream>>>(k, bounds_i);
        kfp_bounds_finish<<<1, 1, 0, stream>>>(bounds_i, bounds_d);
    }
    k.bounds = bounds_d;
    kfp_colour_kernel<<<dim3((k.width + 63) / 64, (k.height + 3) / 4), dim3(256), 0, stream>>>(k);
    return hipGetLastError();
}

}  // namespace fav
