// render_kernel.hip -- the FP64 deep-zoom field kernels for gfx950.
//
// A field is rendered by three kernels on one stream (field_core.h holds the
// per-pixel code):
//
//   plain_kernel    every pixel: set-up, image-series jump, then the
//                   plain-double block loop; interior pixels finish here
//                   (run only when the reference has plain-double tables).
//   fast_kernel     every pixel the plain pass did not finish (or every
//                   pixel, set up here, without it): the linear-BLA loop.  A 256-thread block covers a 16x16 tile
//                   as four 8x8 wavefronts, so a wave's lanes share their
//                   reference index and BLA level and the table loads are
//                   wave-uniform.  Interior pixels finish here; any other
//                   branch hands the pixel to the general queue.
//   general_kernel  the complete loop, resumed from the handed-on state.
//                   Pixels that enter the plain-double tail go to the tail
//                   queue.
//   tail_kernel     the plain-double tail (renderer.cpp:12450-12669).
//
// Each kernel only carries the registers of its own part, so the common
// paths run at higher occupancy than one kernel holding every branch.  The
// general and tail kernels are persistent: a fixed grid whose waves take 64
// queue entries at a time until the queue is drained.
#include <hip/hip_runtime.h>

#include "field_core.h"

// Occupancy targets (waves per SIMD) of the queue-driven kernels.  The
// general kernel takes the register file of one wave per SIMD (about 270
// registers): limited to 2-5 waves it spills, and on boundary-rich fields
// it measured 5-24% slower at every such target.  The tail kernel fits 4
// waves without spilling; 5 spills and measured no faster.
#ifndef FAV_GENERAL_WAVES
#define FAV_GENERAL_WAVES 1
#endif
#ifndef FAV_TAIL_WAVES
#define FAV_TAIL_WAVES 4
#endif

namespace fav {
namespace {

__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x & 63); }

// This thread's pixel of the block's 16x16 tile (four 8x8 waves): the 2-D
// grid's tile (blockIdx.x, blockIdx.y), or with a tile order the quarter
// blockIdx.x % 4 of 32x32 tile tile_order[blockIdx.x / 4].
__device__ __forceinline__ void block_pixel(const RenderParams& p, int& px, int& py) {
    const int wave = threadIdx.x >> 6;
    const int lane = lane_id();
    int x0, y0;
    if (p.tile_order) {
        const uint32_t t = p.tile_order[blockIdx.x >> 2];
        const uint32_t tiles_x = static_cast<uint32_t>((p.width + 31) / 32);
        x0 = static_cast<int>(t % tiles_x) * 32 + static_cast<int>(blockIdx.x & 1) * 16;
        y0 = static_cast<int>(t / tiles_x) * 32 + static_cast<int>((blockIdx.x >> 1) & 1) * 16;
    } else {
        x0 = blockIdx.x * 16;
        y0 = blockIdx.y * 16;
    }
    px = x0 + (wave & 1) * 8 + (lane & 7);
    py = y0 + (wave >> 1) * 8 + (lane >> 3);
}

// Appends `value` for every lane with `want`, one atomic per wave.  Called
// by all lanes of the wave.
__device__ void wave_push(uint32_t* queue, unsigned* count, uint32_t value, bool want) {
    const unsigned long long mask = __ballot(want);
    if (mask == 0) return;
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    unsigned base = 0;
    if (lane_id() == leader) base = atomicAdd(count, static_cast<unsigned>(__popcll(mask)));
    base = __shfl(base, leader);
    if (want) queue[base + __popcll(mask & ((1ULL << lane_id()) - 1ULL))] = value;
}

// Next 64 queue entries for this wave (all lanes call it).  A plain read of
// the cursor first lets waves of a drained queue leave without an atomic.
__device__ unsigned wave_grab(unsigned* cursor, unsigned count) {
    unsigned base = 0;
    if (lane_id() == 0) {
        base = __hip_atomic_load(cursor, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (base < count) base = atomicAdd(cursor, 64U);
    }
    return __shfl(base, 0);
}

// The plain pass: pixel set-up, image-series jump and the plain-double block
// loop (field_core.h run_plain_blocks) on the same 16x16 tiles as
// fast_kernel.  Interior pixels finish here; every other pixel's loop-top
// state goes to the state store for fast_kernel.  This kernel holds only the
// plain loop's registers, not those of the general per-lane loop.
template <bool Planes, bool Stats>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) plain_kernel(RenderParams p) {
    const int wave = threadIdx.x >> 6;
    int px, py;
    block_pixel(p, px, py);
    const size_t index = static_cast<size_t>(py) * static_cast<size_t>(p.width) + px;
    const bool valid = px < p.width && py < p.height && (!p.point_re || index < static_cast<size_t>(p.point_count));
    int32_t* dlog = nullptr;
    if constexpr (Planes) {
        __shared__ int32_t derivative_log[4][core::kDerivativeLog];
        dlog = derivative_log[wave];
    }
    if (valid && static_cast<int64_t>(index) != p.reference_pixel) {
        core::Counters<Stats> st;
        st.clear();
        const SC dc = core::pixel_dc(p, px, py);
        core::PixelState s = core::init_pixel<Planes, Stats>(p, index, dc, st);
        if (core::plain_pixel<Planes, Stats>(p, index, s, dc, st, dlog)) p.state.iter[index] = p.max_iter;
        else core::save_state(p.state, index, s);
        if constexpr (Stats) {
            st.flush_work(p.work, index);
            st.flush(p.stats);
        }
    }
}

// The pixel on the reference (dc = 0) when the plain pass runs: set-up and
// the linear loop (its wave-uniform on-reference block loop) for that one
// pixel, on a side stream alongside plain_kernel, so that its long chain of
// single-lane blocks does not trail the plain pass.  A pixel it does not
// finish goes to the general queue like fast_kernel's.  One wave.
template <bool Planes, bool Stats>
__global__ void __launch_bounds__(64) reference_pixel_kernel(RenderParams p) {
    const size_t index = static_cast<size_t>(p.reference_pixel);
    const bool valid = lane_id() == 0;
    bool handoff = false;
    if (valid) {
        const int py = static_cast<int>(index / static_cast<size_t>(p.width));
        const int px = static_cast<int>(index - static_cast<size_t>(py) * static_cast<size_t>(p.width));
        core::Counters<Stats> st;
        st.clear();
        const SC dc = core::pixel_dc(p, px, py);
        core::PixelState s = core::init_pixel<Planes, Stats>(p, index, dc, st);
        if (core::run_linear<Planes, Stats>(p, index, s, dc, st, nullptr) == core::kHandoff) {
            core::save_state(p.state, index, s);
            handoff = true;
        }
        if constexpr (Stats) {
            st.flush_work(p.work, index);
            st.flush(p.stats);
        }
    }
    wave_push(p.state.general_queue, p.state.counters + 0, static_cast<uint32_t>(index), handoff);
}

// Occupancy targets measured on the 1920x1080 zoom-1e100 field: 6 waves per
// SIMD (80 VGPRs) with planes, 7 (72 VGPRs) without; the spills they cost sit
// in the per-pixel outer loop, not in the block loop.
template <bool Planes, bool Stats>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Planes ? 6 : 7, 8))) fast_kernel(RenderParams p) {
    const int wave = threadIdx.x >> 6;
    int px, py;
    block_pixel(p, px, py);
    const size_t index = static_cast<size_t>(py) * static_cast<size_t>(p.width) + px;
    bool valid = px < p.width && py < p.height && (!p.point_re || index < static_cast<size_t>(p.point_count));
    // pixels the plain pass finished, and the one reference_pixel_kernel renders
    if (valid && p.plain_pass && (static_cast<int64_t>(index) == p.reference_pixel || p.state.iter[index] >= p.max_iter))
        valid = false;
    bool handoff = false;
    int32_t* dlog = nullptr;
    if constexpr (Planes) {
        // one derivative log per wave, 16 KiB for the block's four waves
        __shared__ int32_t derivative_log[4][core::kDerivativeLog];
        dlog = derivative_log[wave];
    }
    if (valid) {
        core::Counters<Stats> st;
        st.clear();
        const SC dc = core::pixel_dc(p, px, py);
        core::PixelState s = p.plain_pass ? core::load_state(p.state, index)
                                          : core::init_pixel<Planes, Stats>(p, index, dc, st);
        if (core::run_linear<Planes, Stats>(p, index, s, dc, st, dlog) == core::kHandoff) {
            core::save_state(p.state, index, s);
            handoff = true;
        }
        if constexpr (Stats) {
            st.flush_work(p.work, index);
            st.flush(p.stats);
        }
    }
    wave_push(p.state.general_queue, p.state.counters + 0, static_cast<uint32_t>(index), handoff);
}

template <bool Planes, bool Stats>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FAV_GENERAL_WAVES, 8)))
general_kernel(RenderParams p) {
    const unsigned count = p.state.counters[0];
    for (;;) {
        const unsigned base = wave_grab(p.state.counters + 1, count);
        if (base >= count) break;
        const unsigned slot = base + static_cast<unsigned>(lane_id());
        bool to_tail = false;
        uint32_t index = 0;
        if (slot < count) {
            index = p.state.general_queue[slot];
            const int py = static_cast<int>(index / static_cast<uint32_t>(p.width));
            const int px = static_cast<int>(index - static_cast<uint32_t>(py) * static_cast<uint32_t>(p.width));
            core::Counters<Stats> st;
            st.clear();
            const SC dc = core::pixel_dc(p, px, py);
            core::PixelState s = core::load_state(p.state, index);
            if (core::run_general<Planes, Stats, false>(p, index, s, dc, st) == core::kTail) {
                core::save_state(p.state, index, s);
                to_tail = true;
            }
            if constexpr (Stats) {
                st.flush_work(p.work, index);
                st.flush(p.stats);
            }
        }
        wave_push(p.state.tail_queue, p.state.counters + 2, index, to_tail);
    }
}

template <bool Planes, bool Stats>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FAV_TAIL_WAVES, 8)))
tail_kernel(RenderParams p) {
    const unsigned count = p.state.counters[2];
    for (;;) {
        const unsigned base = wave_grab(p.state.counters + 3, count);
        if (base >= count) break;
        const unsigned slot = base + static_cast<unsigned>(lane_id());
        if (slot < count) {
            const uint32_t index = p.state.tail_queue[slot];
            const int py = static_cast<int>(index / static_cast<uint32_t>(p.width));
            const int px = static_cast<int>(index - static_cast<uint32_t>(py) * static_cast<uint32_t>(p.width));
            core::Counters<Stats> st;
            st.clear();
            const SC dc = core::pixel_dc(p, px, py);
            core::PixelState s = core::load_state(p.state, index);
            core::run_tail<Planes, Stats>(p, index, s, dc, st);
            if constexpr (Stats) {
                st.flush_work(p.work, index);
                st.flush(p.stats);
            }
        }
    }
}

// Clears the queue counters (a kernel rather than a memset so the sequence
// can be captured in a HIP graph) and starts the time-budget clock.
__global__ void prologue_kernel(unsigned* counters, unsigned long long* budget_start, int* budget_flag) {
    if (threadIdx.x < 4) counters[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        budget_start[0] = wall_clock64();
        budget_flag[0] = 0;
    }
}

template <bool Planes, bool Stats>
hipError_t launch_all(const RenderParams& p, int persistent_blocks, hipStream_t stream, hipStream_t side,
                      hipEvent_t fork, hipEvent_t join, hipEvent_t plain_start, hipEvent_t plain_stop) {
    prologue_kernel<<<1, 64, 0, stream>>>(p.state.counters, const_cast<unsigned long long*>(p.budget_start),
                                          p.budget_flag);
    const dim3 tiles((p.width + 15) / 16, (p.height + 15) / 16);
    const bool side_pixel = p.plain_pass && p.reference_pixel >= 0;
    hipError_t e = hipSuccess;
    if (side_pixel) {
        // after the prologue (queue counters cleared), beside the plain pass
        if ((e = hipEventRecord(fork, stream)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(side, fork, 0)) != hipSuccess) return e;
        reference_pixel_kernel<Planes, Stats><<<1, 64, 0, side>>>(p);
        if ((e = hipEventRecord(join, side)) != hipSuccess) return e;
    }
    const dim3 grid = p.tile_order ? dim3(4 * static_cast<unsigned>(p.tile_count)) : tiles;
    if (p.plain_pass) {
        if (plain_start && (e = hipEventRecord(plain_start, stream)) != hipSuccess) return e;
        plain_kernel<Planes, Stats><<<grid, dim3(256), 0, stream>>>(p);
        if (plain_stop && (e = hipEventRecord(plain_stop, stream)) != hipSuccess) return e;
    }
    fast_kernel<Planes, Stats><<<grid, dim3(256), 0, stream>>>(p);
    if (side_pixel && (e = hipStreamWaitEvent(stream, join, 0)) != hipSuccess) return e;
    general_kernel<Planes, Stats><<<dim3(persistent_blocks), dim3(256), 0, stream>>>(p);
    tail_kernel<Planes, Stats><<<dim3(persistent_blocks), dim3(256), 0, stream>>>(p);
    return hipSuccess;
}

}  // namespace

// Launches the field kernels on `stream` (and, for the pixel on the
// reference during a plain pass, on `side`, joined back before the general
// kernel through the events `fork` and `join`).  The caller validated the
// shapes: width*height <= 1e8 and every pointer in `p` (outputs, planes,
// state store) sized for width x height.  `persistent_blocks` is the grid of
// the queue-driven kernels.  With `plain_start` and `plain_stop` (may be
// null) the plain pass is bracketed by those events on `stream`.
hipError_t launch_field_kernel(const RenderParams& p, bool planes, bool stats, int persistent_blocks,
                               hipStream_t stream, hipStream_t side, hipEvent_t fork, hipEvent_t join,
                               hipEvent_t plain_start, hipEvent_t plain_stop) {
    hipError_t e;
    if (planes) {
        e = stats ? launch_all<true, true>(p, persistent_blocks, stream, side, fork, join, plain_start, plain_stop)
                  : launch_all<true, false>(p, persistent_blocks, stream, side, fork, join, plain_start, plain_stop);
    } else {
        e = stats ? launch_all<false, true>(p, persistent_blocks, stream, side, fork, join, plain_start, plain_stop)
                  : launch_all<false, false>(p, persistent_blocks, stream, side, fork, join, plain_start, plain_stop);
    }
    return e != hipSuccess ? e : hipGetLastError();
}

}  // namespace fav
