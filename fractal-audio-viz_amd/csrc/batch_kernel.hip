// batch_kernel.hip -- many deep-zoom fields (each with its own reference
// tables, size, zoom and cap) rendered by one launch sequence over a device
// tile queue: the atlas keyframes of a segment in one go.
//
//   batch_fast_kernel     persistent; each 256-thread block takes the next
//                         16x16 tile from the queue (any field), stages that
//                         field's launch parameters in LDS when the field
//                         changes, and runs the same per-pixel fast path as
//                         fast_kernel.  Hand-offs go to one queue shared by
//                         all fields as (field, pixel) pairs.  A tile of a
//                         shallow field (zoom below 1e6) runs the direct
//                         renderer's per-pixel code instead, as a
//                         single-field render of that zoom does
//                         (renderer.cpp:19920-19960).
//   batch_general_kernel  the general loop over that queue.
//   batch_tail_kernel     the plain-double tail.
//
// The per-pixel code is field_core.h (direct_core.h for shallow fields),
// unchanged, so every field's result is bit-identical to its single-field
// render.  The tile order (field-major,
// interleaved, or by predicted cost) only changes the schedule.
#include <hip/hip_runtime.h>

#include "direct_core.h"

namespace fav {
namespace {

__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x & 63); }

__device__ void wave_push64(unsigned long long* queue, unsigned* count, unsigned long long value, bool want) {
    const unsigned long long mask = __ballot(want);
    if (mask == 0) return;
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    unsigned base = 0;
    if (lane_id() == leader) base = atomicAdd(count, static_cast<unsigned>(__popcll(mask)));
    base = __shfl(base, leader);
    if (want) queue[base + __popcll(mask & ((1ULL << lane_id()) - 1ULL))] = value;
}

__device__ unsigned wave_grab(unsigned* cursor, unsigned count) {
    unsigned base = 0;
    if (lane_id() == 0) {
        base = __hip_atomic_load(cursor, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (base < count) base = atomicAdd(cursor, 64U);
    }
    return __shfl(base, 0);
}

constexpr int kTileBits = 20;  // tile index within its field (3840x2160 has 32,400 16x16 tiles)

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8)))
batch_fast_kernel(BatchLaunch q) {
    __shared__ RenderParams sp;
    __shared__ unsigned slot;
    const int wave = threadIdx.x >> 6;
    const int lane = lane_id();
    int staged = -1;  // field whose parameters sit in sp (the same in every thread)
    for (;;) {
        if (threadIdx.x == 0) slot = atomicAdd(q.counters + 0, 1U);
        __syncthreads();
        const unsigned t = slot;
        if (t >= q.tile_count) break;
        const uint32_t code = q.tiles[t];
        const int field = static_cast<int>(code >> kTileBits);
        const int tile = static_cast<int>(code & ((1U << kTileBits) - 1U));
        if (q.direct_fields != nullptr && q.direct_fields[field].width > 0) {
            const direct::DirectParams& d = q.direct_fields[field];
            const int tiles_x = (d.width + 15) / 16;
            const int px = (tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
            const int py = (tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
            if (px < d.width && py < d.height) direct::pixel<0, false, false>(d, px, py);
            __syncthreads();  // every wave has read slot before the next tile
            continue;
        }
        if (field != staged) {
            const uint32_t* src = reinterpret_cast<const uint32_t*>(q.fields + field);
            uint32_t* dst = reinterpret_cast<uint32_t*>(&sp);
            for (unsigned i = threadIdx.x; i < sizeof(RenderParams) / 4; i += 256) dst[i] = src[i];
            __syncthreads();
            staged = field;
        }
        const RenderParams& p = sp;
        const int tiles_x = (p.width + 15) / 16;
        const int px = (tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
        const int py = (tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
        const bool valid = px < p.width && py < p.height;
        const size_t index = static_cast<size_t>(py) * static_cast<size_t>(p.width) + px;
        bool handoff = false;
        if (valid) {
            core::Counters<false> st;
            const SC dc = core::pixel_dc(p, px, py);
            core::PixelState s = core::init_pixel<false, false>(p, index, dc, st);
            if (core::run_linear<false, false>(p, index, s, dc, st, nullptr) == core::kHandoff) {
                core::save_state(p.state, index, s);
                handoff = true;
            }
        }
        wave_push64(q.general_queue, q.counters + 1,
                    (static_cast<unsigned long long>(field) << 32) | static_cast<unsigned long long>(index), handoff);
        __syncthreads();  // every wave is done with sp and slot before the next tile
    }
}

__global__ void __launch_bounds__(256) batch_general_kernel(BatchLaunch q) {
    const unsigned count = q.counters[1];
    for (;;) {
        const unsigned base = wave_grab(q.counters + 2, count);
        if (base >= count) break;
        const unsigned slot = base + static_cast<unsigned>(lane_id());
        bool to_tail = false;
        unsigned long long code = 0;
        if (slot < count) {
            code = q.general_queue[slot];
            const RenderParams& p = q.fields[code >> 32];
            const uint32_t index = static_cast<uint32_t>(code);
            const int py = static_cast<int>(index / static_cast<uint32_t>(p.width));
            const int px = static_cast<int>(index - static_cast<uint32_t>(py) * static_cast<uint32_t>(p.width));
            core::Counters<false> st;
            const SC dc = core::pixel_dc(p, px, py);
            core::PixelState s = core::load_state(p.state, index);
            if (core::run_general<false, false, false>(p, index, s, dc, st) == core::kTail) {
                core::save_state(p.state, index, s);
                to_tail = true;
            }
        }
        wave_push64(q.tail_queue, q.counters + 3, code, to_tail);
    }
}

__global__ void __launch_bounds__(256) batch_tail_kernel(BatchLaunch q) {
    const unsigned count = q.counters[3];
    for (;;) {
        const unsigned base = wave_grab(q.counters + 4, count);
        if (base >= count) break;
        const unsigned slot = base + static_cast<unsigned>(lane_id());
        if (slot < count) {
            const unsigned long long code = q.tail_queue[slot];
            const RenderParams& p = q.fields[code >> 32];
            const uint32_t index = static_cast<uint32_t>(code);
            const int py = static_cast<int>(index / static_cast<uint32_t>(p.width));
            const int px = static_cast<int>(index - static_cast<uint32_t>(py) * static_cast<uint32_t>(p.width));
            core::Counters<false> st;
            const SC dc = core::pixel_dc(p, px, py);
            core::PixelState s = core::load_state(p.state, index);
            core::run_tail<false, false>(p, index, s, dc, st);
        }
    }
}

// Clears the tile cursor and the queue counters (a kernel, so the launch
// sequence can be captured in a HIP graph).
__global__ void batch_prologue(unsigned* counters) {
    if (threadIdx.x < 8) counters[threadIdx.x] = 0;
}

}  // namespace

// One batch: q.fields holds field_count RenderParams with tables, offsets,
// outputs and state stores bound (q.direct_fields, when set, the direct
// parameters of the shallow fields); q.tiles the tile_count tile codes
// (field << 20 | tile); the queues hold one slot per pixel of the batch.
hipError_t launch_batch(const BatchLaunch& q, int fast_blocks, int persistent_blocks, hipStream_t stream) {
    batch_prologue<<<1, 64, 0, stream>>>(q.counters);
    batch_fast_kernel<<<dim3(fast_blocks), dim3(256), 0, stream>>>(q);
    batch_general_kernel<<<dim3(persistent_blocks), dim3(256), 0, stream>>>(q);
    batch_tail_kernel<<<dim3(persistent_blocks), dim3(256), 0, stream>>>(q);
    return hipGetLastError();
}

}  // namespace fav
