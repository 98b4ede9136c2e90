// batch_kernel.hip -- many deep-zoom fields (each with its own reference
// tables, size, zoom and cap) rendered by one launch sequence over a device
// tile queue: the atlas keyframes of a segment in one go.
//
//   batch_plain_kernel    persistent; each 256-thread block takes the next
//                         16x16 tile from the queue (any field), stages that
//                         field's launch parameters in LDS when the field
//                         changes, and runs the plain pass of plain_kernel
//                         (set-up, image-series jump, plain-double block
//                         loop): interior pixels finish, the others leave
//                         their loop-top state in the field's state store.
//                         A tile of a shallow field (zoom below 1e6) runs the
//                         direct renderer's per-pixel code instead, as a
//                         single-field render of that zoom does
//                         (renderer.cpp:19920-19960).
//   batch_fast_kernel     the same tile walk: the fast path of fast_kernel
//                         for the pixels the plain pass left (all pixels of
//                         a field without the plain pass).  Hand-offs go to
//                         one queue shared by all fields as (field, pixel)
//                         pairs.
//   batch_general_kernel  the general loop over that queue.
//   batch_tail_kernel     the plain-double tail.
//
// The per-pixel code is field_core.h (direct_core.h for shallow fields),
// unchanged, so every field's result is bit-identical to its single-field
// render.  The tile order (field-major,
// interleaved, or by predicted cost) only changes the schedule.
#include <hip/hip_runtime.h>

#include "direct_core.h"
#include "fx_core.h"

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

// One tile of the queue for a 256-thread block: which field, and this
// thread's pixel in it.  The block's first thread claims the next tile from
// `cursor`; every thread returns after a barrier with the same tile.
struct TileClaim {
    int field;
    int tile;
    bool done;
};

__device__ TileClaim claim_tile(const BatchLaunch& q, unsigned* cursor, unsigned& slot) {
    if (threadIdx.x == 0) slot = atomicAdd(cursor, 1U);
    __syncthreads();
    const unsigned t = slot;
    if (t >= q.tile_count) return TileClaim{0, 0, true};
    const uint32_t code = q.tiles[t];
    return TileClaim{static_cast<int>(code >> kTileBits), static_cast<int>(code & ((1U << kTileBits) - 1U)), false};
}

// Copies field `field`'s RenderParams into the block's LDS copy when another
// field is staged there (the branch is uniform over the block).
__device__ void stage_field(const BatchLaunch& q, RenderParams& sp, int& staged, int field) {
    if (field == staged) return;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(q.fields + field);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&sp);
    for (unsigned i = threadIdx.x; i < sizeof(RenderParams) / 4; i += 256) dst[i] = src[i];
    __syncthreads();
    staged = field;
}

__device__ bool is_direct(const BatchLaunch& q, int field) {
    return q.direct_fields != nullptr && q.direct_fields[field].width > 0;
}

// This thread's pixel of a 16x16 tile (four 8x8 waves, as fast_kernel).
__device__ void tile_pixel(int width, int tile, int& px, int& py) {
    const int wave = threadIdx.x >> 6;
    const int lane = lane_id();
    const int tiles_x = (width + 15) / 16;
    px = (tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    py = (tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
batch_plain_kernel(BatchLaunch q) {
    __shared__ RenderParams sp;
    __shared__ unsigned slot;
    int staged = -1;  // field whose parameters sit in sp (the same in every thread)
    for (;;) {
        const TileClaim c = claim_tile(q, q.counters + 5, slot);
        if (c.done) break;
        int px, py;
        if (is_direct(q, c.field)) {
            const direct::DirectParams& d = q.direct_fields[c.field];
            tile_pixel(d.width, c.tile, px, py);
            if (px < d.width && py < d.height) direct::pixel<0, false, false>(d, px, py);
        } else {
            stage_field(q, sp, staged, c.field);
            const RenderParams& p = sp;
            tile_pixel(p.width, c.tile, px, py);
            if (p.plain_pass && px < p.width && py < p.height) {
                const size_t index = static_cast<size_t>(py) * static_cast<size_t>(p.width) + px;
                core::Counters<false> st;
                const SC dc = core::pixel_dc(p, px, py);
                core::PixelState s = core::init_pixel<false, false>(p, index, dc, st);
                if (core::plain_pixel<false, false>(p, index, s, dc, st, nullptr)) p.state.iter[index] = p.max_iter;
                else core::save_state(p.state, index, s);
            }
        }
        __syncthreads();  // every wave is done with sp and slot before the next tile
    }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8)))
batch_fast_kernel(BatchLaunch q) {
    __shared__ RenderParams sp;
    __shared__ unsigned slot;
    int staged = -1;
    for (;;) {
        const TileClaim c = claim_tile(q, q.counters + 0, slot);
        if (c.done) break;
        if (is_direct(q, c.field)) {  // rendered by the plain pass
            __syncthreads();
            continue;
        }
        stage_field(q, sp, staged, c.field);
        const RenderParams& p = sp;
        int px, py;
        tile_pixel(p.width, c.tile, px, py);
        const size_t index = static_cast<size_t>(py) * static_cast<size_t>(p.width) + px;
        // pixels the plain pass finished are done
        const bool valid = px < p.width && py < p.height && !(p.plain_pass && p.state.iter[index] >= p.max_iter);
        bool handoff = false;
        if (valid) {
            core::Counters<false> st;
            const SC dc = core::pixel_dc(p, px, py);
            core::PixelState s = p.plain_pass ? core::load_state(p.state, index)
                                              : core::init_pixel<false, false>(p, index, dc, st);
            if (core::run_linear<false, false>(p, index, s, dc, st, nullptr) == core::kHandoff) {
                core::save_state(p.state, index, s);
                handoff = true;
            }
        }
        wave_push64(q.general_queue, q.counters + 1,
                    (static_cast<unsigned long long>(c.field) << 32) | static_cast<unsigned long long>(index), handoff);
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

// The floatexp batch: one persistent kernel over the tile queue, each
// perturbation tile through the FP32 floatexp pixel (fx_core.h, with the
// FP64 plain-double tail where the FP64 path enters it), each shallow tile
// through the direct renderer.  No hand-off queues: the floatexp pixel runs
// its whole loop in one kernel.
// 5 waves per SIMD (102 registers, from the compiler's 138 and 3 waves):
// the 3840x2160 zoom-1e100 floatexp batch 77.7 -> 55.3 ms; 4 waves 62.8,
// 6 and 8 waves within 2% of 5.
#ifndef FAV_BATCH_FX_WAVES
#define FAV_BATCH_FX_WAVES 5
#endif
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FAV_BATCH_FX_WAVES, 8)))
batch_fx_kernel(BatchLaunch q) {
    __shared__ RenderParams sp;
    __shared__ unsigned slot;
    int staged = -1;
    for (;;) {
        const TileClaim c = claim_tile(q, q.counters + 5, slot);
        if (c.done) break;
        int px, py;
        if (is_direct(q, c.field)) {
            const direct::DirectParams& d = q.direct_fields[c.field];
            tile_pixel(d.width, c.tile, px, py);
            if (px < d.width && py < d.height) direct::pixel<0, false, false>(d, px, py);
        } else {
            stage_field(q, sp, staged, c.field);
            const RenderParams& p = sp;
            tile_pixel(p.width, c.tile, px, py);
            if (px < p.width && py < p.height) fx::pixel(p, fx::FxOut{p.output, nullptr, 1}, px, py);
        }
        __syncthreads();  // every wave is done with sp and slot before the next tile
    }
}

// Clears the tile cursors ([0] fast, [5] plain) and the queue counters (a
// kernel, so the launch sequence can be captured in a HIP graph).
__global__ void batch_prologue(unsigned* counters) {
    if (threadIdx.x < 8) counters[threadIdx.x] = 0;
}

}  // namespace

// One batch: q.fields holds field_count RenderParams with tables, offsets,
// outputs and state stores bound (with q.floatexp, one persistent kernel of
// fast_blocks blocks renders every tile in floatexp) (q.direct_fields, when set, the direct
// parameters of the shallow fields); q.tiles the tile_count tile codes
// (field << 20 | tile); the queues hold one slot per pixel of the batch.
hipError_t launch_batch(const BatchLaunch& q, int fast_blocks, int persistent_blocks, hipStream_t stream) {
    batch_prologue<<<1, 64, 0, stream>>>(q.counters);
    if (q.floatexp) {
        batch_fx_kernel<<<dim3(fast_blocks), dim3(256), 0, stream>>>(q);
        return hipGetLastError();
    }
    batch_plain_kernel<<<dim3(persistent_blocks), dim3(256), 0, stream>>>(q);
    batch_fast_kernel<<<dim3(fast_blocks), dim3(256), 0, stream>>>(q);
    batch_general_kernel<<<dim3(persistent_blocks), dim3(256), 0, stream>>>(q);
    batch_tail_kernel<<<dim3(persistent_blocks), dim3(256), 0, stream>>>(q);
    return hipGetLastError();
}

}  // namespace fav
