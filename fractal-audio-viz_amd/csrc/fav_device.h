// fav_device.h -- device-side table layouts and the launch interface shared
// by the host library (abi.cpp / reference_build.cpp, g++) and the gfx950
// kernels (render_kernel.hip, hipcc).  Plain structs and pointers only.
#pragma once

#include <cstddef>
#include <cstdint>

namespace fav {

constexpr int kMaxLevels = 16;
constexpr int kStatsFields = 39;  // RenderStats, renderer.cpp:4615-4643

// One compact linear BLA step (LinearBlaStep, renderer.cpp:8525-8530):
// delta' = A*delta + B*dc, valid while |delta|^2 < r2.  64 bytes, one
// cache line, so a wave-uniform lookup is one broadcast load.
struct alignas(64) LinearStepDev {
    double a_re, a_im, b_re, b_im, r2_m;
    int32_t a_e, b_e, r2_e, length;
    int32_t pad[2];
};
static_assert(sizeof(LinearStepDev) == 64, "linear step is one cache line");

// One compact cubic BLA step (FastBlaStep, renderer.cpp:8519-8523): nine
// ScaledComplex coefficients A..I of the bivariate map in (delta, dc).
struct alignas(64) CubicStepDev {
    double re[9];
    double im[9];
    double r2_m;
    int32_t e[9];
    int32_t r2_e, length;
    int32_t pad[5];
};
static_assert(sizeof(CubicStepDev) == 256, "cubic step is four cache lines");

// Level directory of one BLA hierarchy: level l holds size[l] steps starting
// at offset[l] in the flat step array.
struct LevelDir {
    int32_t count;
    int32_t size[kMaxLevels];
    int64_t offset[kMaxLevels];
};

// The per-index values the negligible-regime block loop reads, gathered in
// one 64-byte record so one wave-uniform scalar load fetches them all:
// Z_n, |Z_n|^2 as the loop-top norm, the escape margin, the negligible
// delta exponent and the loop-top flags.
struct alignas(64) OrbitFastDev {
    double z_re, z_im, zn_m, margin_m;
    int32_t z_e, zn_e, margin_e, small_exp, flags;
    int32_t pad[3];
};
static_assert(sizeof(OrbitFastDev) == 64, "orbit record is one cache line");

// The plain-double block loop (field_core.h run_plain_blocks) reads one
// linear BLA tier through one table built from it (table_image.cpp
// make_plain_tables): a 256-byte RungDev per block start r, so that every
// read of a block comes from one base address, and the block of the common
// case (slot 0, below) from the rung's first two cache lines.
//
// * r2[k]: the validity radius of the step of level top - k that starts at
//   r, as a plain double; top is the highest level aligned at r, at most the
//   table's level cap (RenderParams::plain_top).  A slot the reference's
//   lookup skips (length 1, non-finite or zero r2, past the level's end)
//   holds 0, so no lane fits it.  A step the loop may not apply carries the
//   sign bit, so that fitting it ends the loop: A or B outside the operand
//   bounds, a block shorter than 2^level, or an end index where |Z|^2 >= 3
//   (renderer.cpp:13422).
// * small_d: 2^small_exp[r], the negligible-delta bound at r, or 0 where the
//   loop must stop at r (the glitch or escape test fires for a negligible
//   delta, r is at or past the reference's escape, or no block starts at r:
//   r past the level-0 steps).
// * zn_d: |Z_r|^2 (0 below, +inf above the normal range).
// * a_re .. b_im: slot 0's A and B as plain doubles (their exact values);
//   step_low[k - 1]: those of slot k >= 1.
// * stride: 256 << top, the byte distance from this rung to slot 0's end.
//
// The slot-0 loop tests a block through bounds that imply the general
// tests:
// * fast_r2: |delta|^2 < fast_r2 implies every loop-top test of a slot-0
//   block at r: |Z|^2 >= |delta|^2 (fast_r2 <= the double after zn_d),
//   |delta|^2 below slot 0's radius, the step applicable and top >= 1 (else
//   fast_r2 = 0), and, in tables built for approximation order 3,
//   |delta|^2 < 2^-161 (renderer.cpp:13252-13273); all times (1 - 2^-48),
//   so that a |delta|^2 within 2^-50 of the reference's rounding of it
//   (the slot-0 loop forms it with one FMA) passes only when that does.
// * the product gap: |m|^2 / |c|^2 within [gap_lo, gap_hi] puts |A m| /
//   |B c| within 2^-58 .. 2^58, where gap_lo = |B|^2/|A|^2 2^-116 and
//   gap_hi = |B|^2/|A|^2 2^116, narrowed to [2^-600, 2^600] (0 and 0 when
//   slot 0 is not applicable).
// Bounds the loop compares with are kept as the high word of a double
// whose low word is 0 ("hi" below), so that the scalar unit can scale them
// by a power of two with one integer subtraction and take their minimum
// with one integer minimum: an upper bound is rounded down to that form,
// a lower bound up.
// * end_hi: the upper bound on |delta'|^2 of slot 0's block, min(small^2
//   (1 - 2^-48), fast_r2 of the end index's rung), where small is small_d
//   of the end index: |delta'|^2 below it puts the larger component of
//   delta' below that small_d and passes the next block's fast_r2 test.
//   The high word of -0.0 (no |delta'|^2 is below it) where that minimum is
//   outside [2^-900, 2^200].
// * gap_lo_end, gap_hi_end: gap_lo and gap_hi of the end index's rung, the
//   next block's product-gap bounds on |m'|^2 / |c|^2; gap_lo_own and
//   gap_hi_own: this rung's own (for the first block of a run).
struct alignas(32) PlainStepDev {
    double a_re, a_im, b_re, b_im;
};
static_assert(sizeof(PlainStepDev) == 32, "plain step is half a cache line");

constexpr int kRungSlots = 5;
struct alignas(256) RungDev {
    // first cache line: what a slot-0 block reads, and fast_r2
    int32_t end_hi;
    int32_t gap_hi_end;
    double a_re, a_im, b_re, b_im;
    uint32_t stride;
    int32_t gap_lo_end;
    double fast_r2;
    int32_t top;
    int32_t gap_lo_own;
    // second: the slot search
    int32_t gap_hi_own;
    int32_t pad;
    double zn_d;
    double r2[kRungSlots];
    double small_d;
    PlainStepDev step_low[kRungSlots - 1];
};
static_assert(sizeof(RungDev) == 256, "rung is four cache lines");
static_assert(offsetof(RungDev, gap_lo_own) == 60, "a slot-0 block reads the first cache line");
static_assert(offsetof(RungDev, step_low) == 128, "steps of slots 1.. are aligned");

// Device copy of one reference (orbit + tables), structure of arrays.  The
// per-index FloatExp parts of Z, the reference escape margin |Z|^2-R^2 and
// the glitch threshold 1e-7*|Z|^2 are pure functions of the orbit; they are
// evaluated once on the host with the reference's exact operations, so the
// kernel reads them instead of re-deriving them at every step.
struct OrbitDev {
    const double* re;        // ScaledComplex mantissas of Z_n
    const double* im;
    const int32_t* e;        // shared exponent of Z_n
    const double* zr_m;      // FloatExp parts of Re Z_n, Im Z_n
    const double* zi_m;
    const int32_t* zr_e;
    const int32_t* zi_e;
    const double* margin_m;  // |Z_n|^2 - R^2 (FloatExp)
    const int32_t* margin_e;
    const double* glitch_m;  // 1e-7*|Z_n|^2 (FloatExp; 0 disables the test)
    const int32_t* glitch_e;
    const double* real_double;  // ldexp(re, e): the plain-double tail orbit
    const double* imag_double;
    const double* zn_m;         // R^2 + (|Z|^2 - R^2): the loop-top norm for a negligible delta
    const int32_t* zn_e;
    const int32_t* small_exp;   // delta.e <= small_exp[n]: Z + delta, margin and norm are Z's own
    const int32_t* flags;       // bit 0: glitch test fires, bit 1: escape test fires (negligible delta)
    const OrbitFastDev* fast;   // the same values per index as 64-byte records
    int32_t size;
    int32_t map_end;
};

struct SeriesDev {
    const double* re;  // coefficients, index 0..count-1
    const double* im;
    const int32_t* e;
    int32_t count;
    int32_t iteration;
    int32_t available;
    double radius2_m;
    int32_t radius2_e;
};

// Resumable per-pixel state (structure of arrays, one slot per pixel) and
// the two hand-over queues of the three field kernels.  counters[0]/[1] are
// the general queue's length and read cursor, counters[2]/[3] the tail
// queue's.
struct StateStore {
    double *d_re, *d_im, *g_re, *g_im, *t_re, *t_im, *tn_m, *pn_m, *mg_m;
    int32_t *d_e, *g_e, *t_e, *tn_e, *pn_e, *mg_e, *ref, *iter, *misc;
    uint32_t* general_queue;
    uint32_t* tail_queue;
    unsigned* counters;
};

// Bytes of StateStore storage per pixel (9 doubles, 9 ints, 2 queue slots).
// The walk of the pixel on the reference (delta = dc = 0, field_core.h
// run_reference_blocks) through one BLA table at one block-length limit,
// one record per block start r: the step the lookup picks there with
// |delta|^2 = 0 when the limit is the full one, and what the walk reads of
// it, so that a block of that walk costs one scalar read instead of a chain
// of them.  length > 0 also says that the walk's other tests at r pass: no
// glitch or escape flag at r, r below the map end, and an end index inside
// the orbit without the endpoint flag.  length 0: the walk takes the
// lookup's path at r (table_image.cpp make_ref_walk).
struct alignas(64) RefWalkDev {
    double a_re, a_im, b_re, b_im;  // the step's A and B (ScaledComplex mantissas)
    int32_t a_e, b_e;
    int32_t length;
    int32_t pad[5];
};
static_assert(sizeof(RefWalkDev) == 64, "walk record is one cache line");

constexpr int32_t kOrbitGlitchFlag = 1;
constexpr int32_t kOrbitEscapeFlag = 2;
constexpr int32_t kOrbitEndpointFlag = 4;  // |Z_n|^2 >= 3: a BLA block may not end here
// LinearStepDev::pad[0] bits, set when the tables are flattened
constexpr int32_t kStepUsable = 1;           // length > 1 and 0 < r2 < inf
constexpr int32_t kStepNormalCoefficients = 2;  // A and B are normal ScaledComplex values
constexpr size_t kStateBytesPerPixel = 9 * 8 + 9 * 4 + 2 * 4;

// One alternate-formula linear BLA step (AlternateLinearBlaStep,
// renderer.cpp:8520-8525): d' = M d + P dc with M = c[0..3] and P = c[4..7]
// real 2x2 row-major FloatExp matrices, valid while |d|^2 < r2.
struct alignas(64) AltStepDev {
    double m[8];
    double r2_m;
    int32_t e[8];
    int32_t r2_e, length;
    int32_t pad[4];
};
static_assert(sizeof(AltStepDev) == 128, "alternate step is two cache lines");

// Launch parameters of the alternate-formula kernel (Julia, burning ship,
// tricorn; render_alternate_reference_impl, renderer.cpp:12064-12448).
struct AltRenderParams {
    const double* re;  // ScaledComplex orbit
    const double* im;
    const int32_t* e;
    const double* margin_m;  // MPFR escape margin |Z|^2 - R^2 per orbit entry
    const int32_t* margin_e;
    int32_t orbit_size;
    int32_t formula;
    const AltStepDev* steps;
    LevelDir dir;
    SeriesDev series;  // Julia only
    const double* x_off_m;
    const int32_t* x_off_e;
    const double* y_off_m;
    const int32_t* y_off_e;
    float* output;
    int* budget_flag;
    const unsigned long long* budget_start;
    unsigned long long budget_ticks;
    double parameter_re, parameter_im;  // c as ScaledComplex
    int32_t parameter_e;
    int32_t alt_start;
    double input_r2_m;
    int32_t input_r2_e;
    int32_t width, height, max_iter;
    int32_t escape_radius_mode;
    int32_t strict;
    int32_t cycle_enabled;
    int32_t bla_enabled;
    int32_t max_block;
    int32_t series_available;
    double output_bias;
    // schedule: 32x32 tiles launched in tile_order (null = row-major grid);
    // work, when not null, receives each pixel's loop passes
    const uint32_t* tile_order;
    int32_t tile_count;
    int32_t* work;
    // when not null: each pixel's escape iteration (max_iter inside, -1 for
    // a NaN pixel)
    int32_t* counts;
    // preview budget: when > 0 a pixel stops after this many loop passes
    // (output NaN, work = the budget)
    int32_t work_limit;
    // point renders: per-point dc (null for a pixel grid); valid pixel
    // indices are below point_count when point_re is set
    const double* point_re;
    const double* point_im;
    const int32_t* point_e;
    int32_t point_count;
};

// Everything one field launch needs, passed by value as the kernel argument.
struct RenderParams {
    StateStore state;
    OrbitDev orbit;
    SeriesDev series;
    const double* point_re;       // point renders: per-point dc (null for a pixel grid)
    const double* point_im;
    const int32_t* point_e;
    int32_t point_count;          // valid pixel indices are below this when point_re is set
    const LinearStepDev* linear;  // tier selected on the host (ordinary/deep)
    LevelDir linear_dir;
    const CubicStepDev* cubic;
    LevelDir cubic_dir;
    const double* x_off_m;  // per-column FloatExp offsets (width entries)
    const int32_t* x_off_e;
    const double* y_off_m;  // per-row FloatExp offsets (height entries)
    const int32_t* y_off_e;
    float* output;
    int64_t* plane_iteration;  // planes (all null when not requested)
    double* plane_phase;
    double* plane_de_x;
    double* plane_de_y;
    double* plane_test1;
    double* plane_test2;
    unsigned long long* stats;   // kStatsFields counters or null
    int* budget_flag;            // set once the time budget has expired
    const unsigned long long* budget_start;  // wall_clock64 at launch start
    unsigned long long budget_ticks;         // 0 = no time budget
    double output_bias;
    double pixel_spacing_m;
    int32_t pixel_spacing_e;
    int32_t width, height;
    int32_t max_iter;
    int32_t escape_radius_mode;
    int32_t strict;
    int32_t allow_recovery;
    int32_t disable_bla;
    int32_t cycle_enabled;
    int32_t strict_cycle;
    int32_t max_bla_length;
    int32_t max_linear_length;
    int32_t approximation_order;
    int32_t fast_linear;  // linear-BLA fast kernel on (off with cycle checks or a time budget)
    // plain-double block loop tables of the selected linear tier (rungs null:
    // the loop is off and the scaled-operand block loop runs instead)
    const RungDev* rungs;
    // 1: plain_kernel ran first; fast_kernel resumes each pixel from the
    // state store (state.iter >= max_iter marks a pixel it finished)
    int32_t plain_pass;
    int32_t plain_top;  // level cap the rung tables were built for
    // with the plain pass: the one grid pixel whose dc is zero (it sits on
    // the reference), rendered by reference_pixel_kernel on a side stream
    // while the plain pass runs; -1 when there is none
    int64_t reference_pixel;
    // schedule: when tile_order is set, plain_kernel and fast_kernel run a
    // 1-D grid of four 16x16 blocks per 32x32 tile, tiles in tile_order
    // (tile_count row-major 32x32 tile indices); null = the 2-D grid
    const uint32_t* tile_order;
    int32_t tile_count;
    // with statistics on: each pixel's loop passes (BLA blocks, exact,
    // replayed and tail steps) added to work[pixel] by every kernel that
    // runs it (the caller zeroes it); null = no work map
    int32_t* work;
    // the walk of the pixel on the reference through the table of the
    // approximation order at the block-length limit ref_walk_limit (null:
    // every block of that walk takes the lookup)
    const RefWalkDev* ref_walk;
    int32_t ref_walk_limit;
};

namespace direct { struct DirectParams; }

// One batched launch over several fields (batch_kernel.hip).
struct BatchLaunch {
    const RenderParams* fields;  // field_count parameter sets in HBM
    // field_count direct-renderer parameter sets (zoom below 1e6: plain
    // doubles, no reference orbit), width 0 for a perturbation field; null
    // when the batch has no such field
    const direct::DirectParams* direct_fields;
    const uint32_t* tiles;       // tile_count codes: field << 20 | tile index (row-major 16x16 tiles)
    unsigned tile_count;
    unsigned field_count;
    unsigned long long* general_queue;  // (field << 32 | pixel) hand-offs, one slot per batch pixel
    unsigned long long* tail_queue;
    unsigned* counters;  // [0] fast tile cursor, [1]/[2] general count/cursor, [3]/[4] tail count/cursor,
                         // [5] plain tile cursor
    // 1: the perturbation fields render through the FP32 floatexp pixel
    // (fx_core.h) in one persistent kernel over the tile queue
    int32_t floatexp;
};

}  // namespace fav
