// table_image.h -- host-side flat images of a reference's render tables and
// the scalar launch parameters of one field render.  Nothing here touches
// the HIP runtime: device_host.cpp copies the images into HBM, and the
// test-only CPU harness (host_check.cpp) points the same kernel code at
// them directly.
#pragma once

#include <vector>

#include "fav_render.h"

namespace fav {

struct TableImage {
    std::vector<double> re, im;  // ScaledComplex mantissas of the orbit
    std::vector<int32_t> e;
    std::vector<LinearStepDev> linear[2];
    LevelDir linear_dir[2]{};
    std::vector<CubicStepDev> cubic;
    LevelDir cubic_dir{};
    std::vector<double> series_re, series_im;
    std::vector<int32_t> series_e;
    std::vector<OrbitFastDev> fast;  // per-index records of the block loop
    std::vector<AltStepDev> alt;  // alternate formulas only
    LevelDir alt_dir{};
};

void make_table_image(const ReferenceContext& context, TableImage& image);

// The plain-double block loop's rung table (fav_device.h) of linear tier
// `tier`, with rungs starting at level min(aligned level, top_cap), for
// approximation order 3 (order3) or 1; empty when the orbit has no
// per-index records.
void make_plain_tables(const ReferenceContext& context, int tier, int top_cap, bool order3,
                       std::vector<RungDev>& rungs);

// The walk of the pixel on the reference (fav_device.h RefWalkDev) through
// linear tier `table` (0 or 1) or the cubic table (2) at block-length limit
// `limit`; empty when the orbit has no per-index records.
void make_ref_walk(const ReferenceContext& context, int table, int limit, std::vector<RefWalkDev>& walk);

// Per-render parameters: everything in RenderParams except the table, output
// and scratch pointers, plus the per-column / per-row FloatExp offsets
// (width entries, then height entries).
struct PreparedRender {
    RenderParams params{};
    int linear_tier = 0;
    std::vector<double> offset_m;
    std::vector<int32_t> offset_e;
};

void prepare_render(const ReferenceContext& context, const FieldRequest& request, PreparedRender& out);

// Alternate formulas: the scalar part of AltRenderParams (table and output
// pointers left null) and the per-column / per-row FloatExp offsets
// (renderer.cpp:12082-12110, 12126-12140).
struct PreparedAlternate {
    AltRenderParams params{};
    std::vector<double> offset_m;
    std::vector<int32_t> offset_e;
};

void prepare_alternate(const ReferenceContext& context, const FieldRequest& request, PreparedAlternate& out);

// Points p.state at per-pixel arrays inside `block`, which must hold
// pixels * kStateBytesPerPixel + 20 * 256 bytes (each array is 256-byte aligned).
void bind_state_store(RenderParams& p, void* block, size_t pixels);

}  // namespace fav
