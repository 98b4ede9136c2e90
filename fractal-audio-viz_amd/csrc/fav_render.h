// fav_render.h -- host interface between the C ABI layer (abi.cpp) and the
// device launch code (device_host.cpp).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include <cstdint>

#include "../../include/fractal_hip.h"
#include "fav_host.h"

namespace fav {

// Sets the text fractal_last_error() returns on this thread.
void set_last_error(const std::string& message);

struct FieldRequest {
    float* output;                 // host or device pointer (see device_pointers)
    int width, height;
    const char* zoom_text;
    int max_iter;
    int series_order;
    int series_block;
    FractalRenderOptions options;  // validated copy
    FractalRenderPlanes* planes;   // null when no planes are requested
    uint64_t* stats;               // kStatsFields host counters or null
    bool device_pointers;          // output/planes already live in HBM
    // point renders (renderer.cpp:20019-20186): point_count normalised dc
    // offsets; the field is then point_count values, laid out by the device
    // as a width x height grid of whole rows (the last row partly used)
    const SC* points = nullptr;
    int point_count = 0;
    FE point_radius{};             // sqrt of the largest |dc|^2
    // Mandelbrot schedule hooks: 32x32 tiles launched in tile_order (checked
    // by the caller; null = the 2-D grid) and each pixel's loop passes
    // written to `work` (host or device like `output`; null = none)
    const uint32_t* tile_order = nullptr;
    size_t tile_count = 0;
    int32_t* work = nullptr;
};

struct LastTiming {
    double kernel_ms;   // HIP-event time of the field kernel on the library stream
    double total_ms;    // wall time from launch to results on the host
    double pixels;
    double plain_ms;    // HIP-event time of the plain pass alone (0 when it did not run)
};

void render_field(ReferenceContext& context, const FieldRequest& request);

// FP32 floatexp field of a handle of any formula: request.output (field)
// and `counts` (escape iteration per pixel, max_iter inside, -1 unresolved)
// are host or device pointers as request.device_pointers says; no planes.
// Mandelbrot map: with double_tail the pixels finish in the FP64
// plain-double tail where the FP64 path enters it.
void render_field_floatexp(ReferenceContext& context, const FieldRequest& request, int32_t* counts,
                           bool double_tail = true);
void render_alternate_floatexp(ReferenceContext& context, const FieldRequest& request, int32_t* counts);

// A prepared batch of Mandelbrot fields rendered by one launch sequence
// (batch_kernel.hip).  Owns the device copies of every field's launch
// parameters, offsets, state stores and the shared queues; the references
// stay alive while the batch does.
struct BatchFieldRequest {
    std::shared_ptr<ReferenceContext> context;
    FieldRequest request;  // device_pointers output, no planes, no stats
    // zoom below 1e6: rendered by the direct double renderer at the
    // reference's centre (renderer.cpp:19920-19960), at this zoom
    bool direct = false;
    long double direct_zoom = 0.0L;
};

class PreparedBatch;
std::shared_ptr<PreparedBatch> prepare_batch(std::vector<BatchFieldRequest> fields);
// tile order: 0 field-major, 1 fields interleaved tile by tile, 2 by the
// given per-tile cost, heaviest first (costs in field-major tile order)
void batch_set_order(PreparedBatch& batch, int order, const float* costs, size_t cost_count);
size_t batch_tile_count(const PreparedBatch& batch);
// floatexp: the perturbation fields render through the FP32 floatexp pixel
// in one persistent kernel (fast_blocks_per_cu blocks per CU); false: the
// FP64 launch sequence
void batch_set_floatexp(PreparedBatch& batch, bool floatexp);
// one render; fast_blocks_per_cu <= 0 keeps the current setting
void batch_render(PreparedBatch& batch, int fast_blocks_per_cu, bool use_graph);

// A segment of Mandelbrot fields whose references are built here: worker
// threads build each field's MPFR reference and stage its tables in pinned
// memory; the calling thread copies a staged reference to HBM on the side
// stream (hipMemcpyAsync) and renders it on the compute stream while the
// workers build the next ones.  With pipelined false every reference is
// built before the first render (the unpipelined baseline).  Outputs are
// device pointers; a field below zoom 1e6 renders through the direct
// renderer at its centre, as a handle render does.
struct SegmentField {
    std::string x, y, zoom;
    float* output;  // device memory, width * height floats
    int width, height, max_iter, precision_bits, series_order, series_block;
};
struct SegmentTiming {
    double wall_ms = 0.0;        // the whole call
    double build_ms = 0.0;       // summed host time of the reference builds and staging
    double render_ms = 0.0;      // summed HIP-event time of the field launches
    double built_ms = 0.0;       // wall time until the last reference was staged
};
void render_segment(const std::vector<SegmentField>& fields, const FractalRenderOptions& options, int threads,
                    bool pipelined, SegmentTiming& timing);

// Throws unless tile_order lists each of the ceil(w/32)*ceil(h/32)
// row-major 32x32 tiles exactly once.
void check_tile_order(const uint32_t* tile_order, size_t tile_count, int width, int height);

// Alternate-formula field (Julia, burning ship, tricorn) of a reference
// handle; request.planes must be null.
// counts (host, when set): each pixel's escape iteration, max_iter inside,
// -1 for a NaN pixel.
void render_alternate(ReferenceContext& context, const FieldRequest& request, const uint32_t* tile_order = nullptr,
                      size_t tile_count = 0, int32_t* work = nullptr, int32_t* counts = nullptr);
// Learned schedule of an alternate-formula field: the quarter-size preview
// with `work_limit` passes per pixel, its tile features (host `features`,
// tiles * 65 floats) ...
void alternate_preview_features(ReferenceContext& context, const FieldRequest& request, int work_limit,
                                float* features);
// ... and the whole pipeline: preview, features, cost model (host weights),
// heaviest-first order (copied to order_out when set) and the ordered
// render.  timings (5 doubles, when set): preview + features ms, model ms,
// ordered render ms, wall ms up to the ordered launch, total wall ms.
void render_alternate_learned(ReferenceContext& context, const FieldRequest& request, const float* weights,
                              int work_limit, uint32_t* order_out, double* timings);
// The same two for a Mandelbrot perturbation field (zoom >= 1e6): the
// preview runs the field kernels at quarter size with cap
// min(max_iter, work_limit) and its per-pixel loop passes as the work map;
// the ordered render is render_field with the order as request.tile_order.
void field_preview_features(ReferenceContext& context, const FieldRequest& request, int work_limit,
                            float* features);
void render_field_learned(ReferenceContext& context, const FieldRequest& request, const float* weights,
                          int work_limit, uint32_t* order_out, double* timings);

// Learned tile-cost model (sched_mlp.h shapes); every pointer is device
// memory and the caller checks n and steps.  loss_history gets one mean
// squared error per training step (the loss before that step's update).
void mlp_forward(const float* weights, const float* x, int n, float* y);
void mlp_gradient(const float* weights, const float* x, const float* t, int n, float* grad, float* loss);
void mlp_train(float* weights, float* adam_m, float* adam_v, const float* x, const float* t, int n, int steps,
               float learning_rate, int first_step, float* loss_history);

// KFP colour transfer on the device.  `params` carries host pointers for
// field / lut / output unless device_pointers is set (then field and output
// are device pointers; the LUT is always a host pointer).
namespace kfp { struct KfpParams; }
namespace direct { struct DirectParams; }

// Raw KFP atlas colour pass (atlas_core.h): `compose` carries host pointers
// to the parent (parent_count floats) and child (child_count floats) tiles
// and the geometry; axes[0..3] are the parent x / y and child x / y maps;
// `colour` the KFP parameters (its field and output are replaced).  RGB is
// copied to the host buffer `output`.
namespace atlas { struct ComposeParams; struct Axis; }
void atlas_colourise_kfp_raw(const atlas::ComposeParams& compose, size_t parent_count, size_t child_count,
                             const atlas::Axis* axes, const kfp::KfpParams& colour, uint8_t* output);


// Plane-aware KFP atlas colour pass (atlas_core.h compose_planes_pixel,
// then the KFP colour kernel).  `compose` carries the geometry (its
// pointers are filled here); tiles[0] is the primary tile, tiles[1] the
// secondary one (field null when absent); axes are primary x / y and
// secondary x / y; `colour` the KFP constants for the composed planes;
// `texture` the texture image the composed planes carry (or null).
namespace atlas { struct PlanesCompose; }
struct PlaneTile {
    const float* field = nullptr;
    int width = 0, height = 0;
    const FractalKfpPlanes* planes = nullptr;
};
void atlas_colourise_kfp_planes(const atlas::PlanesCompose& compose, const PlaneTile* tiles, const atlas::Axis* axes,
                                const kfp::KfpParams& colour, const FractalKfpPlanes* texture, uint8_t* output);

// Explicit-rectangle KFP atlas (kfp_core.h rect_atlas_pixel): `parent` and
// `child` carry host field / LUT / output pointers (child.field null when
// there is no child); edge_x / edge_y the child's seam weights (empty
// without a feather).
namespace kfp { struct RectAtlas; }
void atlas_colourise_kfp_rect(const kfp::KfpParams& parent, const kfp::KfpParams& child, const kfp::RectAtlas& g,
                              const std::vector<float>& edge_x, const std::vector<float>& edge_y);

// Ordinary-palette colouriser and crop / atlas compositors (aurora_core.h).
// `params` carries the mode, geometry and palette controls; the sources,
// axes (parent x / y, child x / y; empty when unused), seam weights and
// palette are host arrays uploaded here; results go to the host buffers
// out_field (float modes) or out_rgb (RGB modes; kAccents reads it too).
namespace aurora { struct Params; }
struct AuroraInputs {
    const float* parent_field = nullptr;
    size_t parent_field_count = 0;
    const float* child_field = nullptr;
    size_t child_field_count = 0;
    const uint8_t* parent_rgb = nullptr;
    size_t parent_rgb_bytes = 0;
    const uint8_t* child_rgb = nullptr;
    size_t child_rgb_bytes = 0;
    const atlas::Axis* axes[4] = {nullptr, nullptr, nullptr, nullptr};
    const std::vector<float>* edge_x = nullptr;
    const std::vector<float>* edge_y = nullptr;
    const std::vector<uint8_t>* palette = nullptr;
    float* out_field = nullptr;
    uint8_t* out_rgb = nullptr;
};
void aurora_run(const aurora::Params& params, const AuroraInputs& inputs);

// Shallow direct render on the device: x/y are the per-column/per-row
// coordinates; `params` output/plane pointers are host pointers unless
// device_pointers is set.
void render_direct_on_device(const direct::DirectParams& params, const double* x, const double* y, bool planes,
                             bool device_pointers);
void colourise_kfp_on_device(const kfp::KfpParams& params, bool device_pointers,
                             const FractalKfpPlanes* planes = nullptr);
// colourise_kfp_on_device with device pointers and no device allocation
// (scratch in a file-scope array; the LUT is a host pointer staged there).
void colourise_kfp_device_noalloc(const kfp::KfpParams& params);
LastTiming last_timing();
bool device_available();

}  // namespace fav
