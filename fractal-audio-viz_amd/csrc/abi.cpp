// abi.cpp -- extern "C" entry points of libfractal_hip.so.
//
// Names, argument meaning, validation order, return codes and the
// thread-local error text follow the reference's C ABI (renderer.h,
// renderer.cpp:15870-15924, 19568-20474).  No exception crosses the ABI.
#include <algorithm>
#include <cstddef>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <locale>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>

#include "../../include/fractal_hip.h"
#include "fav_host.h"
#include "fav_render.h"
#include "sched_mlp.h"
#include "atlas_core.h"
#include "kfp_prepare.h"
#include "direct_prepare.h"

namespace {

using fav::ReferenceContext;

constexpr int kMaxIterations = 10'000'000;   // renderer.cpp:88-93
constexpr int kMaxThreads = 4096;
constexpr int kMaxPrecisionBits = 131'072;
constexpr int kMaxPixels = 100'000'000;
constexpr int kMaxPoints = 100000000;  // renderer.cpp:92
constexpr size_t kMaxTextLength = 50'000;

thread_local std::string last_error;  // renderer.cpp:4596-4599

void set_error(const std::string& message) { last_error = message; }

bool valid_text(const char* text) {  // renderer.cpp:207-213
    if (!text) return false;
    for (size_t n = 0; n <= kMaxTextLength; ++n)
        if (text[n] == '\0') return true;
    return false;
}

bool valid_dimensions(int w, int h) {  // renderer.cpp:170-175
    if (w <= 0 || h <= 0) return false;
    return static_cast<uint64_t>(w) <= static_cast<uint64_t>(kMaxPixels) / static_cast<uint64_t>(h);
}
bool valid_iterations(int n) { return n > 0 && n <= kMaxIterations; }
bool valid_threads(int n) { return n >= 0 && n <= kMaxThreads; }
bool valid_precision(int bits) { return bits >= 128 && bits <= kMaxPrecisionBits; }
bool valid_series(int order, int block) { return order >= 1 && order <= 32 && block >= 2 && block <= 4096; }

void parse_coordinate(const char* text, const char* label) {  // renderer.cpp:214-241
    if (!valid_text(text)) throw std::runtime_error("native coordinate text is too long or null");
    std::istringstream stream(text);
    stream.imbue(std::locale::classic());
    long double value = 0.0L;
    stream >> value;
    bool ok = !stream.fail();
    if (ok) {
        stream >> std::ws;
        ok = stream.eof();
    }
    if (!ok || !std::isfinite(value)) throw std::runtime_error(std::string("invalid ") + label + " coordinate");
}

FractalRenderOptions default_options() {  // renderer.cpp:4708-4727
    FractalRenderOptions o{};
    o.struct_size = sizeof(FractalRenderOptions);
    o.version = FRACTAL_RENDER_OPTIONS_VERSION;
    o.strict = 1;
    o.series_min_terms = 8;
    o.series_max_terms = 32;
    o.max_bla_length = 64;
    o.max_linear_bla_length = 4096;
    o.backend = 2;
    return o;
}

FractalRenderOptions checked_options(const FractalRenderOptions* supplied) {  // renderer.cpp:4729-4765
    FractalRenderOptions o = default_options();
    if (!supplied) return o;
    if (supplied->struct_size < sizeof(FractalRenderOptions) || supplied->version != FRACTAL_RENDER_OPTIONS_VERSION)
        throw std::runtime_error("unsupported FractalRenderOptions version or size");
    o = *supplied;
    auto flag = [](int32_t v, const char* label) {
        if (v != 0 && v != 1) throw std::runtime_error(std::string(label) + " must be 0 or 1");
    };
    flag(o.strict, "strict render flag");
    flag(o.allow_recovery, "recovery flag");
    flag(o.disable_bla, "disable-BLA flag");
    flag(o.disable_cycle, "disable-cycle flag");
    flag(o.strict_cycle, "strict-cycle flag");
    if (o.escape_radius_mode != 0 && o.escape_radius_mode != 1) throw std::runtime_error("unknown escape-radius mode");
    if (o.coordinate_mode != 0 && o.coordinate_mode != 1) throw std::runtime_error("unknown coordinate mode");
    o.series_min_terms = std::clamp(o.series_min_terms, 1, 32);
    o.series_max_terms = std::clamp(o.series_max_terms, o.series_min_terms, 32);
    o.max_bla_length = std::clamp(o.max_bla_length, 1, 64);
    o.max_linear_bla_length = std::clamp(o.max_linear_bla_length, 1, 4096);
    if (o.time_budget_ms < 0) throw std::runtime_error("render time budget cannot be negative");
    if (!std::isfinite(o.output_bias) || o.output_bias < 0.0)
        throw std::runtime_error("render output bias must be finite and non-negative");
    return o;
}

void require_hip_backend(const FractalRenderOptions& o) {  // renderer.cpp:19835-19837
    if (o.backend != 0 && o.backend != 1 && o.backend != 2) throw std::runtime_error("unknown native render backend");
    if (o.backend != 2)
        throw std::runtime_error(
            "this library is the HIP backend (backend=2); the scalar/AVX2 CPU backends are not part of it");
}

bool valid_planes(const FractalRenderPlanes* p, int w, int h) {  // renderer.cpp:7493-7507
    if (!p || !valid_dimensions(w, h)) return false;
    return p->struct_size >= sizeof(FractalRenderPlanes) && p->version == FRACTAL_RENDER_PLANES_VERSION
        && p->orbit_iteration && p->phase && p->de_x && p->de_y && p->test1 && p->test2;
}

// ---- opaque handle registry, renderer.cpp:9791-9836 -----------------------
std::mutex registry_mutex;
std::unordered_map<uintptr_t, std::shared_ptr<ReferenceContext>> registry;
uintptr_t next_handle = 0x1000U;

void* register_reference(std::unique_ptr<ReferenceContext> context) {
    auto shared = std::shared_ptr<ReferenceContext>(std::move(context));
    std::lock_guard<std::mutex> lock(registry_mutex);
    for (;;) {
        const uintptr_t token = next_handle++;
        if (token == 0 || registry.count(token)) continue;
        registry.emplace(token, std::move(shared));
        return reinterpret_cast<void*>(token);
    }
}

// batches share the token space (and counter) of reference handles
std::unordered_map<uintptr_t, std::shared_ptr<fav::PreparedBatch>> batches;

void* register_batch(std::shared_ptr<fav::PreparedBatch> batch) {
    std::lock_guard<std::mutex> lock(registry_mutex);
    for (;;) {
        const uintptr_t token = next_handle++;
        if (token == 0 || registry.count(token) || batches.count(token)) continue;
        batches.emplace(token, std::move(batch));
        return reinterpret_cast<void*>(token);
    }
}

std::shared_ptr<fav::PreparedBatch> acquire_batch(void* handle) {
    if (!handle) return {};
    std::lock_guard<std::mutex> lock(registry_mutex);
    auto it = batches.find(reinterpret_cast<uintptr_t>(handle));
    return it == batches.end() ? nullptr : it->second;
}

std::shared_ptr<ReferenceContext> acquire(void* handle) {
    if (!handle) return {};
    std::lock_guard<std::mutex> lock(registry_mutex);
    auto it = registry.find(reinterpret_cast<uintptr_t>(handle));
    return it == registry.end() ? nullptr : it->second;
}

std::shared_ptr<ReferenceContext> remove(void* handle) {
    if (!handle) return {};
    std::lock_guard<std::mutex> lock(registry_mutex);
    auto it = registry.find(reinterpret_cast<uintptr_t>(handle));
    if (it == registry.end()) return {};
    auto context = std::move(it->second);
    registry.erase(it);
    return context;
}

// ---- opt-in statistics, renderer.cpp:4645-4706 ----------------------------
std::atomic<bool> stats_enabled{false};
std::mutex stats_mutex;
uint64_t last_stats[fav::kStatsFields] = {};

void publish_stats(const uint64_t* values) {
    std::lock_guard<std::mutex> lock(stats_mutex);
    std::memcpy(last_stats, values, sizeof(last_stats));
}

int copy_stats(uint64_t* values, int capacity, int fields) {
    if (!values || capacity < fields) return -1;
    std::lock_guard<std::mutex> lock(stats_mutex);
    std::memcpy(values, last_stats, sizeof(uint64_t) * static_cast<size_t>(fields));
    return fields;
}

void* create(const char* x, const char* y, const char* zoom, int max_iter, int bits, int series_order,
             bool reusable, int formula, const char* julia_re, const char* julia_im,
             const FractalRenderOptions* supplied, bool use_options, const char* what,
             int bla_epsilon_exponent = -38) {
    try {
        if (!x || !y || !zoom || !julia_re || !julia_im || !valid_iterations(max_iter) || !valid_precision(bits)
            || !valid_series(series_order, 2) || formula < 0 || formula > 3)
            throw std::runtime_error(what);
        const FractalRenderOptions o = use_options ? checked_options(supplied) : default_options();
        if (!valid_text(x) || !valid_text(y) || !valid_text(zoom))
            throw std::runtime_error("native reference text is too long or null");
        parse_coordinate(x, "real");
        parse_coordinate(y, "imaginary");
        auto context = fav::build_reference(x, y, zoom, max_iter, bits, series_order, reusable, formula,
                                            use_options ? o.escape_radius_mode : 0,
                                            use_options ? o.coordinate_mode : 0, julia_re, julia_im,
                                            bla_epsilon_exponent);
        void* handle = register_reference(std::move(context));
        set_error("");
        return handle;
    } catch (const std::exception& e) {
        set_error(e.what());
        return nullptr;
    } catch (...) {
        set_error("native reference creation failed with an unknown exception");
        return nullptr;
    }
}

int render_impl(float* output, int width, int height, const char* zoom, void* handle, int max_iter,
                int threads, int series_order, int series_block, const FractalRenderOptions* supplied,
                FractalRenderPlanes* planes, bool device_pointers) {
    try {
        if (!output || !zoom || !handle || !valid_dimensions(width, height) || !valid_iterations(max_iter)
            || !valid_threads(threads) || !valid_series(series_order, series_block)
            || (planes != nullptr && !valid_planes(planes, width, height)))
            throw std::runtime_error("invalid native render dimensions or handle");
        const FractalRenderOptions o = checked_options(supplied);
        require_hip_backend(o);
        const auto context = acquire(handle);
        if (!context) throw std::runtime_error("invalid or already-destroyed reference handle");
        if (planes != nullptr && context->formula != 0)  // renderer.cpp:19842-19848
            throw std::runtime_error("render metadata requires the native scalar backend for this formula");
        if (context->escape_radius_mode != o.escape_radius_mode)
            throw std::runtime_error("reference escape-radius mode does not match render options");
        if (max_iter > context->requested_max_iter)
            throw std::runtime_error("render iteration count exceeds prepared reference");
        if (!valid_text(zoom)) throw std::runtime_error("native zoom text is too long or null");
        const fav::FE z = fav::parse_zoom(zoom, context->precision_bits);
        const long double log10_zoom =
            (std::log(static_cast<long double>(z.m)) + static_cast<long double>(z.e) * 0.693147180559945309417232121458176568L)
            / std::log(10.0L);
        if (log10_zoom < 6.0L) {
            // renderer.cpp:19920-19960: shallow views of a handle use the
            // direct double-precision renderer at the handle's centre
            fav::direct::DirectSetup d = fav::direct::make_setup_values(
                width, height, fav::direct::parse_zoom_long(zoom), context->x_center, context->y_center, max_iter,
                context->formula, context->julia_re, context->julia_im, o.output_bias, o.escape_radius_mode,
                o.coordinate_mode, static_cast<uint32_t>(o.reserved[0]));
            d.params.output = output;
            if (planes) {
                d.params.plane_iteration = planes->orbit_iteration;
                d.params.plane_phase = planes->phase;
                d.params.plane_de_x = planes->de_x;
                d.params.plane_de_y = planes->de_y;
                d.params.plane_test1 = planes->test1;
                d.params.plane_test2 = planes->test2;
            }
            fav::render_direct_on_device(d.params, d.x.data(), d.y.data(), planes != nullptr, device_pointers);
            set_error("");
            return 0;
        }
        uint64_t stats[fav::kStatsFields] = {};
        const bool collect = stats_enabled.load(std::memory_order_relaxed);
        fav::FieldRequest request{output, width, height, zoom, max_iter, series_order, series_block, o,
                                  planes, collect ? stats : nullptr, device_pointers};
        if (context->formula != 0)
            fav::render_alternate(*context, request);
        else
            fav::render_field(*context, request);
        if (collect) publish_stats(stats);
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    } catch (...) {
        set_error("native reference render failed with an unknown exception");
        return 1;
    }
}

// fractal_render_points_impl, renderer.cpp:20019-20186: the same checks and
// the deep field loop (Mandelbrot or alternate formula) over caller-supplied
// dc offsets instead of a pixel grid (project coordinates, no planes on this
// backend).
int render_points_impl(float* output, int point_count, const char* zoom, const double* real_mantissa,
                       const double* imag_mantissa, const int32_t* exponents, void* handle, int max_iter,
                       int threads, int series_order, int series_block, const FractalRenderOptions* supplied,
                       FractalRenderPlanes* planes) {
    try {
        if (!output || point_count <= 0 || point_count > kMaxPoints || !zoom || !real_mantissa || !imag_mantissa
            || !exponents || !handle || (planes != nullptr && !valid_planes(planes, point_count, 1)))
            throw std::runtime_error("invalid native point-render arguments");
        if (!valid_iterations(max_iter) || !valid_threads(threads) || !valid_series(series_order, series_block))
            throw std::runtime_error("invalid native point-render limits");
        const FractalRenderOptions o = checked_options(supplied);
        if (o.backend != 0 && o.backend != 1 && o.backend != 2) throw std::runtime_error("unknown native render backend");
        if (planes != nullptr && o.backend != 0)
            throw std::runtime_error("point metadata requires the native scalar backend");
        require_hip_backend(o);
        const auto context = acquire(handle);
        if (!context) throw std::runtime_error("invalid or already-destroyed reference handle");
        if (max_iter > context->requested_max_iter)
            throw std::runtime_error("point render iteration count exceeds prepared reference");
        if (!valid_text(zoom)) throw std::runtime_error("native zoom text is too long or null");
        (void)fav::parse_zoom(zoom, context->precision_bits);  // malformed zoom text fails before the copy
        std::vector<fav::SC> points(static_cast<size_t>(point_count));
        fav::SN largest{0.0, 0};
        for (int i = 0; i < point_count; ++i) {
            fav::SC point{real_mantissa[i], imag_mantissa[i], exponents[i]};
            if (!fav::sc_finite(point)) throw std::runtime_error("non-finite point offset");
            fav::sc_normalize(point);
            points[static_cast<size_t>(i)] = point;
            const fav::SN radius = fav::sc_norm_squared(point);
            if (fav::sc_compare_norm(radius, largest) > 0) largest = radius;
        }
        if (o.coordinate_mode != 0)  // as the reference's GPU point path (renderer.cpp:20093-20097)
            throw std::runtime_error("point rendering on the HIP backend requires project coordinates");
        // the points as rows of up to 256, so wavefronts are filled
        const int row = point_count < 256 ? point_count : 256;
        uint64_t stats[fav::kStatsFields] = {};
        const bool collect = stats_enabled.load(std::memory_order_relaxed);
        fav::FieldRequest request{output, row, (point_count + row - 1) / row, zoom, max_iter, series_order,
                                  series_block, o, nullptr, collect ? stats : nullptr, false};
        request.points = points.data();
        request.point_count = point_count;
        request.point_radius = fav::fe_sqrt(fav::FE{largest.m, largest.e});
        if (context->formula != 0)  // render_alternate_reference_impl, renderer.cpp:20105-20136
            fav::render_alternate(*context, request);
        else
            fav::render_field(*context, request);
        if (collect) publish_stats(stats);
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    } catch (...) {
        set_error("native point render failed with an unknown exception");
        return 1;
    }
}

// fractal_atlas_colourise_kfp, renderer.cpp:16750-17321: an output-sized
// parent and a child tile at an explicit rectangle, each coloured from its
// own field and transfer range, blended across the feather.
int atlas_kfp_rect(const float* parent, int parent_width, int parent_height, const float* child, int child_width,
                   int child_height, uint8_t* output, int output_width, int output_height, int max_iter,
                   int child_left, int child_top, int feather, double phase, double vocal, double instrumental,
                   double pitch, const FractalKfpOptions* options, const uint8_t* lut, int lut_size, int threads) {
    try {
        if (!parent || !output || !valid_dimensions(parent_width, parent_height)
            || !valid_dimensions(output_width, output_height) || parent_width != output_width
            || parent_height != output_height || !valid_iterations(max_iter) || !valid_threads(threads)
            || !std::isfinite(phase) || !std::isfinite(vocal) || !std::isfinite(instrumental) || !std::isfinite(pitch)
            || !lut || !fav::kfp::valid_options(options, lut_size))
            throw std::runtime_error("invalid native KFP atlas dimensions or controls");
        const bool use_child = child != nullptr;
        if (use_child) {
            if (!valid_dimensions(child_width, child_height) || child_left < 0 || child_top < 0
                || child_width > output_width - child_left || child_height > output_height - child_top || feather < 0
                || feather > std::min(child_width, child_height))
                throw std::runtime_error("invalid native KFP atlas child rectangle");
        } else if (child_width != 0 || child_height != 0 || child_left != 0 || child_top != 0 || feather != 0) {
            throw std::runtime_error("invalid native KFP atlas child absence");
        }
        if (options->field_bias > static_cast<double>(max_iter))
            throw std::runtime_error("native KFP field bias exceeds iteration cap");
        const fav::kfp::KfpParams p = fav::kfp::make_params(*options, parent, output, output_width, output_height,
                                                            max_iter, lut, lut_size);
        fav::kfp::KfpParams c{};
        if (use_child) c = fav::kfp::make_params(*options, child, output, child_width, child_height, max_iter, lut,
                                                 lut_size);
        fav::kfp::RectAtlas g{};
        g.use_child = use_child;
        g.child_left = child_left;
        g.child_top = child_top;
        g.feather = feather;
        std::vector<float> ex, ey;
        if (use_child && feather >= 2) {  // renderer.cpp:16887-16905
            auto fill = [feather](std::vector<float>& v, int n) {
                v.resize(static_cast<size_t>(n));
                for (int i = 0; i < n; ++i) {
                    const int edge = std::min(i, n - 1 - i);
                    const float linear = std::min(1.0F, static_cast<float>(edge) / static_cast<float>(feather));
                    v[static_cast<size_t>(i)] = linear * linear * (3.0F - 2.0F * linear);
                }
            };
            fill(ex, child_width);
            fill(ey, child_height);
        }
        fav::atlas_colourise_kfp_rect(p, c, g, ex, ey);
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    } catch (...) {
        set_error("native KFP atlas colourizer failed with an unknown exception");
        return 1;
    }
}

// fractal_crop_colourise_kfp, renderer.cpp:17803-17968: the centred crop of
// one scalar tile, interior-preserving with the KFP field bias as the tile's
// own, then the KFP pass -- the raw atlas pass without a child.
int crop_kfp(const float* source, int source_width, int source_height, uint8_t* output, int output_width,
             int output_height, double zoom_factor, int max_iter, double phase, double vocal, double instrumental,
             double pitch, const FractalKfpOptions* options, const uint8_t* lut, int lut_size, int threads) {
    try {
        if (!source || !output || !valid_dimensions(source_width, source_height)
            || !valid_dimensions(output_width, output_height) || !valid_iterations(max_iter) || !valid_threads(threads)
            || !std::isfinite(zoom_factor) || zoom_factor <= 0.0 || !std::isfinite(phase) || !std::isfinite(vocal)
            || !std::isfinite(instrumental) || !std::isfinite(pitch) || !lut || !fav::kfp::valid_options(options, lut_size))
            throw std::runtime_error("invalid native KFP crop dimensions or controls");
        zoom_factor = std::max(zoom_factor, 1.0);
        if (options->field_bias > static_cast<double>(max_iter))
            throw std::runtime_error("native KFP field bias exceeds iteration cap");
        fav::atlas::ComposeParams a{};
        a.parent = source;
        a.parent_bias = options->field_bias;
        a.field_bias = options->field_bias;
        a.parent_width = source_width;
        a.parent_max_iter = max_iter;
        a.width = output_width;
        a.height = output_height;
        a.max_iter = max_iter;
        fav::atlas::Axis axes[4];
        fav::atlas::fill_axis(axes[0], source_width, output_width, zoom_factor);
        fav::atlas::fill_axis(axes[1], source_height, output_height, zoom_factor);
        const fav::kfp::KfpParams k = fav::kfp::make_params(*options, nullptr, output, output_width, output_height,
                                                            max_iter, lut, lut_size);
        fav::atlas_colourise_kfp_raw(a, static_cast<size_t>(source_width) * static_cast<size_t>(source_height), 0,
                                     axes, k, output);
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    } catch (...) {
        set_error("native KFP crop colourizer failed with an unknown exception");
        return 1;
    }
}

// fractal_atlas_colourise_kfp_planes, renderer.cpp:17323-17801: the
// reference's checks; tile selection, geometry and maps on the host; the
// compose of the field and planes and the KFP pass on the device.
int atlas_kfp_planes(const float* parent, int parent_width, int parent_height, int parent_max_iter,
                     const FractalKfpPlanes* parent_planes, const float* child, int child_width, int child_height,
                     int child_max_iter, const FractalKfpPlanes* child_planes, uint8_t* output, int output_width,
                     int output_height, double parent_zoom, double child_fraction, double child_zoom, int max_iter,
                     int centered_input, double phase, double vocal, double instrumental, double pitch,
                     const FractalKfpOptions* options, const uint8_t* lut, int lut_size, int threads) {
    try {
        if (!parent || !output || !parent_planes || !valid_dimensions(parent_width, parent_height)
            || !valid_dimensions(output_width, output_height) || !valid_iterations(parent_max_iter)
            || !valid_iterations(max_iter) || !valid_threads(threads) || centered_input < 0 || centered_input > 1
            || !std::isfinite(parent_zoom) || parent_zoom <= 0.0 || !std::isfinite(child_fraction)
            || child_fraction < 0.0 || child_fraction > 1.0 || !std::isfinite(child_zoom) || child_zoom <= 0.0
            || !std::isfinite(phase) || !std::isfinite(vocal) || !std::isfinite(instrumental) || !std::isfinite(pitch)
            || !lut || !fav::kfp::valid_options(options, lut_size)
            || !fav::kfp::valid_planes(parent_planes, parent_width, parent_height))
            throw std::runtime_error("invalid native plane-aware KFP atlas dimensions or controls");
        const bool use_child = child != nullptr && child_fraction > 0.0;
        if (use_child) {
            if (!valid_dimensions(child_width, child_height) || !valid_iterations(child_max_iter) || !child_planes
                || !fav::kfp::valid_planes(child_planes, child_width, child_height))
                throw std::runtime_error("invalid native plane-aware KFP atlas child tile");
        } else if (child != nullptr || child_planes != nullptr || child_width != 0 || child_height != 0
                   || child_max_iter != 0) {
            throw std::runtime_error("invalid native plane-aware KFP atlas child absence");
        }
        if (options->field_bias > static_cast<double>(max_iter))
            throw std::runtime_error("native KFP field bias exceeds iteration cap");
        parent_zoom = std::max(parent_zoom, 1.0);
        child_zoom = std::max(child_zoom, 1.0);
        const bool full_child = use_child && child_fraction >= 0.999999;
        fav::PlaneTile tiles[2];
        tiles[0] = fav::PlaneTile{full_child ? child : parent, full_child ? child_width : parent_width,
                                  full_child ? child_height : parent_height, full_child ? child_planes : parent_planes};
        if (use_child && !full_child) tiles[1] = fav::PlaneTile{child, child_width, child_height, child_planes};
        const int primary_max_iter = full_child ? child_max_iter : parent_max_iter;
        const int secondary_max_iter = use_child && !full_child ? child_max_iter : 0;
        const int target_max_iter = full_child ? child_max_iter : max_iter;
        fav::atlas::PlanesCompose a{};
        a.primary.width = tiles[0].width;
        a.primary.max_iter = primary_max_iter;
        a.primary.bias = centered_input ? static_cast<double>(primary_max_iter) : 0.0;
        a.secondary.width = tiles[1].width;
        a.secondary.max_iter = secondary_max_iter;
        a.secondary.bias = centered_input ? static_cast<double>(secondary_max_iter) : 0.0;
        a.use_secondary = tiles[1].field != nullptr;
        a.visible_width = a.use_secondary
            ? std::max(1, static_cast<int>(std::lround(static_cast<double>(output_width) * child_fraction))) : 0;
        a.visible_height = a.use_secondary
            ? std::max(1, static_cast<int>(std::lround(static_cast<double>(output_height) * child_fraction))) : 0;
        a.visible_left = a.use_secondary ? (output_width - a.visible_width) / 2 : 0;
        a.visible_top = a.use_secondary ? (output_height - a.visible_height) / 2 : 0;
        a.feather = a.use_secondary ? fav::atlas::seam_feather(a.visible_width, a.visible_height) : 0;
        a.width = output_width;
        a.height = output_height;
        a.target_max_iter = target_max_iter;
        a.field_bias = options->field_bias;
        fav::atlas::Axis axes[4];
        const double primary_zoom = full_child ? child_zoom : parent_zoom;
        fav::atlas::fill_axis(axes[0], tiles[0].width, output_width, primary_zoom);
        fav::atlas::fill_axis(axes[1], tiles[0].height, output_height, primary_zoom);
        if (a.use_secondary) {
            fav::atlas::fill_axis(axes[2], child_width, a.visible_width, child_zoom);
            fav::atlas::fill_axis(axes[3], child_height, a.visible_height, child_zoom);
        }
        // the composed planes: present where either tile has them; the
        // texture of the primary tile, else the secondary's (17513-17522)
        const FractalKfpPlanes* primary_planes = tiles[0].planes;
        const FractalKfpPlanes* secondary_planes = tiles[1].planes;
        const FractalKfpPlanes* texture = primary_planes;
        if (!texture->texture_rgb && secondary_planes && secondary_planes->texture_rgb) texture = secondary_planes;
        FractalKfpPlanes shape{};
        shape.struct_size = sizeof(FractalKfpPlanes);
        shape.version = FRACTAL_KFP_PLANES_VERSION;
        static const double present = 0.0;
        const bool has_bailout = primary_planes->bailout || (secondary_planes && secondary_planes->bailout);
        shape.bailout = has_bailout ? &present : nullptr;
        shape.texture_rgb = texture->texture_rgb;
        shape.texture_width = texture->texture_width;
        shape.texture_height = texture->texture_height;
        shape.texture_stride = texture->texture_stride;
        fav::kfp::KfpParams k = fav::kfp::make_params(*options, nullptr, output, output_width, output_height,
                                                      target_max_iter, lut, lut_size);
        fav::kfp::set_plane_context(k, *options, &shape);
        fav::atlas_colourise_kfp_planes(a, tiles, axes, k, texture, output);
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    } catch (...) {
        set_error("native plane-aware KFP atlas colourizer failed with an unknown exception");
        return 1;
    }
}

int one_shot(float* output, int width, int height, const char* zoom, const char* x, const char* y, int max_iter,
             int bits, int use_perturbation, int threads, int formula, double julia_re, double julia_im,
             const FractalRenderOptions* supplied, FractalRenderPlanes* planes) {
    // render_fractal_ex[_planes], renderer.cpp:20279-20474
    try {
        if (!output || !zoom || !x || !y || !valid_dimensions(width, height)
            || (planes && !valid_planes(planes, width, height)) || !valid_iterations(max_iter)
            || !valid_precision(bits) || use_perturbation < 0 || use_perturbation > 1 || !valid_threads(threads)
            || formula < 0 || formula > 3 || !std::isfinite(julia_re) || !std::isfinite(julia_im))
            throw std::runtime_error("invalid native render dimensions, formula, or argument");
        const FractalRenderOptions o = checked_options(supplied);
        require_hip_backend(o);
        // renderer.cpp:20399-20405 allows planes off the scalar backend for
        // Mandelbrot only; the HIP direct kernel also carries Julia's
        // analytic derivative, so Julia planes are accepted when rendered
        // directly.  Burning ship / tricorn planes are rejected as there.
        if (planes && formula != 0 && !(formula == 1 && !use_perturbation))
            throw std::runtime_error("render metadata requires the native scalar backend for this formula");
        if (!use_perturbation) {  // render_direct[_with_planes], renderer.cpp:20311-20328, 20406-20425
            fav::direct::DirectSetup d = fav::direct::make_setup(
                width, height, zoom, x, y, max_iter, formula, julia_re, julia_im, o.output_bias,
                o.escape_radius_mode, o.coordinate_mode, static_cast<uint32_t>(o.reserved[0]));
            d.params.output = output;
            if (planes) {
                d.params.plane_iteration = planes->orbit_iteration;
                d.params.plane_phase = planes->phase;
                d.params.plane_de_x = planes->de_x;
                d.params.plane_de_y = planes->de_y;
                d.params.plane_test1 = planes->test1;
                d.params.plane_test2 = planes->test2;
            }
            fav::render_direct_on_device(d.params, d.x.data(), d.y.data(), planes != nullptr, false);
            set_error("");
            return 0;
        }
        char jr[64], ji[64];
        std::snprintf(jr, sizeof(jr), "%.17g", julia_re);
        std::snprintf(ji, sizeof(ji), "%.17g", julia_im);
        void* handle = fractal_create_reference_ex_options(x, y, zoom, max_iter, bits, 8, formula, jr, ji, &o);
        if (!handle) throw std::runtime_error(last_error);
        const int status = render_impl(output, width, height, zoom, handle, max_iter, threads, 8, 32, &o, planes, false);
        const std::string render_error = last_error;
        fractal_destroy_reference(handle);
        if (status != 0) throw std::runtime_error(render_error);
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    } catch (...) {
        set_error("native render failed with an unknown exception");
        return 1;
    }
}

// fractal_colourise_kfp_impl for a scalar field, renderer.cpp:16314-16545.
int colourise_kfp(const float* field, uint8_t* output, int width, int height, int max_iter, double phase, double vocal,
                  double instrumental, double pitch, const FractalKfpOptions* options, const FractalKfpPlanes* planes,
                  const uint8_t* lut, int lut_size, int threads, bool device_pointers) {
    try {
        if (!field || !output || !valid_dimensions(width, height) || !valid_iterations(max_iter)
            || !valid_threads(threads) || !std::isfinite(phase) || !std::isfinite(vocal)
            || !std::isfinite(instrumental) || !std::isfinite(pitch) || !lut || !fav::kfp::valid_options(options, lut_size)
            || !fav::kfp::valid_planes(planes, width, height))
            throw std::runtime_error("invalid native KFP colour dimensions or controls");
        if (planes != nullptr && device_pointers)
            throw std::runtime_error("Kalles planes are taken by the host-pointer KFP entries");
        fav::kfp::KfpParams k = fav::kfp::make_params(*options, field, output, width, height, max_iter, lut,
                                                      lut_size);
        fav::kfp::set_plane_context(k, *options, planes);
        fav::colourise_kfp_on_device(k, device_pointers, planes);
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    } catch (...) {
        set_error("native KFP colouriser failed with an unknown exception");
        return 1;
    }
}

// fractal_atlas_colourise_kfp_raw, renderer.cpp:17969-18265: the same
// validation and geometry, the compose step and the KFP pass on the device.
int atlas_kfp_raw(const float* parent, int parent_width, int parent_height, int parent_max_iter, const float* child,
                  int child_width, int child_height, int child_max_iter, uint8_t* output, int output_width,
                  int output_height, double parent_zoom, double child_fraction, double child_zoom, int max_iter,
                  double phase, double vocal, double instrumental, double pitch, const FractalKfpOptions* options,
                  const uint8_t* lut, int lut_size, int threads) {
    try {
        if (!parent || !output || !valid_dimensions(parent_width, parent_height)
            || !valid_dimensions(output_width, output_height) || !valid_iterations(parent_max_iter)
            || !valid_iterations(max_iter) || !valid_threads(threads) || !std::isfinite(parent_zoom)
            || parent_zoom <= 0.0 || !std::isfinite(child_fraction) || child_fraction < 0.0 || child_fraction > 1.0
            || !std::isfinite(child_zoom) || child_zoom <= 0.0 || !std::isfinite(phase) || !std::isfinite(vocal)
            || !std::isfinite(instrumental) || !std::isfinite(pitch) || !lut
            || !fav::kfp::valid_options(options, lut_size))
            throw std::runtime_error("invalid native raw KFP atlas dimensions or controls");
        parent_zoom = std::max(parent_zoom, 1.0);
        child_zoom = std::max(child_zoom, 1.0);
        const bool use_child = child != nullptr && child_fraction > 0.0;
        if (use_child && (!valid_dimensions(child_width, child_height) || !valid_iterations(child_max_iter)))
            throw std::runtime_error("invalid native raw KFP atlas child tile");
        if (!use_child && child != nullptr)
            throw std::runtime_error("raw KFP atlas child needs a positive visible fraction");
        if (!use_child && (child_width != 0 || child_height != 0 || child_max_iter != 0))
            throw std::runtime_error("invalid native raw KFP atlas child absence");
        if (options->field_bias > static_cast<double>(max_iter))
            throw std::runtime_error("native KFP field bias exceeds iteration cap");
        // centred tiles are stored relative to their own caps
        const bool centered = options->field_bias > 0.0;
        fav::atlas::ComposeParams a{};
        a.parent = parent;
        a.child = use_child ? child : nullptr;
        a.parent_bias = centered ? static_cast<double>(parent_max_iter) : 0.0;
        a.child_bias = centered ? static_cast<double>(child_max_iter) : 0.0;
        a.field_bias = options->field_bias;
        a.parent_width = parent_width;
        a.parent_max_iter = parent_max_iter;
        a.child_width = use_child ? child_width : 0;
        a.child_max_iter = use_child ? child_max_iter : 0;
        a.width = output_width;
        a.height = output_height;
        a.max_iter = max_iter;
        fav::atlas::Axis axes[4];
        fav::atlas::fill_axis(axes[0], parent_width, output_width, parent_zoom);
        fav::atlas::fill_axis(axes[1], parent_height, output_height, parent_zoom);
        if (use_child) {
            const int visible_w = std::max(1, static_cast<int>(std::lround(static_cast<double>(output_width) * child_fraction)));
            const int visible_h = std::max(1, static_cast<int>(std::lround(static_cast<double>(output_height) * child_fraction)));
            const int left = (output_width - visible_w) / 2, top = (output_height - visible_h) / 2;
            const int halo = std::min(2, std::min(std::min(left, output_width - left - visible_w),
                                                  std::min(top, output_height - top - visible_h)));
            a.use_child = 1;
            a.visible_left = left;
            a.visible_top = top;
            a.visible_width = visible_w;
            a.visible_height = visible_h;
            a.child_left = left - halo;
            a.child_top = top - halo;
            a.feather = fav::atlas::seam_feather(visible_w, visible_h);
            fav::atlas::fill_axis_window(axes[2], child_width, visible_w + 2 * halo, -static_cast<double>(halo),
                                         static_cast<double>(visible_w), child_zoom);
            fav::atlas::fill_axis_window(axes[3], child_height, visible_h + 2 * halo, -static_cast<double>(halo),
                                         static_cast<double>(visible_h), child_zoom);
        }
        const fav::kfp::KfpParams k = fav::kfp::make_params(*options, nullptr, output, output_width, output_height,
                                                            max_iter, lut, lut_size);
        fav::atlas_colourise_kfp_raw(
            a, static_cast<size_t>(parent_width) * static_cast<size_t>(parent_height),
            use_child ? static_cast<size_t>(child_width) * static_cast<size_t>(child_height) : 0, axes, k, output);
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    } catch (...) {
        set_error("native raw KFP atlas colourizer failed with an unknown exception");
        return 1;
    }
}

}  // namespace

namespace fav {
// The thread-local error text of the C ABI, for entry points defined in
// other files.
void set_last_error(const std::string& message) { set_error(message); }
}  // namespace fav

extern "C" {

int fractal_colourise_kfp(const float* field, uint8_t* output, int width, int height, int max_iter, double phase,
                          double vocal, double instrumental, double pitch, const FractalKfpOptions* options,
                          const uint8_t* lut, int lut_size, int threads) {
    return colourise_kfp(field, output, width, height, max_iter, phase, vocal, instrumental, pitch, options, nullptr,
                         lut, lut_size, threads, false);
}

int fractal_colourise_kfp_opencl(const float* field, uint8_t* output, int width, int height, int max_iter,
                                 double phase, double vocal, double instrumental, double pitch,
                                 const FractalKfpOptions* options, const uint8_t* lut, int lut_size, int threads) {
    return colourise_kfp(field, output, width, height, max_iter, phase, vocal, instrumental, pitch, options, nullptr,
                         lut, lut_size, threads, false);
}

int fractal_colourise_kfp_precise(const float* field, uint8_t* output, int width, int height, int max_iter,
                                  double phase, double vocal, double instrumental, double pitch,
                                  const FractalKfpOptions* options, const uint8_t* lut, int lut_size, int threads) {
    return colourise_kfp(field, output, width, height, max_iter, phase, vocal, instrumental, pitch, options, nullptr,
                         lut, lut_size, threads, false);
}

int fractal_colourise_kfp_planes(const float* field, uint8_t* output, int width, int height, int max_iter,
                                 double phase, double vocal, double instrumental, double pitch,
                                 const FractalKfpOptions* options, const FractalKfpPlanes* planes,
                                 const uint8_t* lut, int lut_size, int threads) {
    return colourise_kfp(field, output, width, height, max_iter, phase, vocal, instrumental, pitch, options, planes,
                         lut, lut_size, threads, false);
}

int fractal_atlas_colourise_kfp_raw(const float* parent, int parent_width, int parent_height, int parent_max_iter,
                                    const float* child, int child_width, int child_height, int child_max_iter,
                                    uint8_t* output, int output_width, int output_height, double parent_zoom,
                                    double child_fraction, double child_zoom, int max_iter, double phase,
                                    double vocal, double instrumental, double pitch,
                                    const FractalKfpOptions* options, const uint8_t* lut, int lut_size,
                                    int threads) {
    return atlas_kfp_raw(parent, parent_width, parent_height, parent_max_iter, child, child_width, child_height,
                         child_max_iter, output, output_width, output_height, parent_zoom, child_fraction,
                         child_zoom, max_iter, phase, vocal, instrumental, pitch, options, lut, lut_size, threads);
}

int fractal_atlas_colourise_kfp_raw_precise(const float* parent, int parent_width, int parent_height,
                                            int parent_max_iter, const float* child, int child_width,
                                            int child_height, int child_max_iter, uint8_t* output,
                                            int output_width, int output_height, double parent_zoom,
                                            double child_fraction, double child_zoom, int max_iter, double phase,
                                            double vocal, double instrumental, double pitch,
                                            const FractalKfpOptions* options, const uint8_t* lut, int lut_size,
                                            int threads) {
    return atlas_kfp_raw(parent, parent_width, parent_height, parent_max_iter, child, child_width, child_height,
                         child_max_iter, output, output_width, output_height, parent_zoom, child_fraction,
                         child_zoom, max_iter, phase, vocal, instrumental, pitch, options, lut, lut_size, threads);
}

int fractal_atlas_colourise_kfp(const float* parent, int parent_width, int parent_height, const float* child,
                                int child_width, int child_height, uint8_t* output, int output_width,
                                int output_height, int max_iter, int child_left, int child_top, int feather,
                                double phase, double vocal, double instrumental, double pitch,
                                const FractalKfpOptions* options, const uint8_t* lut, int lut_size, int threads) {
    return atlas_kfp_rect(parent, parent_width, parent_height, child, child_width, child_height, output, output_width,
                          output_height, max_iter, child_left, child_top, feather, phase, vocal, instrumental, pitch,
                          options, lut, lut_size, threads);
}

// renderer.cpp:18298-18327: the same pass.
int fractal_atlas_colourise_kfp_precise(const float* parent, int parent_width, int parent_height, const float* child,
                                        int child_width, int child_height, uint8_t* output, int output_width,
                                        int output_height, int max_iter, int child_left, int child_top, int feather,
                                        double phase, double vocal, double instrumental, double pitch,
                                        const FractalKfpOptions* options, const uint8_t* lut, int lut_size,
                                        int threads) {
    return atlas_kfp_rect(parent, parent_width, parent_height, child, child_width, child_height, output, output_width,
                          output_height, max_iter, child_left, child_top, feather, phase, vocal, instrumental, pitch,
                          options, lut, lut_size, threads);
}

int fractal_crop_colourise_kfp(const float* source, int source_width, int source_height, uint8_t* output,
                               int output_width, int output_height, double zoom_factor, int max_iter, double phase,
                               double vocal, double instrumental, double pitch, const FractalKfpOptions* options,
                               const uint8_t* lut, int lut_size, int threads) {
    return crop_kfp(source, source_width, source_height, output, output_width, output_height, zoom_factor, max_iter,
                    phase, vocal, instrumental, pitch, options, lut, lut_size, threads);
}

// renderer.cpp:18273-18296: the same pass.
int fractal_crop_colourise_kfp_precise(const float* source, int source_width, int source_height, uint8_t* output,
                                       int output_width, int output_height, double zoom_factor, int max_iter,
                                       double phase, double vocal, double instrumental, double pitch,
                                       const FractalKfpOptions* options, const uint8_t* lut, int lut_size,
                                       int threads) {
    return crop_kfp(source, source_width, source_height, output, output_width, output_height, zoom_factor, max_iter,
                    phase, vocal, instrumental, pitch, options, lut, lut_size, threads);
}

int fractal_atlas_colourise_kfp_planes(const float* parent, int parent_width, int parent_height, int parent_max_iter,
                                       const FractalKfpPlanes* parent_planes, const float* child, int child_width,
                                       int child_height, int child_max_iter, const FractalKfpPlanes* child_planes,
                                       uint8_t* output, int output_width, int output_height, double parent_zoom,
                                       double child_fraction, double child_zoom, int max_iter, int centered_input,
                                       double phase, double vocal, double instrumental, double pitch,
                                       const FractalKfpOptions* options, const uint8_t* lut, int lut_size,
                                       int threads) {
    return atlas_kfp_planes(parent, parent_width, parent_height, parent_max_iter, parent_planes, child, child_width,
                            child_height, child_max_iter, child_planes, output, output_width, output_height,
                            parent_zoom, child_fraction, child_zoom, max_iter, centered_input, phase, vocal,
                            instrumental, pitch, options, lut, lut_size, threads);
}

// renderer.cpp:18329-18364: the same pass (the KFP transfer here is always
// the precise one).
int fractal_atlas_colourise_kfp_planes_precise(
    const float* parent, int parent_width, int parent_height, int parent_max_iter,
    const FractalKfpPlanes* parent_planes, const float* child, int child_width, int child_height, int child_max_iter,
    const FractalKfpPlanes* child_planes, uint8_t* output, int output_width, int output_height, double parent_zoom,
    double child_fraction, double child_zoom, int max_iter, int centered_input, double phase, double vocal,
    double instrumental, double pitch, const FractalKfpOptions* options, const uint8_t* lut, int lut_size,
    int threads) {
    return atlas_kfp_planes(parent, parent_width, parent_height, parent_max_iter, parent_planes, child, child_width,
                            child_height, child_max_iter, child_planes, output, output_width, output_height,
                            parent_zoom, child_fraction, child_zoom, max_iter, centered_input, phase, vocal,
                            instrumental, pitch, options, lut, lut_size, threads);
}

int fractal_hip_colourise_kfp_device(const float* device_field, uint8_t* device_rgb, int width, int height,
                                     int max_iter, const FractalKfpOptions* options, const uint8_t* lut,
                                     int lut_size) {
    return colourise_kfp(device_field, device_rgb, width, height, max_iter, 0.0, 0.0, 0.0, 0.0, options, nullptr,
                         lut, lut_size, 0, true);
}

int fractal_hip_colourise_kfp_device_noalloc(const float* device_field, uint8_t* device_rgb, int width, int height,
                                             int max_iter, const FractalKfpOptions* options, const uint8_t* lut,
                                             int lut_size) {
    try {
        if (!device_field || !device_rgb || !valid_dimensions(width, height) || !valid_iterations(max_iter) || !lut
            || !fav::kfp::valid_options(options, lut_size))
            throw std::runtime_error("invalid native KFP colour dimensions or controls");
        fav::kfp::KfpParams k = fav::kfp::make_params(*options, device_field, device_rgb, width, height, max_iter, lut,
                                                      lut_size);
        fav::kfp::set_plane_context(k, *options, nullptr);
        fav::colourise_kfp_device_noalloc(k);
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    } catch (...) {
        set_error("native allocation-free KFP colouriser failed with an unknown exception");
        return 1;
    }
}


int fractal_abi_version(void) { return FRACTAL_ABI_VERSION; }

int fractal_render_options_version(void) { return FRACTAL_RENDER_OPTIONS_VERSION; }

void fractal_render_options_default(FractalRenderOptions* options) {
    if (options) *options = default_options();
}

int fractal_backend_capabilities(void) {
    try {
        const int caps = fav::device_available() ? (4 | 8) : 0;
        set_error(caps ? "" : "no gfx950 HIP device is visible");
        return caps;
    } catch (...) {
        set_error("HIP capability query failed");
        return 0;
    }
}

const char* fractal_last_error(void) { return last_error.c_str(); }

void fractal_set_stats_enabled(int enabled) { stats_enabled.store(enabled != 0, std::memory_order_relaxed); }
int fractal_get_last_stats(uint64_t* values, int capacity) { return copy_stats(values, capacity, 16); }
int fractal_get_last_stats_ex(uint64_t* values, int capacity) { return copy_stats(values, capacity, fav::kStatsFields); }

void* fractal_create_reference(const char* x, const char* y, const char* zoom, int max_iter, int bits,
                               int series_order) {
    return create(x, y, zoom, max_iter, bits, series_order, false, 0, "0", "0", nullptr, false,
                  "invalid reference configuration");
}

void* fractal_create_reference_reusable(const char* x, const char* y, const char* zoom, int max_iter, int bits,
                                        int series_order) {
    return create(x, y, zoom, max_iter, bits, series_order, true, 0, "0", "0", nullptr, false,
                  "invalid reusable reference configuration");
}

void* fractal_create_reference_reusable_options(const char* x, const char* y, const char* zoom, int max_iter,
                                                int bits, int series_order, const FractalRenderOptions* options) {
    return create(x, y, zoom, max_iter, bits, series_order, true, 0, "0", "0", options, true,
                  "invalid reusable reference configuration");
}

void* fractal_create_reference_ex(const char* x, const char* y, const char* zoom, int max_iter, int bits,
                                  int series_order, int formula, const char* julia_re, const char* julia_im) {
    return create(x, y, zoom, max_iter, bits, series_order, formula != FRACTAL_FORMULA_MANDELBROT, formula,
                  julia_re, julia_im, nullptr, false, "invalid formula-aware reference configuration");
}

void* fractal_create_reference_ex_options(const char* x, const char* y, const char* zoom, int max_iter, int bits,
                                          int series_order, int formula, const char* julia_re,
                                          const char* julia_im, const FractalRenderOptions* options) {
    return create(x, y, zoom, max_iter, bits, series_order, formula != FRACTAL_FORMULA_MANDELBROT, formula,
                  julia_re, julia_im, options, true, "invalid formula-aware reference configuration");
}

void* fractal_hip_create_reference_tuned(const char* x, const char* y, const char* zoom, int max_iter, int bits,
                                         int series_order, int formula, const char* julia_re, const char* julia_im,
                                         const FractalRenderOptions* options, int bla_epsilon_exponent) {
    return create(x, y, zoom, max_iter, bits, series_order, formula != FRACTAL_FORMULA_MANDELBROT, formula,
                  julia_re, julia_im, options, true, "invalid formula-aware reference configuration",
                  bla_epsilon_exponent);
}

void* fractal_clone_reference(void* source_handle, const char* zoom) {
    try {
        if (!source_handle || !zoom) throw std::runtime_error("invalid reference tier clone configuration");
        const auto source = acquire(source_handle);
        if (!source) throw std::runtime_error("invalid or already-destroyed reference handle");
        auto context = fav::clone_reference(*source, zoom);
        void* handle = register_reference(std::move(context));
        set_error("");
        return handle;
    } catch (const std::exception& e) {
        set_error(e.what());
        return nullptr;
    } catch (...) {
        set_error("native reference clone failed with an unknown exception");
        return nullptr;
    }
}

void fractal_destroy_reference(void* handle) {
    try {
        if (!handle) {
            set_error("");
            return;
        }
        if (!remove(handle)) {
            set_error("invalid or already-destroyed reference handle");
            return;
        }
        set_error("");
    } catch (...) {
        set_error("native reference destruction failed with an unknown exception");
    }
}

int fractal_get_reference_stats(void* handle, uint64_t* values, int capacity) {
    try {
        if (!values || capacity < 5) {
            set_error("reference statistics buffer is too small or null");
            return -1;
        }
        const auto c = acquire(handle);
        if (!c) {
            set_error("invalid or already-destroyed reference handle");
            return -1;
        }
        values[0] = c->reference_build_ns;
        values[1] = c->series_build_ns;
        values[2] = c->bla_build_ns;
        values[3] = c->series_enabled ? static_cast<uint64_t>(c->series_iteration) : 0;
        values[4] = c->series_enabled ? static_cast<uint64_t>(c->series_order) : 0;
        set_error("");
        return 5;
    } catch (...) {
        set_error("reference statistics lookup failed with an unknown exception");
        return -1;
    }
}

int fractal_render_reference_ex_planes(float* output, int width, int height, const char* zoom, void* handle,
                                       int max_iter, int threads, int series_order, int series_block,
                                       const FractalRenderOptions* options, FractalRenderPlanes* planes) {
    return render_impl(output, width, height, zoom, handle, max_iter, threads, series_order, series_block, options,
                       planes, false);
}

int fractal_render_reference_ex(float* output, int width, int height, const char* zoom, void* handle, int max_iter,
                                int threads, int series_order, int series_block, const FractalRenderOptions* options) {
    return render_impl(output, width, height, zoom, handle, max_iter, threads, series_order, series_block, options,
                       nullptr, false);
}

int fractal_render_mandelbrot_reference_ex(float* output, int width, int height, const char* zoom, void* handle,
                                           int max_iter, int threads, int series_order, int series_block,
                                           const FractalRenderOptions* options) {
    return render_impl(output, width, height, zoom, handle, max_iter, threads, series_order, series_block, options,
                       nullptr, false);
}

int render_mandelbrot_reference(float* output, int width, int height, const char* zoom, void* handle, int max_iter,
                                int threads, int series_order, int series_block) {
    return render_impl(output, width, height, zoom, handle, max_iter, threads, series_order, series_block, nullptr,
                       nullptr, false);
}

int render_fractal_ex(float* output, int width, int height, const char* zoom, const char* x, const char* y,
                      int max_iter, int bits, int use_perturbation, int threads, int formula, double julia_re,
                      double julia_im, const FractalRenderOptions* options) {
    return one_shot(output, width, height, zoom, x, y, max_iter, bits, use_perturbation, threads, formula, julia_re,
                    julia_im, options, nullptr);
}

int render_mandelbrot_ex(float* output, int width, int height, const char* zoom, const char* x, const char* y,
                         int max_iter, int bits, int use_perturbation, int threads,
                         const FractalRenderOptions* options) {  // renderer.cpp:20478-20506
    return one_shot(output, width, height, zoom, x, y, max_iter, bits, use_perturbation, threads, 0, 0.0, 0.0,
                    options, nullptr);
}

int render_mandelbrot(float* output, int width, int height, const char* zoom, const char* x, const char* y,
                      int max_iter, int bits, int use_perturbation, int threads) {  // renderer.cpp:20508-20533
    return one_shot(output, width, height, zoom, x, y, max_iter, bits, use_perturbation, threads, 0, 0.0, 0.0,
                    nullptr, nullptr);
}

int render_fractal_ex_planes(float* output, int width, int height, const char* zoom, const char* x, const char* y,
                             int max_iter, int bits, int use_perturbation, int threads, int formula,
                             double julia_re, double julia_im, const FractalRenderOptions* options,
                             FractalRenderPlanes* planes) {
    if (!planes) {
        set_error("render metadata requires valid native render arguments");
        return 1;
    }
    return one_shot(output, width, height, zoom, x, y, max_iter, bits, use_perturbation, threads, formula, julia_re,
                    julia_im, options, planes);
}

// ---- extensions (fav_extensions.h) ----------------------------------------

}  // extern "C"

namespace {

constexpr int kMaxPreviewWork = 1 << 24;

// Checks shared by the scheduled, preview and learned entries: a handle of
// any formula, zoom >= 1e6 (perturbation fields), a prepared cap and
// matching escape mode.
std::shared_ptr<ReferenceContext> scheduled_context(void* handle, const char* zoom, int width, int height,
                                                    int max_iter, int series_order, int series_block,
                                                    const FractalRenderOptions& o) {
    if (!zoom || !handle || !valid_dimensions(width, height) || !valid_iterations(max_iter)
        || !valid_series(series_order, series_block))
        throw std::runtime_error("invalid native render dimensions or handle");
    require_hip_backend(o);
    auto context = acquire(handle);
    if (!context) throw std::runtime_error("invalid or already-destroyed reference handle");
    if (context->escape_radius_mode != o.escape_radius_mode)
        throw std::runtime_error("reference escape-radius mode does not match render options");
    if (max_iter > context->requested_max_iter)
        throw std::runtime_error("render iteration count exceeds prepared reference");
    if (!valid_text(zoom)) throw std::runtime_error("native zoom text is too long or null");
    const fav::FE z = fav::parse_zoom(zoom, context->precision_bits);
    const long double log10_zoom =
        (std::log(static_cast<long double>(z.m)) + static_cast<long double>(z.e) * 0.693147180559945309417232121458176568L)
        / std::log(10.0L);
    if (log10_zoom < 6.0L) throw std::runtime_error("scheduled renders are perturbation fields (zoom >= 1e6)");
    return context;
}

// Alternate-formula perturbation renders need zoom >= 1e6 (shallower views
// of a handle render through the direct renderer).
void require_perturbation_zoom(const ReferenceContext& context, const char* zoom) {
    const fav::FE z = fav::parse_zoom(zoom, context.precision_bits);
    const long double log10_zoom =
        (std::log(static_cast<long double>(z.m)) + static_cast<long double>(z.e) * 0.693147180559945309417232121458176568L)
        / std::log(10.0L);
    if (log10_zoom < 6.0L) throw std::runtime_error("this render needs a perturbation field (zoom >= 1e6)");
}

int guarded(const char* what, const std::function<void()>& body) {
    try {
        body();
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    } catch (...) {
        set_error(std::string(what) + " failed with an unknown exception");
        return 1;
    }
}

}  // namespace

extern "C" {

int fractal_hip_render_reference_scheduled(float* output, int width, int height, const char* zoom, void* handle,
                                           int max_iter, int series_order, int series_block,
                                           const FractalRenderOptions* options, const uint32_t* tile_order,
                                           int tile_count, int32_t* work) {
    return guarded("native scheduled render", [&] {
        if (!output || tile_count < 0 || (tile_order && tile_count == 0))
            throw std::runtime_error("invalid native render dimensions or handle");
        const FractalRenderOptions o = checked_options(options);
        const auto context = scheduled_context(handle, zoom, width, height, max_iter, series_order, series_block, o);
        if (tile_order) fav::check_tile_order(tile_order, static_cast<size_t>(tile_count), width, height);
        uint64_t stats[fav::kStatsFields] = {};
        const bool collect = stats_enabled.load(std::memory_order_relaxed);
        fav::FieldRequest request{output, width, height, zoom, max_iter, series_order, series_block, o,
                                  nullptr, collect ? stats : nullptr, false};
        if (context->formula != 0) {
            fav::render_alternate(*context, request, tile_order, static_cast<size_t>(tile_count), work);
        } else {
            request.tile_order = tile_order;
            request.tile_count = static_cast<size_t>(tile_count);
            request.work = work;
            fav::render_field(*context, request);
        }
        if (collect) publish_stats(stats);
    });
}

int fractal_hip_render_reference_counts(float* output, int32_t* counts, int width, int height, const char* zoom,
                                        void* handle, int max_iter, int series_order, int series_block,
                                        const FractalRenderOptions* options) {
    return guarded("native counts render", [&] {
        if (!output || !counts) throw std::runtime_error("output and counts must not be null");
        const FractalRenderOptions o = checked_options(options);
        const auto context = scheduled_context(handle, zoom, width, height, max_iter, series_order, series_block, o);
        fav::FieldRequest request{output, width, height, zoom, max_iter, series_order, series_block, o,
                                  nullptr, nullptr, false};
        if (context->formula != 0) {
            fav::render_alternate(*context, request, nullptr, 0, nullptr, counts);
            return;
        }
        // the Mandelbrot map: from the iteration plane (the escape iteration is
        // orbit_iteration + 1, the cap inside), -1 where the field is NaN
        const size_t n = static_cast<size_t>(width) * static_cast<size_t>(height);
        std::vector<int64_t> iteration(n);
        std::vector<double> scratch(5 * n);
        FractalRenderPlanes planes{};
        planes.struct_size = sizeof(FractalRenderPlanes);
        planes.version = FRACTAL_RENDER_PLANES_VERSION;
        planes.orbit_iteration = iteration.data();
        planes.phase = scratch.data();
        planes.de_x = scratch.data() + n;
        planes.de_y = scratch.data() + 2 * n;
        planes.test1 = scratch.data() + 3 * n;
        planes.test2 = scratch.data() + 4 * n;
        request.planes = &planes;
        fav::render_field(*context, request);
        for (size_t i = 0; i < n; ++i)
            counts[i] = output[i] != output[i] ? -1
                : static_cast<int32_t>(iteration[i] < max_iter ? iteration[i] + 1 : max_iter);
    });
}

int fractal_hip_preview_features(float* features, int width, int height, const char* zoom, void* handle,
                                 int max_iter, int series_order, int series_block,
                                 const FractalRenderOptions* options, int work_limit) {
    return guarded("native preview", [&] {
        if (!features) throw std::runtime_error("feature buffer must not be null");
        if (work_limit <= 0 || work_limit > kMaxPreviewWork)
            throw std::runtime_error("preview work limit must be in [1, 16777216]");
        const FractalRenderOptions o = checked_options(options);
        const auto context = scheduled_context(handle, zoom, width, height, max_iter, series_order, series_block, o);
        fav::FieldRequest request{nullptr, width, height, zoom, max_iter, series_order, series_block, o,
                                  nullptr, nullptr, false};
        if (context->formula != 0) fav::alternate_preview_features(*context, request, work_limit, features);
        else fav::field_preview_features(*context, request, work_limit, features);
    });
}

int fractal_hip_render_reference_learned(float* output, int width, int height, const char* zoom, void* handle,
                                         int max_iter, int series_order, int series_block,
                                         const FractalRenderOptions* options, const float* weights, int work_limit,
                                         uint32_t* tile_order, double* timings) {
    return guarded("native learned render", [&] {
        if (!output || !weights) throw std::runtime_error("output and weights must not be null");
        if (work_limit <= 0 || work_limit > kMaxPreviewWork)
            throw std::runtime_error("preview work limit must be in [1, 16777216]");
        const FractalRenderOptions o = checked_options(options);
        const auto context = scheduled_context(handle, zoom, width, height, max_iter, series_order, series_block, o);
        fav::FieldRequest request{output, width, height, zoom, max_iter, series_order, series_block, o,
                                  nullptr, nullptr, false};
        if (context->formula != 0)
            fav::render_alternate_learned(*context, request, weights, work_limit, tile_order, timings);
        else
            fav::render_field_learned(*context, request, weights, work_limit, tile_order, timings);
    });
}

void* fractal_hip_batch_create(const FractalHipBatchField* fields, int count, const FractalRenderOptions* options) {
    try {
        if (!fields || count <= 0 || count > 4095) throw std::runtime_error("invalid batch field list");
        const FractalRenderOptions o = checked_options(options);
        require_hip_backend(o);
        std::vector<fav::BatchFieldRequest> list;
        for (int i = 0; i < count; ++i) {
            const FractalHipBatchField& f = fields[i];
            if (!f.device_output || !f.zoom_text || !valid_dimensions(f.width, f.height)
                || !valid_iterations(f.max_iter) || !valid_series(f.series_order, f.series_block)
                || !valid_text(f.zoom_text))
                throw std::runtime_error("invalid native render dimensions or handle");
            auto context = acquire(f.handle);
            if (!context) throw std::runtime_error("invalid or already-destroyed reference handle");
            if (context->escape_radius_mode != o.escape_radius_mode)
                throw std::runtime_error("reference escape-radius mode does not match render options");
            if (f.max_iter > context->requested_max_iter)
                throw std::runtime_error("render iteration count exceeds prepared reference");
            const fav::FE z = fav::parse_zoom(f.zoom_text, context->precision_bits);
            const long double log10_zoom =
                (std::log(static_cast<long double>(z.m)) + static_cast<long double>(z.e) * 0.693147180559945309417232121458176568L)
                / std::log(10.0L);
            fav::FieldRequest request{f.device_output, f.width, f.height, f.zoom_text, f.max_iter, f.series_order,
                                      f.series_block, o, nullptr, nullptr, true};
            fav::BatchFieldRequest entry{std::move(context), request};
            if (log10_zoom < 6.0L) {  // the handle render's shallow branch (fractal_render_reference_ex above)
                entry.direct = true;
                entry.direct_zoom = fav::direct::parse_zoom_long(f.zoom_text);
            }
            list.push_back(std::move(entry));
        }
        void* handle = register_batch(fav::prepare_batch(std::move(list)));
        set_error("");
        return handle;
    } catch (const std::exception& e) {
        set_error(e.what());
        return nullptr;
    } catch (...) {
        set_error("native batch creation failed with an unknown exception");
        return nullptr;
    }
}

int fractal_hip_render_segment(const FractalHipSegmentField* fields, int count, const FractalRenderOptions* options,
                               int build_threads, int pipelined, double* timings) {
    try {
        if (!fields || count <= 0 || count > 4095 || build_threads < 1 || build_threads > 256)
            throw std::runtime_error("invalid segment field list");
        const FractalRenderOptions o = checked_options(options);
        require_hip_backend(o);
        std::vector<fav::SegmentField> list;
        list.reserve(static_cast<size_t>(count));
        for (int i = 0; i < count; ++i) {
            const FractalHipSegmentField& f = fields[i];
            if (!f.x_center || !f.y_center || !f.zoom_text || !f.device_output || !valid_dimensions(f.width, f.height)
                || !valid_iterations(f.max_iter) || !valid_precision(f.precision_bits)
                || !valid_series(f.series_order, f.series_block) || !valid_text(f.x_center) || !valid_text(f.y_center)
                || !valid_text(f.zoom_text))
                throw std::runtime_error("invalid native render dimensions or reference configuration");
            parse_coordinate(f.x_center, "real");
            parse_coordinate(f.y_center, "imaginary");
            list.push_back(fav::SegmentField{f.x_center, f.y_center, f.zoom_text, f.device_output, f.width, f.height,
                                             f.max_iter, f.precision_bits, f.series_order, f.series_block});
        }
        fav::SegmentTiming t;
        fav::render_segment(list, o, build_threads, pipelined != 0, t);
        if (timings) {
            timings[0] = t.wall_ms;
            timings[1] = t.build_ms;
            timings[2] = t.render_ms;
            timings[3] = t.built_ms;
        }
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    } catch (...) {
        set_error("native segment render failed with an unknown exception");
        return 1;
    }
}

int fractal_hip_batch_tile_count(void* batch) {
    const auto b = acquire_batch(batch);
    if (!b) {
        set_error("invalid or already-destroyed batch handle");
        return -1;
    }
    set_error("");
    return static_cast<int>(fav::batch_tile_count(*b));
}

int fractal_hip_batch_set_order(void* batch, int order, const float* tile_costs, int cost_count) {
    try {
        const auto b = acquire_batch(batch);
        if (!b) throw std::runtime_error("invalid or already-destroyed batch handle");
        fav::batch_set_order(*b, order, tile_costs, cost_count > 0 ? static_cast<size_t>(cost_count) : 0);
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    }
}

int fractal_hip_batch_set_precision(void* batch, int floatexp) {
    try {
        const auto b = acquire_batch(batch);
        if (!b) throw std::runtime_error("invalid or already-destroyed batch handle");
        if (floatexp != 0 && floatexp != 1) throw std::runtime_error("batch precision must be 0 (FP64) or 1 (floatexp)");
        fav::batch_set_floatexp(*b, floatexp != 0);
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    }
}

int fractal_hip_batch_render(void* batch, int fast_blocks_per_cu, int use_graph) {
    try {
        const auto b = acquire_batch(batch);
        if (!b) throw std::runtime_error("invalid or already-destroyed batch handle");
        if (fast_blocks_per_cu < 0 || fast_blocks_per_cu > 64 || (use_graph != 0 && use_graph != 1))
            throw std::runtime_error("invalid batch launch settings");
        fav::batch_render(*b, fast_blocks_per_cu, use_graph != 0);
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    }
}

void fractal_hip_batch_destroy(void* batch) {
    if (!batch) return;
    std::shared_ptr<fav::PreparedBatch> doomed;
    {
        std::lock_guard<std::mutex> lock(registry_mutex);
        auto it = batches.find(reinterpret_cast<uintptr_t>(batch));
        if (it == batches.end()) {
            set_error("invalid or already-destroyed batch handle");
            return;
        }
        doomed = std::move(it->second);
        batches.erase(it);
    }
    set_error("");
}

int fractal_hip_render_reference_floatexp(float* output, int32_t* counts, int width, int height, const char* zoom,
                                          void* handle, int max_iter, int series_order, int series_block,
                                          const FractalRenderOptions* options, int device_pointers) {
    try {
        if (!output || !counts || !zoom || !handle || !valid_dimensions(width, height) || !valid_iterations(max_iter)
            || !valid_series(series_order, series_block) || (device_pointers != 0 && device_pointers != 1))
            throw std::runtime_error("invalid native render dimensions or handle");
        const FractalRenderOptions o = checked_options(options);
        require_hip_backend(o);
        const auto context = acquire(handle);
        if (!context) throw std::runtime_error("invalid or already-destroyed reference handle");
        if (context->escape_radius_mode != o.escape_radius_mode)
            throw std::runtime_error("reference escape-radius mode does not match render options");
        if (max_iter > context->requested_max_iter)
            throw std::runtime_error("render iteration count exceeds prepared reference");
        if (!valid_text(zoom)) throw std::runtime_error("native zoom text is too long or null");
        if (context->formula != 0) require_perturbation_zoom(*context, zoom);
        fav::FieldRequest request{output, width, height, zoom, max_iter, series_order, series_block, o,
                                  nullptr, nullptr, device_pointers != 0};
        fav::render_field_floatexp(*context, request, counts);
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    } catch (...) {
        set_error("native floatexp render failed with an unknown exception");
        return 1;
    }
}

// ---- learned tile-cost model (no counterpart in the reference) ----

// static: inside the extern "C" block an anonymous namespace would still
// give the helper C linkage and export it.
static constexpr int kMaxMlpSamples = 1 << 22;

static void check_samples(int count) {
    if (count <= 0 || count > kMaxMlpSamples)
        throw std::runtime_error("cost model sample count must be in [1, 4194304]");
}

int fractal_hip_cost_model_parameter_count(void) { return fav::mlp::kParams; }

int fractal_hip_cost_model_forward(const float* weights, const float* features, int count, float* costs) {
    return guarded("cost model forward", [&] {
        if (!weights || !features || !costs) throw std::runtime_error("cost model buffers must not be null");
        check_samples(count);
        fav::mlp_forward(weights, features, count, costs);
    });
}

int fractal_hip_cost_model_gradient(const float* weights, const float* features, const float* targets, int count,
                                    float* gradient, float* loss) {
    return guarded("cost model gradient", [&] {
        if (!weights || !features || !targets || !gradient || !loss)
            throw std::runtime_error("cost model buffers must not be null");
        check_samples(count);
        fav::mlp_gradient(weights, features, targets, count, gradient, loss);
    });
}

int fractal_hip_cost_model_train(float* weights, float* adam_m, float* adam_v, const float* features,
                                 const float* targets, int count, int steps, float learning_rate, int first_step,
                                 float* loss_history) {
    return guarded("cost model training", [&] {
        if (!weights || !adam_m || !adam_v || !features || !targets || !loss_history)
            throw std::runtime_error("cost model buffers must not be null");
        check_samples(count);
        if (steps <= 0 || steps > 100000) throw std::runtime_error("training steps must be in [1, 100000]");
        if (first_step < 1 || first_step > 100000000) throw std::runtime_error("first Adam step must be >= 1");
        if (!(learning_rate > 0.0F) || !(learning_rate < 1.0F))
            throw std::runtime_error("learning rate must be in (0, 1)");
        fav::mlp_train(weights, adam_m, adam_v, features, targets, count, steps, learning_rate, first_step,
                       loss_history);
    });
}

int fractal_render_points(float* output, int point_count, const char* zoom, const double* real_mantissa,
                          const double* imag_mantissa, const int32_t* exponents, void* handle, int max_iter,
                          int threads, int series_order, int series_block, const FractalRenderOptions* options) {
    return render_points_impl(output, point_count, zoom, real_mantissa, imag_mantissa, exponents, handle, max_iter,
                              threads, series_order, series_block, options, nullptr);
}

int fractal_render_points_ex_planes(float* output, int point_count, const char* zoom, const double* real_mantissa,
                                    const double* imag_mantissa, const int32_t* exponents, void* handle,
                                    int max_iter, int threads, int series_order, int series_block,
                                    const FractalRenderOptions* options, FractalRenderPlanes* planes) {
    return render_points_impl(output, point_count, zoom, real_mantissa, imag_mantissa, exponents, handle, max_iter,
                              threads, series_order, series_block, options, planes);
}

int fractal_hip_render_reference_device(float* device_output, int width, int height, const char* zoom, void* handle,
                                        int max_iter, int series_order, int series_block,
                                        const FractalRenderOptions* options, FractalRenderPlanes* device_planes) {
    return render_impl(device_output, width, height, zoom, handle, max_iter, 0, series_order, series_block, options,
                       device_planes, true);
}

int fractal_hip_last_timing(double* values, int capacity) {
    if (!values || capacity < 3) {
        set_error("timing buffer is too small or null");
        return -1;
    }
    const fav::LastTiming t = fav::last_timing();
    values[0] = t.kernel_ms;
    values[1] = t.total_ms;
    values[2] = t.pixels;
    if (capacity >= 4) values[3] = t.plain_ms;
    set_error("");
    return capacity >= 4 ? 4 : 3;
}

int fractal_hip_reference_upload(void* handle) {
    try {
        const auto c = acquire(handle);
        if (!c) throw std::runtime_error("invalid or already-destroyed reference handle");
        fav::ensure_device_tables(*c);
        set_error("");
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    }
}

}  // extern "C"
