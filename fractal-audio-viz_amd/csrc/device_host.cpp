// device_host.cpp -- HBM residency of reference tables and the host side of
// one field launch: viewport geometry, per-row / per-column offsets,
// scratch buffers, the kernel launch on the library stream, HIP-event
// timing and the copy of results back to the caller.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include <hip/hip_runtime_api.h>

#include "fav_host.h"
#include "fav_render.h"

namespace fav {

hipError_t launch_field_kernel(const RenderParams& p, bool planes, bool stats, hipStream_t stream);

namespace {

void check(hipError_t status, const char* what) {
    if (status != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(status));
}

uint64_t now_ns() {
    return static_cast<uint64_t>(std::chrono::duration_cast<std::chrono::nanoseconds>(
        std::chrono::steady_clock::now().time_since_epoch()).count());
}

// Grow-only device buffer.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    void* get(size_t need) {
        if (need <= bytes && ptr) return ptr;
        if (ptr) check(hipFree(ptr), "hipFree");
        ptr = nullptr;
        bytes = 0;
        const size_t rounded = std::max<size_t>(need, 256);
        check(hipMalloc(&ptr, rounded), "hipMalloc scratch");
        bytes = rounded;
        return ptr;
    }
};

// Per-process launch state.  Renders are serialised by `mutex` (the
// reference's OpenCL backend serialises the same way, renderer.cpp:9064);
// the device does the parallel work.
struct LaunchState {
    std::mutex mutex;
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
    DeviceBuffer offsets, output, planes, stats, budget;
    double wall_clock_khz = 0.0;
    LastTiming timing{};
};

LaunchState& state() {
    static LaunchState s;
    return s;
}

void init_locked(LaunchState& s) {
    if (s.ready) return;
    int count = 0;
    check(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (count <= 0) throw std::runtime_error("no HIP device is visible");
    check(hipGetDevice(&s.device), "hipGetDevice");
    check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate");
    check(hipEventCreate(&s.start), "hipEventCreate");
    check(hipEventCreate(&s.stop), "hipEventCreate");
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s.device) != hipSuccess || khz <= 0)
        khz = 100000;  // 100 MHz constant clock of gfx9 parts
    s.wall_clock_khz = static_cast<double>(khz);
    s.ready = true;
}

template <typename T>
size_t append(std::vector<unsigned char>& blob, const std::vector<T>& v) {
    const size_t offset = (blob.size() + 255) & ~size_t(255);
    blob.resize(offset + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(blob.data() + offset, v.data(), v.size() * sizeof(T));
    return offset;
}

}  // namespace

DeviceTables::~DeviceTables() {
    if (orbit_block) (void)hipFree(orbit_block);
    for (void* p : linear_block) if (p) (void)hipFree(p);
    if (cubic_block) (void)hipFree(cubic_block);
    if (series_block) (void)hipFree(series_block);
}

namespace {

void* upload_blob(const std::vector<unsigned char>& blob, hipStream_t stream) {
    void* d = nullptr;
    check(hipMalloc(&d, std::max<size_t>(blob.size(), 256)), "hipMalloc tables");
    if (!blob.empty()) check(hipMemcpyAsync(d, blob.data(), blob.size(), hipMemcpyHostToDevice, stream), "upload");
    return d;
}

template <typename Step>
void* upload_levels(const std::vector<std::vector<Step>>& levels, LevelDir& dir, const Step*& base,
                    hipStream_t stream, size_t& bytes) {
    if (levels.size() > static_cast<size_t>(kMaxLevels))
        throw std::runtime_error("BLA hierarchy deeper than the device level directory");
    std::vector<Step> flat;
    dir = LevelDir{};
    dir.count = static_cast<int32_t>(levels.size());
    for (size_t l = 0; l < levels.size(); ++l) {
        dir.offset[l] = static_cast<int64_t>(flat.size());
        dir.size[l] = static_cast<int32_t>(levels[l].size());
        flat.insert(flat.end(), levels[l].begin(), levels[l].end());
    }
    void* d = nullptr;
    const size_t n = std::max<size_t>(flat.size(), 1) * sizeof(Step);
    check(hipMalloc(&d, n), "hipMalloc BLA");
    if (!flat.empty())
        check(hipMemcpyAsync(d, flat.data(), flat.size() * sizeof(Step), hipMemcpyHostToDevice, stream), "upload BLA");
    base = static_cast<const Step*>(d);
    bytes += n;
    return d;
}

}  // namespace

void ensure_device_tables(ReferenceContext& c) {
    std::lock_guard<std::mutex> lock(c.device_mutex);
    if (c.device) return;
    LaunchState& s = state();
    {
        std::lock_guard<std::mutex> slock(s.mutex);
        init_locked(s);
    }
    const uint64_t t0 = now_ns();
    auto t = std::make_unique<DeviceTables>();
    t->device = s.device;
    const OrbitData& o = *c.orbit;
    std::vector<unsigned char> blob;
    std::vector<double> re(o.scaled.size()), im(o.scaled.size());
    std::vector<int32_t> ex(o.scaled.size());
    for (size_t i = 0; i < o.scaled.size(); ++i) { re[i] = o.scaled[i].re; im[i] = o.scaled[i].im; ex[i] = o.scaled[i].e; }
    const size_t off_re = append(blob, re), off_im = append(blob, im), off_e = append(blob, ex);
    const size_t off_zr_m = append(blob, o.zr_m), off_zi_m = append(blob, o.zi_m);
    const size_t off_zr_e = append(blob, o.zr_e), off_zi_e = append(blob, o.zi_e);
    const size_t off_mm = append(blob, o.margin_m), off_me = append(blob, o.margin_e);
    const size_t off_gm = append(blob, o.glitch_m), off_ge = append(blob, o.glitch_e);
    const size_t off_rd = append(blob, o.real_double), off_id = append(blob, o.imag_double);
    t->orbit_block = upload_blob(blob, s.stream);
    t->bytes += blob.size();
    auto* base = static_cast<unsigned char*>(t->orbit_block);
    t->orbit.re = reinterpret_cast<const double*>(base + off_re);
    t->orbit.im = reinterpret_cast<const double*>(base + off_im);
    t->orbit.e = reinterpret_cast<const int32_t*>(base + off_e);
    t->orbit.zr_m = reinterpret_cast<const double*>(base + off_zr_m);
    t->orbit.zi_m = reinterpret_cast<const double*>(base + off_zi_m);
    t->orbit.zr_e = reinterpret_cast<const int32_t*>(base + off_zr_e);
    t->orbit.zi_e = reinterpret_cast<const int32_t*>(base + off_zi_e);
    t->orbit.margin_m = reinterpret_cast<const double*>(base + off_mm);
    t->orbit.margin_e = reinterpret_cast<const int32_t*>(base + off_me);
    t->orbit.glitch_m = reinterpret_cast<const double*>(base + off_gm);
    t->orbit.glitch_e = reinterpret_cast<const int32_t*>(base + off_ge);
    t->orbit.real_double = reinterpret_cast<const double*>(base + off_rd);
    t->orbit.imag_double = reinterpret_cast<const double*>(base + off_id);
    t->orbit.size = static_cast<int32_t>(o.scaled.size());
    t->orbit.map_end = c.map_end;
    // series coefficients
    std::vector<unsigned char> sblob;
    std::vector<double> sre, sim;
    std::vector<int32_t> se;
    for (const SC& k : c.series_coefficients) { sre.push_back(k.re); sim.push_back(k.im); se.push_back(k.e); }
    const size_t s_re = append(sblob, sre), s_im = append(sblob, sim), s_e = append(sblob, se);
    t->series_block = upload_blob(sblob, s.stream);
    auto* sbase = static_cast<unsigned char*>(t->series_block);
    t->series.re = reinterpret_cast<const double*>(sbase + s_re);
    t->series.im = reinterpret_cast<const double*>(sbase + s_im);
    t->series.e = reinterpret_cast<const int32_t*>(sbase + s_e);
    t->series.count = static_cast<int32_t>(c.series_coefficients.size());
    t->series.iteration = c.series_iteration;
    t->series.radius2_m = c.series_radius_squared.m;
    t->series.radius2_e = c.series_radius_squared.e;
    t->cubic_block = upload_levels(c.cubic.levels, t->cubic_dir, t->cubic, s.stream, t->bytes);
    for (int tier = 0; tier < 2; ++tier)
        t->linear_block[tier] = upload_levels(c.linear[tier].levels, t->linear_dir[tier], t->linear[tier],
                                              s.stream, t->bytes);
    check(hipStreamSynchronize(s.stream), "table upload");
    c.upload_ns = now_ns() - t0;
    c.device = std::move(t);
}

// One deep field render on the device, renderer.cpp:12672-13583 host part.
void render_field(ReferenceContext& c, const FieldRequest& r) {
    const int w = r.width, h = r.height;
    const FractalRenderOptions& opt = r.options;
    const FE zoom = parse_zoom(r.zoom_text, c.precision_bits);
    // viewport, renderer.cpp:12689-12705
    const double factor = c.coordinate_mode == 1 ? 4.0 : 2.8;
    const FE inverse_zoom = fe_div(fe_from_parts(1.0, 0), zoom);
    const FE view_height = fe_muld(inverse_zoom, factor);
    const FE view_width = fe_muld(view_height, static_cast<double>(w) / static_cast<double>(h));
    const FE pixel_spacing = fe_div(view_height, fe_from_parts(static_cast<double>(h), 0));
    const FE current_input_radius = fe_muld(inverse_zoom, factor);
    const bool covers = fe_compare(current_input_radius, c.input_radius) <= 0;
    const bool use_deep_linear = fe_compare(current_input_radius, c.deep_input_radius) <= 0;
    const bool use_ultra = fe_compare(current_input_radius, fe_muld(c.input_radius, 1.0e-40)) <= 0;
    // per-row / per-column offsets, renderer.cpp:12764-12783
    std::vector<double> off_m(static_cast<size_t>(w) + h);
    std::vector<int32_t> off_e(static_cast<size_t>(w) + h);
    for (int px = 0; px < w; ++px) {
        const double origin = opt.coordinate_mode == 1 ? static_cast<double>(w / 2) : static_cast<double>(w - 1) / 2.0;
        const FE v = fe_muld(view_width, (static_cast<double>(px) - origin) / static_cast<double>(w));
        off_m[static_cast<size_t>(px)] = v.m;
        off_e[static_cast<size_t>(px)] = v.e;
    }
    for (int py = 0; py < h; ++py) {
        const double origin = opt.coordinate_mode == 1 ? static_cast<double>(h / 2) : static_cast<double>(h - 1) / 2.0;
        const FE v = fe_muld(view_height, -(static_cast<double>(py) - origin) / static_cast<double>(h));
        off_m[static_cast<size_t>(w + py)] = v.m;
        off_e[static_cast<size_t>(w + py)] = v.e;
    }
    ensure_device_tables(c);
    const DeviceTables& t = *c.device;

    RenderParams p{};
    p.orbit = t.orbit;
    p.series = t.series;
    p.series.available = c.series_enabled && c.series_order >= opt.series_min_terms
        && c.series_order <= opt.series_max_terms
        && fe_compare(current_input_radius, fe_sqrt(c.series_radius_squared)) <= 0
        && c.series_iteration < r.max_iter;
    const int tier = use_deep_linear ? 1 : 0;
    p.linear = t.linear[tier];
    p.linear_dir = t.linear_dir[tier];
    p.cubic = t.cubic;
    p.cubic_dir = t.cubic_dir;
    p.output_bias = opt.output_bias;
    p.pixel_spacing_m = pixel_spacing.m;
    p.pixel_spacing_e = pixel_spacing.e;
    p.width = w;
    p.height = h;
    p.max_iter = r.max_iter;
    p.escape_radius_mode = opt.escape_radius_mode;
    p.strict = opt.strict != 0;
    p.allow_recovery = opt.allow_recovery != 0;
    p.disable_bla = !covers || opt.disable_bla != 0;
    p.cycle_enabled = opt.disable_cycle == 0 && (opt.strict == 0 || opt.strict_cycle != 0);
    p.strict_cycle = opt.strict_cycle != 0;
    p.max_bla_length = std::clamp(std::min(r.series_block, opt.max_bla_length), 2, 64);
    p.max_linear_length = std::clamp(std::min(r.series_block, opt.max_linear_bla_length), 2,
                                     use_ultra ? 4096 : (use_deep_linear ? 1024 : 64));
    p.approximation_order = std::clamp(r.series_order, 1, 3);

    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    const size_t pixels = static_cast<size_t>(w) * static_cast<size_t>(h);
    // offsets: doubles first, then the int32 exponents
    char* doff = static_cast<char*>(s.offsets.get(off_m.size() * 12 + 512));
    const size_t exp_offset = (off_m.size() * 8 + 255) & ~size_t(255);
    check(hipMemcpyAsync(doff, off_m.data(), off_m.size() * 8, hipMemcpyHostToDevice, s.stream), "offsets");
    check(hipMemcpyAsync(doff + exp_offset, off_e.data(), off_e.size() * 4, hipMemcpyHostToDevice, s.stream), "offsets");
    p.x_off_m = reinterpret_cast<const double*>(doff);
    p.y_off_m = p.x_off_m + w;
    p.x_off_e = reinterpret_cast<const int32_t*>(doff + exp_offset);
    p.y_off_e = p.x_off_e + w;

    const bool want_planes = r.planes != nullptr;
    if (r.device_pointers) {
        p.output = r.output;
        if (want_planes) {
            p.plane_iteration = r.planes->orbit_iteration;
            p.plane_phase = r.planes->phase;
            p.plane_de_x = r.planes->de_x;
            p.plane_de_y = r.planes->de_y;
            p.plane_test1 = r.planes->test1;
            p.plane_test2 = r.planes->test2;
        }
    } else {
        p.output = static_cast<float*>(s.output.get(pixels * sizeof(float)));
        if (want_planes) {
            char* base = static_cast<char*>(s.planes.get(pixels * 48));
            p.plane_iteration = reinterpret_cast<int64_t*>(base);
            p.plane_phase = reinterpret_cast<double*>(base + pixels * 8);
            p.plane_de_x = p.plane_phase + pixels;
            p.plane_de_y = p.plane_de_x + pixels;
            p.plane_test1 = p.plane_de_y + pixels;
            p.plane_test2 = p.plane_test1 + pixels;
        }
    }
    if (r.stats) {
        p.stats = static_cast<unsigned long long*>(s.stats.get(kStatsFields * 8));
        check(hipMemsetAsync(p.stats, 0, kStatsFields * 8, s.stream), "stats clear");
    }
    void* budget = s.budget.get(64);
    p.budget_start = static_cast<const unsigned long long*>(budget);
    p.budget_flag = reinterpret_cast<int*>(static_cast<char*>(budget) + 8);
    p.budget_ticks = opt.time_budget_ms > 0
        ? static_cast<unsigned long long>(opt.time_budget_ms * s.wall_clock_khz)
        : 0ULL;

    const uint64_t wall0 = now_ns();
    check(hipEventRecord(s.start, s.stream), "event");
    check(launch_field_kernel(p, want_planes, r.stats != nullptr, s.stream), "field kernel launch");
    check(hipEventRecord(s.stop, s.stream), "event");
    if (!r.device_pointers) {
        check(hipMemcpyAsync(r.output, p.output, pixels * sizeof(float), hipMemcpyDeviceToHost, s.stream), "D2H field");
        if (want_planes) {
            check(hipMemcpyAsync(r.planes->orbit_iteration, p.plane_iteration, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
            check(hipMemcpyAsync(r.planes->phase, p.plane_phase, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
            check(hipMemcpyAsync(r.planes->de_x, p.plane_de_x, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
            check(hipMemcpyAsync(r.planes->de_y, p.plane_de_y, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
            check(hipMemcpyAsync(r.planes->test1, p.plane_test1, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
            check(hipMemcpyAsync(r.planes->test2, p.plane_test2, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
        }
    }
    if (r.stats) {
        check(hipMemcpyAsync(r.stats, p.stats, kStatsFields * 8, hipMemcpyDeviceToHost, s.stream), "D2H stats");
    }
    check(hipStreamSynchronize(s.stream), "field render");
    float kernel_ms = 0.0f;
    check(hipEventElapsedTime(&kernel_ms, s.start, s.stop), "event time");
    s.timing.kernel_ms = kernel_ms;
    s.timing.total_ms = static_cast<double>(now_ns() - wall0) * 1e-6;
    s.timing.pixels = static_cast<double>(pixels);
    if (r.stats) r.stats[22] = static_cast<uint64_t>(static_cast<double>(kernel_ms) * 1e6);  // render_ns
}

LastTiming last_timing() {
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    return s.timing;
}

bool device_available() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return false;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return false;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

}  // namespace fav
