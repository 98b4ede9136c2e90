// device_host.cpp -- HBM residency of reference tables and the host side of
// one field launch: viewport geometry, per-row / per-column offsets,
// scratch buffers, the kernel launch on the library stream, HIP-event
// timing and the copy of results back to the caller.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <exception>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include <hip/hip_runtime_api.h>

#include "fav_host.h"
#include "fav_render.h"
#include "table_image.h"
#include "kfp_core.h"
#include "direct_core.h"
#include "direct_prepare.h"
#include "fx_core.h"
#include "sched_mlp.h"
#include "atlas_core.h"
#include "aurora_core.h"

namespace fav {

hipError_t launch_field_kernel(const RenderParams& p, bool planes, bool stats, int persistent_blocks,
                               hipStream_t stream, hipStream_t side, hipEvent_t fork, hipEvent_t join,
                               hipEvent_t plain_start = nullptr, hipEvent_t plain_stop = nullptr);
hipError_t launch_kfp_kernel(const kfp::KfpParams& params, void* scratch, int compute_units, hipStream_t stream);
hipError_t kfp_noalloc_scratch(int device, unsigned char** out, size_t* lut_offset, size_t* bytes);
hipError_t launch_direct_kernel(const direct::DirectParams& p, bool planes, hipStream_t stream);
hipError_t launch_alt_kernel(const AltRenderParams& p, hipStream_t stream);
hipError_t launch_fx_kernel(const RenderParams& p, const fx::FxOut& out, hipStream_t stream);
hipError_t launch_fx_alt_kernel(const AltRenderParams& p, const fx::FxOut& out, hipStream_t stream);
hipError_t launch_batch(const BatchLaunch& q, int fast_blocks, int persistent_blocks, hipStream_t stream);
hipError_t launch_atlas_compose(const atlas::ComposeParams& a, hipStream_t stream);
hipError_t launch_aurora(const aurora::Params& p, hipStream_t stream);
hipError_t launch_atlas_planes(const atlas::PlanesCompose& a, hipStream_t stream);
hipError_t launch_kfp_rect_atlas(kfp::KfpParams* host_params, kfp::KfpParams* params, const kfp::RectAtlas& g,
                                 void* scratch, int compute_units, hipStream_t stream);

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
    hipStream_t side = nullptr;  // default priority, like `stream`
    hipEvent_t fork = nullptr, join = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
    hipEvent_t plain_start = nullptr, plain_stop = nullptr;  // around the plain pass
    hipEvent_t mark[3] = {nullptr, nullptr, nullptr};
    DeviceBuffer offsets, output, planes, stats, budget, state, kfp_field, kfp_lut, kfp_rgb, kfp_scratch, direct_coords, fx_counts, tile_order, work,
        mlp_y, mlp_act, mlp_delta, mlp_err2, mlp_partial, mlp_grad, mlp_features, mlp_weights, preview_offsets,
        preview_output, preview_work, preview_state, preview_stats,
        atlas_tiles, atlas_axes, atlas_field, points, aurora_in, aurora_out, kfp_planes;
    int compute_units = 256;
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
    check(hipStreamCreateWithFlags(&s.side, hipStreamNonBlocking), "hipStreamCreate");
    check(hipEventCreateWithFlags(&s.fork, hipEventDisableTiming), "hipEventCreate");
    check(hipEventCreateWithFlags(&s.join, hipEventDisableTiming), "hipEventCreate");
    check(hipEventCreate(&s.start), "hipEventCreate");
    check(hipEventCreate(&s.stop), "hipEventCreate");
    check(hipEventCreate(&s.plain_start), "hipEventCreate");
    check(hipEventCreate(&s.plain_stop), "hipEventCreate");
    for (hipEvent_t& e : s.mark) check(hipEventCreate(&e), "hipEventCreate");
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s.device) != hipSuccess || khz <= 0)
        khz = 100000;  // 100 MHz constant clock of gfx9 parts
    s.wall_clock_khz = static_cast<double>(khz);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s.device) == hipSuccess && cus > 0)
        s.compute_units = cus;
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
    if (alt_block) (void)hipFree(alt_block);
    if (fast_block) (void)hipFree(fast_block);
    for (auto& tier : rung_block)
        for (auto& cap : tier)
            for (void* p : cap)
                if (p) (void)hipFree(p);
    for (auto& table : walk_block)
        for (void* p : table)
            if (p) (void)hipFree(p);
}

namespace {

template <typename T>
void upload_vector(const std::vector<T>& v, void*& block, const T*& view, hipStream_t stream, size_t& bytes) {
    const size_t n = std::max<size_t>(v.size(), 1) * sizeof(T);
    check(hipMalloc(&block, n), "hipMalloc tables");
    if (!v.empty()) check(hipMemcpyAsync(block, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream), "upload");
    view = static_cast<const T*>(block);
    bytes += n;
}

// A reference's device tables as host images: the orbit arrays packed into
// one blob (each at a 256-byte boundary), the series coefficients into
// another, and the step tables.  Host work only, safe on any thread.
struct HostTables {
    TableImage img;
    std::vector<unsigned char> orbit_blob, series_blob;
    size_t off_re = 0, off_im = 0, off_e = 0, off_zr_m = 0, off_zi_m = 0, off_zr_e = 0, off_zi_e = 0;
    size_t off_mm = 0, off_me = 0, off_gm = 0, off_ge = 0, off_rd = 0, off_id = 0, off_znm = 0, off_zne = 0;
    size_t off_small = 0, off_flags = 0, off_am = 0, off_ae = 0;
    size_t s_re = 0, s_im = 0, s_e = 0;
};

void build_host_tables(const ReferenceContext& c, HostTables& h) {
    make_table_image(c, h.img);
    const OrbitData& o = *c.orbit;
    std::vector<unsigned char>& blob = h.orbit_blob;
    h.off_re = append(blob, h.img.re);
    h.off_im = append(blob, h.img.im);
    h.off_e = append(blob, h.img.e);
    h.off_zr_m = append(blob, o.zr_m);
    h.off_zi_m = append(blob, o.zi_m);
    h.off_zr_e = append(blob, o.zr_e);
    h.off_zi_e = append(blob, o.zi_e);
    h.off_mm = append(blob, o.margin_m);
    h.off_me = append(blob, o.margin_e);
    h.off_gm = append(blob, o.glitch_m);
    h.off_ge = append(blob, o.glitch_e);
    h.off_rd = append(blob, o.real_double);
    h.off_id = append(blob, o.imag_double);
    h.off_znm = append(blob, o.zn_m);
    h.off_zne = append(blob, o.zn_e);
    h.off_small = append(blob, o.small_exp);
    h.off_flags = append(blob, o.flags);
    h.off_am = append(blob, o.alt_margin_m);
    h.off_ae = append(blob, o.alt_margin_e);
    h.s_re = append(h.series_blob, h.img.series_re);
    h.s_im = append(h.series_blob, h.img.series_im);
    h.s_e = append(h.series_blob, h.img.series_e);
}

// The parts of HostTables in upload order.
enum TablePart { kPartOrbit, kPartAlt, kPartFast, kPartSeries, kPartCubic, kPartLinear0, kPartLinear1, kTableParts };

// The host images copied into one pinned (page-locked) buffer, each part at
// a 256-byte boundary, so that the copies to HBM are asynchronous DMA
// transfers.  `done` marks the end of the copies once they are issued.
struct StagedTables {
    unsigned char* pinned = nullptr;
    size_t offset[kTableParts] = {}, bytes[kTableParts] = {};
    hipEvent_t done = nullptr;
    StagedTables() = default;
    StagedTables(const StagedTables&) = delete;
    StagedTables& operator=(const StagedTables&) = delete;
    ~StagedTables() { release(); }
    void release() {
        if (done) (void)hipEventDestroy(done);
        if (pinned) (void)hipHostFree(pinned);
        done = nullptr;
        pinned = nullptr;
    }
};

template <typename T>
std::pair<const void*, size_t> table_part(const std::vector<T>& v) {
    return {v.data(), v.size() * sizeof(T)};
}

void stage_tables(const HostTables& h, StagedTables& s) {
    const std::pair<const void*, size_t> parts[kTableParts] = {
        table_part(h.orbit_blob), table_part(h.img.alt), table_part(h.img.fast), table_part(h.series_blob),
        table_part(h.img.cubic), table_part(h.img.linear[0]), table_part(h.img.linear[1])};
    size_t total = 0;
    for (int i = 0; i < kTableParts; ++i) {
        s.offset[i] = total;
        s.bytes[i] = parts[i].second;
        total = (total + parts[i].second + 255) & ~size_t(255);
    }
    check(hipHostMalloc(reinterpret_cast<void**>(&s.pinned), std::max<size_t>(total, 256), hipHostMallocPortable),
          "hipHostMalloc table staging");
    for (int i = 0; i < kTableParts; ++i)
        if (s.bytes[i]) std::memcpy(s.pinned + s.offset[i], parts[i].first, s.bytes[i]);
}

// HBM blocks for the staged parts and their copies on `stream`; the views
// of the returned tables point into those blocks.  Records `s.done` after
// the copies.
std::unique_ptr<DeviceTables> upload_staged(const ReferenceContext& c, const HostTables& h, StagedTables& s,
                                            int device, hipStream_t stream) {
    auto t = std::make_unique<DeviceTables>();
    t->device = device;
    auto copy = [&](int part, void*& block) {
        const size_t n = std::max<size_t>(s.bytes[part], 1);
        check(hipMalloc(&block, n), "hipMalloc tables");
        if (s.bytes[part])
            check(hipMemcpyAsync(block, s.pinned + s.offset[part], s.bytes[part], hipMemcpyHostToDevice, stream),
                  "table upload");
        t->bytes += n;
        return static_cast<const unsigned char*>(block);
    };
    const unsigned char* base = copy(kPartOrbit, t->orbit_block);
    t->orbit.re = reinterpret_cast<const double*>(base + h.off_re);
    t->orbit.im = reinterpret_cast<const double*>(base + h.off_im);
    t->orbit.e = reinterpret_cast<const int32_t*>(base + h.off_e);
    t->orbit.zr_m = reinterpret_cast<const double*>(base + h.off_zr_m);
    t->orbit.zi_m = reinterpret_cast<const double*>(base + h.off_zi_m);
    t->orbit.zr_e = reinterpret_cast<const int32_t*>(base + h.off_zr_e);
    t->orbit.zi_e = reinterpret_cast<const int32_t*>(base + h.off_zi_e);
    t->orbit.margin_m = reinterpret_cast<const double*>(base + h.off_mm);
    t->orbit.margin_e = reinterpret_cast<const int32_t*>(base + h.off_me);
    t->orbit.glitch_m = reinterpret_cast<const double*>(base + h.off_gm);
    t->orbit.glitch_e = reinterpret_cast<const int32_t*>(base + h.off_ge);
    t->orbit.real_double = reinterpret_cast<const double*>(base + h.off_rd);
    t->orbit.imag_double = reinterpret_cast<const double*>(base + h.off_id);
    t->orbit.zn_m = reinterpret_cast<const double*>(base + h.off_znm);
    t->orbit.zn_e = reinterpret_cast<const int32_t*>(base + h.off_zne);
    t->orbit.small_exp = reinterpret_cast<const int32_t*>(base + h.off_small);
    t->orbit.flags = reinterpret_cast<const int32_t*>(base + h.off_flags);
    t->orbit.size = static_cast<int32_t>(c.orbit->scaled.size());
    t->orbit.map_end = c.map_end;
    t->alt_margin_m = reinterpret_cast<const double*>(base + h.off_am);
    t->alt_margin_e = reinterpret_cast<const int32_t*>(base + h.off_ae);
    t->alt = reinterpret_cast<const AltStepDev*>(copy(kPartAlt, t->alt_block));
    t->orbit.fast = reinterpret_cast<const OrbitFastDev*>(copy(kPartFast, t->fast_block));
    t->alt_dir = h.img.alt_dir;
    const unsigned char* sbase = copy(kPartSeries, t->series_block);
    t->series.re = reinterpret_cast<const double*>(sbase + h.s_re);
    t->series.im = reinterpret_cast<const double*>(sbase + h.s_im);
    t->series.e = reinterpret_cast<const int32_t*>(sbase + h.s_e);
    t->cubic = reinterpret_cast<const CubicStepDev*>(copy(kPartCubic, t->cubic_block));
    t->cubic_dir = h.img.cubic_dir;
    for (int tier = 0; tier < 2; ++tier) {
        t->linear[tier] = reinterpret_cast<const LinearStepDev*>(copy(kPartLinear0 + tier, t->linear_block[tier]));
        t->linear_dir[tier] = h.img.linear_dir[tier];
    }
    check(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "hipEventCreate");
    check(hipEventRecord(s.done, stream), "hipEventRecord");
    return t;
}

}  // namespace

// Host images, pinned staging and the copies on the side stream (default
// priority, like the compute stream), waited for before returning.
void ensure_device_tables(ReferenceContext& c) {
    std::lock_guard<std::mutex> lock(c.device_mutex);
    if (c.device) return;
    LaunchState& s = state();
    {
        std::lock_guard<std::mutex> slock(s.mutex);
        init_locked(s);
    }
    const uint64_t t0 = now_ns();
    HostTables h;
    build_host_tables(c, h);
    StagedTables staged;
    stage_tables(h, staged);
    auto t = upload_staged(c, h, staged, s.device, s.side);
    check(hipEventSynchronize(staged.done), "table upload");
    c.upload_ns = now_ns() - t0;
    c.device = std::move(t);
}

// The plain-double block loop's rung table of linear tier `tier` for level
// cap `cap` and approximation order 3 (order3) or 1 (table_image.cpp
// make_plain_tables), built and copied to HBM on first use; null when the
// reference has none.
void ensure_plain_tables(ReferenceContext& c, int tier, int cap, bool order3, const RungDev*& rungs) {
    std::lock_guard<std::mutex> lock(c.device_mutex);
    DeviceTables& t = *c.device;
    const int o = order3 ? 1 : 0;
    if (!t.plain_built[tier][cap][o]) {
        std::vector<RungDev> r;
        make_plain_tables(c, tier, cap, order3, r);
        if (!r.empty()) {
            LaunchState& s = state();
            upload_vector(r, t.rung_block[tier][cap][o], t.rungs[tier][cap][o], s.stream, t.bytes);
            check(hipStreamSynchronize(s.stream), "plain table upload");
        }
        t.plain_built[tier][cap][o] = true;
    }
    rungs = t.rungs[tier][cap][o];
}

// The walk of the pixel on the reference through `table` (linear tier 0 or
// 1, cubic 2) at a block-length limit of floor(log2) `level`.
void ensure_ref_walk(ReferenceContext& c, int table, int level, const RefWalkDev*& walk) {
    std::lock_guard<std::mutex> lock(c.device_mutex);
    DeviceTables& t = *c.device;
    if (!t.walk_built[table][level]) {
        std::vector<RefWalkDev> w;
        make_ref_walk(c, table, 1 << level, w);
        if (!w.empty()) {
            LaunchState& s = state();
            upload_vector(w, t.walk_block[table][level], t.walks[table][level], s.stream, t.bytes);
            check(hipStreamSynchronize(s.stream), "walk table upload");
        }
        t.walk_built[table][level] = true;
    }
    walk = t.walks[table][level];
}

// One deep field render on the device, host part of renderer.cpp:12672-13583.
// A point render's offsets in HBM: re then im (doubles), then exponents.
// The host copies stay alive in `staging` until the caller synchronises.
struct PointStaging {
    std::vector<double> m;
    std::vector<int32_t> e;
};

void upload_points(LaunchState& s, const FieldRequest& r, PointStaging& staging, const double*& re,
                   const double*& im, const int32_t*& e) {
    const size_t n = static_cast<size_t>(r.point_count);
    staging.m.resize(2 * n);
    staging.e.resize(n);
    for (size_t i = 0; i < n; ++i) {
        staging.m[i] = r.points[i].re;
        staging.m[n + i] = r.points[i].im;
        staging.e[i] = r.points[i].e;
    }
    const size_t e_at = (2 * n * 8 + 255) & ~size_t(255);
    char* dp = static_cast<char*>(s.points.get(e_at + n * 4 + 256));
    check(hipMemcpyAsync(dp, staging.m.data(), 2 * n * 8, hipMemcpyHostToDevice, s.stream), "H2D points");
    check(hipMemcpyAsync(dp + e_at, staging.e.data(), n * 4, hipMemcpyHostToDevice, s.stream), "H2D points");
    re = reinterpret_cast<const double*>(dp);
    im = re + n;
    e = reinterpret_cast<const int32_t*>(dp + e_at);
}

namespace {

// Which rung table and on-reference walk a field render binds: the linear
// tier, the level cap (floor(log2) of the linear length limit) and the
// order-3 flag of the rung table; the walk's table (linear tier, or 2 for
// the cubic table) and floor(log2) of its length limit.
struct FieldTableKeys {
    int tier, cap, order3, walk_table, walk_level;
};

FieldTableKeys field_table_keys(const PreparedRender& prep) {
    const RenderParams& q = prep.params;
    const bool linear = q.approximation_order <= 1;
    return FieldTableKeys{prep.linear_tier, 31 - __builtin_clz(static_cast<unsigned>(q.max_linear_length)),
                          q.approximation_order >= 3 ? 1 : 0, linear ? prep.linear_tier : 2,
                          31 - __builtin_clz(static_cast<unsigned>(linear ? q.max_linear_length : q.max_bla_length))};
}

// The host images of a field's rung table and walk, built off the render
// thread (render_segment's workers) and uploaded by install_prebuilt.
struct PrebuiltTables {
    FieldTableKeys keys{};
    std::vector<RungDev> rungs;
    std::vector<RefWalkDev> walk;
};

void prebuild_field_tables(const ReferenceContext& c, const PreparedRender& prep, PrebuiltTables& out) {
    out.keys = field_table_keys(prep);
    make_plain_tables(c, out.keys.tier, out.keys.cap, out.keys.order3 != 0, out.rungs);
    if (out.keys.walk_level < kMaxLevels) make_ref_walk(c, out.keys.walk_table, 1 << out.keys.walk_level, out.walk);
}

// Uploads prebuilt tables the reference does not hold yet (its device tables
// must exist), so that the render's ensure_* calls find them built.
void install_prebuilt(ReferenceContext& c, const PrebuiltTables& pre) {
    std::lock_guard<std::mutex> lock(c.device_mutex);
    DeviceTables& t = *c.device;
    LaunchState& s = state();
    const FieldTableKeys& k = pre.keys;
    bool uploaded = false;
    if (!t.plain_built[k.tier][k.cap][k.order3]) {
        if (!pre.rungs.empty()) {
            upload_vector(pre.rungs, t.rung_block[k.tier][k.cap][k.order3], t.rungs[k.tier][k.cap][k.order3],
                          s.stream, t.bytes);
            uploaded = true;
        }
        t.plain_built[k.tier][k.cap][k.order3] = true;
    }
    if (k.walk_level < kMaxLevels && !t.walk_built[k.walk_table][k.walk_level]) {
        if (!pre.walk.empty()) {
            upload_vector(pre.walk, t.walk_block[k.walk_table][k.walk_level], t.walks[k.walk_table][k.walk_level],
                          s.stream, t.bytes);
            uploaded = true;
        }
        t.walk_built[k.walk_table][k.walk_level] = true;
    }
    if (uploaded) check(hipStreamSynchronize(s.stream), "prebuilt table upload");
}

// Launch parameters of one Mandelbrot field: the per-frame values of `prep`
// with the reference's device tables bound (uploaded first, with the
// plain-double rung tables of the selected tier on first use).
RenderParams bind_field_tables(ReferenceContext& c, const PreparedRender& prep) {
    ensure_device_tables(c);
    const DeviceTables& t = *c.device;
    RenderParams p = prep.params;
    p.orbit = t.orbit;
    const SeriesDev scalar_series = p.series;
    p.series = t.series;
    p.series.count = scalar_series.count;
    p.series.iteration = scalar_series.iteration;
    p.series.available = scalar_series.available;
    p.series.radius2_m = scalar_series.radius2_m;
    p.series.radius2_e = scalar_series.radius2_e;
    p.linear = t.linear[prep.linear_tier];
    p.linear_dir = t.linear_dir[prep.linear_tier];
    {
        // floor(log2) of the block length limit (at least 2)
        const int cap = 31 - __builtin_clz(static_cast<unsigned>(p.max_linear_length));
        ensure_plain_tables(c, prep.linear_tier, cap, p.approximation_order >= 3, p.rungs);
        p.plain_top = cap;
        p.plain_pass = p.rungs != nullptr && p.fast_linear ? 1 : 0;
    }
    p.cubic = t.cubic;
    p.cubic_dir = t.cubic_dir;
    {
        // the on-reference walk through the table of the approximation
        // order; its lookup reads the length limit only through floor(log2)
        const bool linear = p.approximation_order <= 1;
        const int level = 31 - __builtin_clz(static_cast<unsigned>(linear ? p.max_linear_length : p.max_bla_length));
        if (level < kMaxLevels) {
            ensure_ref_walk(c, linear ? prep.linear_tier : 2, level, p.ref_walk);
            p.ref_walk_limit = 1 << level;
        }
    }
    return p;
}

// The per-frame uploads of one field on the library stream (lock held):
// pixel offsets into `offsets`, the points of a point render, the pixel on
// the reference, the time budget and the state store `state`.  `prep` and
// `points` must stay alive until the stream has run the copies.
void bind_field_frame(RenderParams& p, const PreparedRender& prep, const FieldRequest& r, LaunchState& s,
                      DeviceBuffer& offsets, DeviceBuffer& state_buffer, PointStaging& points) {
    const int w = r.width, h = r.height;
    const size_t pixels = static_cast<size_t>(w) * static_cast<size_t>(h);
    const std::vector<double>& off_m = prep.offset_m;
    const std::vector<int32_t>& off_e = prep.offset_e;
    // offsets: doubles first, then the int32 exponents
    char* doff = static_cast<char*>(offsets.get(off_m.size() * 12 + 512));
    const size_t exp_offset = (off_m.size() * 8 + 255) & ~size_t(255);
    check(hipMemcpyAsync(doff, off_m.data(), off_m.size() * 8, hipMemcpyHostToDevice, s.stream), "offsets");
    check(hipMemcpyAsync(doff + exp_offset, off_e.data(), off_e.size() * 4, hipMemcpyHostToDevice, s.stream), "offsets");
    p.x_off_m = reinterpret_cast<const double*>(doff);
    p.y_off_m = p.x_off_m + w;
    p.x_off_e = reinterpret_cast<const int32_t*>(doff + exp_offset);
    p.y_off_e = p.x_off_e + w;
    p.reference_pixel = -1;
    if (r.points) {
        upload_points(s, r, points, p.point_re, p.point_im, p.point_e);
        p.point_count = r.point_count;
        // a point with dc = 0 stays on fast_kernel's path
    } else if (p.plain_pass) {
        // the pixel whose column and row offsets are both zero (dc = 0)
        int zx = -1, zy = -1;
        for (int x = 0; x < w && zx < 0; ++x) if (off_m[static_cast<size_t>(x)] == 0.0) zx = x;
        for (int y = 0; y < h && zy < 0; ++y) if (off_m[static_cast<size_t>(w + y)] == 0.0) zy = y;
        if (zx >= 0 && zy >= 0) p.reference_pixel = static_cast<int64_t>(zy) * w + zx;
    }
    void* budget = s.budget.get(256);
    p.budget_start = static_cast<const unsigned long long*>(budget);
    p.budget_flag = reinterpret_cast<int*>(static_cast<char*>(budget) + 8);
    p.budget_ticks = r.options.time_budget_ms > 0
        ? static_cast<unsigned long long>(r.options.time_budget_ms * s.wall_clock_khz)
        : 0ULL;
    bind_state_store(p, state_buffer.get(pixels * kStateBytesPerPixel + 20 * 256), pixels);
}

// Uploads a 32x32 tile order (checked by the caller) on the library stream.
void bind_tile_order(RenderParams& p, LaunchState& s, const uint32_t* order, size_t count) {
    void* d = s.tile_order.get(count * 4);
    check(hipMemcpyAsync(d, order, count * 4, hipMemcpyHostToDevice, s.stream), "H2D tile order");
    p.tile_order = static_cast<const uint32_t*>(d);
    p.tile_count = static_cast<int32_t>(count);
}

}  // namespace

namespace {

// One Mandelbrot field render with `p` bound to the reference's tables
// (bind_field_tables over `prep`); lock held.
void render_field_locked(const FieldRequest& r, const PreparedRender& prep, RenderParams p, LaunchState& s) {
    const int w = r.width, h = r.height;
    const size_t pixels = static_cast<size_t>(w) * static_cast<size_t>(h);
    PointStaging points;
    bind_field_frame(p, prep, r, s, s.offsets, s.state, points);
    const size_t out_count = r.points ? static_cast<size_t>(r.point_count) : pixels;
    if (r.tile_order) bind_tile_order(p, s, r.tile_order, r.tile_count);

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
    // the work map needs the statistics kernels (their per-pixel counters)
    const bool counters = r.stats != nullptr || r.work != nullptr;
    if (counters) {
        p.stats = static_cast<unsigned long long*>(s.stats.get(kStatsFields * 8));
        check(hipMemsetAsync(p.stats, 0, kStatsFields * 8, s.stream), "stats clear");
    }
    if (r.work) {
        p.work = r.device_pointers ? r.work : static_cast<int32_t*>(s.work.get(pixels * 4));
        check(hipMemsetAsync(p.work, 0, pixels * 4, s.stream), "work clear");
    }

    const uint64_t wall0 = now_ns();
    check(hipEventRecord(s.start, s.stream), "event");
    check(launch_field_kernel(p, want_planes, counters, s.compute_units * 8, s.stream, s.side, s.fork, s.join,
                              s.plain_start, s.plain_stop),
          "field kernel launch");
    check(hipEventRecord(s.stop, s.stream), "event");
    if (!r.device_pointers) {
        check(hipMemcpyAsync(r.output, p.output, out_count * sizeof(float), hipMemcpyDeviceToHost, s.stream),
              "D2H field");
        if (want_planes) {
            check(hipMemcpyAsync(r.planes->orbit_iteration, p.plane_iteration, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
            check(hipMemcpyAsync(r.planes->phase, p.plane_phase, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
            check(hipMemcpyAsync(r.planes->de_x, p.plane_de_x, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
            check(hipMemcpyAsync(r.planes->de_y, p.plane_de_y, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
            check(hipMemcpyAsync(r.planes->test1, p.plane_test1, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
            check(hipMemcpyAsync(r.planes->test2, p.plane_test2, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
        }
        if (r.work) check(hipMemcpyAsync(r.work, p.work, pixels * 4, hipMemcpyDeviceToHost, s.stream), "D2H work");
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
    s.timing.plain_ms = 0.0;
    if (p.plain_pass) {
        float plain_ms = 0.0f;
        check(hipEventElapsedTime(&plain_ms, s.plain_start, s.plain_stop), "event time");
        s.timing.plain_ms = plain_ms;
    }
    if (r.stats) r.stats[22] = static_cast<uint64_t>(static_cast<double>(kernel_ms) * 1e6);  // render_ns
}

}  // namespace

void render_field(ReferenceContext& c, const FieldRequest& r) {
    PreparedRender prep;
    prepare_render(c, r, prep);
    const RenderParams p = bind_field_tables(c, prep);
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    render_field_locked(r, prep, p, s);
}

void colourise_kfp_on_device(const kfp::KfpParams& host, bool device_pointers, const FractalKfpPlanes* planes) {
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    const size_t pixels = static_cast<size_t>(host.width) * static_cast<size_t>(host.height);
    kfp::KfpParams k = host;
    if (!device_pointers) {
        void* f = s.kfp_field.get(pixels * sizeof(float));
        check(hipMemcpyAsync(f, host.field, pixels * sizeof(float), hipMemcpyHostToDevice, s.stream), "H2D field");
        k.field = static_cast<const float*>(f);
        k.output = static_cast<uint8_t*>(s.kfp_rgb.get(pixels * 3));
    }
    if (planes && !device_pointers) {  // the present planes, 8 bytes per pixel each, then the texture
        const void* sources[9] = {planes->orbit_iteration, planes->iteration, planes->bailout, planes->transition,
                                  planes->phase, planes->de_x, planes->de_y, planes->test1, planes->test2};
        const void** targets[9] = {reinterpret_cast<const void**>(&k.orbit_iteration),
                                   reinterpret_cast<const void**>(&k.iteration),
                                   reinterpret_cast<const void**>(&k.bailout),
                                   reinterpret_cast<const void**>(&k.transition),
                                   reinterpret_cast<const void**>(&k.phase),
                                   reinterpret_cast<const void**>(&k.de_x),
                                   reinterpret_cast<const void**>(&k.de_y),
                                   reinterpret_cast<const void**>(&k.test1),
                                   reinterpret_cast<const void**>(&k.test2)};
        const size_t plane_bytes = (pixels * 8 + 255) & ~size_t(255);
        const size_t texture_bytes = k.texture_active
            ? static_cast<size_t>(planes->texture_stride) * static_cast<size_t>(planes->texture_height) : 0;
        char* block = static_cast<char*>(s.kfp_planes.get(plane_bytes * 9 + texture_bytes + 256));
        for (int i = 0; i < 9; ++i) {
            *targets[i] = nullptr;
            if (!sources[i]) continue;
            check(hipMemcpyAsync(block + plane_bytes * i, sources[i], pixels * 8, hipMemcpyHostToDevice, s.stream),
                  "H2D KFP plane");
            *targets[i] = block + plane_bytes * i;
        }
        k.texture = nullptr;
        if (texture_bytes) {
            check(hipMemcpyAsync(block + plane_bytes * 9, planes->texture_rgb, texture_bytes, hipMemcpyHostToDevice,
                                 s.stream),
                  "H2D KFP texture");
            k.texture = reinterpret_cast<const uint8_t*>(block + plane_bytes * 9);
        }
    }
    void* l = s.kfp_lut.get(static_cast<size_t>(host.lut_size) * 3);
    check(hipMemcpyAsync(l, host.lut, static_cast<size_t>(host.lut_size) * 3, hipMemcpyHostToDevice, s.stream), "H2D lut");
    k.lut = static_cast<const uint8_t*>(l);
    void* scratch = s.kfp_scratch.get(64);
    const uint64_t wall0 = now_ns();
    check(hipEventRecord(s.start, s.stream), "event");
    check(launch_kfp_kernel(k, scratch, s.compute_units, s.stream), "KFP colour launch");
    check(hipEventRecord(s.stop, s.stream), "event");
    if (!device_pointers)
        check(hipMemcpyAsync(host.output, k.output, pixels * 3, hipMemcpyDeviceToHost, s.stream), "D2H rgb");
    check(hipStreamSynchronize(s.stream), "KFP colour");
    float kernel_ms = 0.0f;
    check(hipEventElapsedTime(&kernel_ms, s.start, s.stop), "event time");
    s.timing.kernel_ms = kernel_ms;
    s.timing.total_ms = static_cast<double>(now_ns() - wall0) * 1e-6;
    s.timing.pixels = static_cast<double>(pixels);
}

// The KFP colour pass of fractal_hip_colourise_kfp_device without any
// device allocation: the LUT is staged into, and the range reduction works
// in, the file-scope scratch array of kfp_kernel.hip (its address looked up
// once per device).  Field and RGB are the caller's device pointers.  Calls
// on one device share that scratch, so overlapping calls race.
void colourise_kfp_device_noalloc(const kfp::KfpParams& host) {
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    unsigned char* scratch = nullptr;
    size_t lut_offset = 0, bytes = 0;
    check(kfp_noalloc_scratch(s.device, &scratch, &lut_offset, &bytes), "hipGetSymbolAddress(scratch)");
    const size_t lut_bytes = static_cast<size_t>(host.lut_size) * 3;
    if (lut_offset + lut_bytes > bytes)
        throw std::runtime_error("allocation-free KFP colour: the LUT exceeds the 65536-entry scratch");
    kfp::KfpParams k = host;
    check(hipMemcpyAsync(scratch + lut_offset, host.lut, lut_bytes, hipMemcpyHostToDevice, s.stream), "H2D lut");
    k.lut = scratch + lut_offset;
    const size_t pixels = static_cast<size_t>(host.width) * static_cast<size_t>(host.height);
    const uint64_t wall0 = now_ns();
    check(hipEventRecord(s.start, s.stream), "event");
    check(launch_kfp_kernel(k, scratch, s.compute_units, s.stream), "KFP colour launch");
    check(hipEventRecord(s.stop, s.stream), "event");
    check(hipStreamSynchronize(s.stream), "KFP colour");
    float kernel_ms = 0.0f;
    check(hipEventElapsedTime(&kernel_ms, s.start, s.stop), "event time");
    s.timing.kernel_ms = kernel_ms;
    s.timing.total_ms = static_cast<double>(now_ns() - wall0) * 1e-6;
    s.timing.pixels = static_cast<double>(pixels);
}

void atlas_colourise_kfp_planes(const atlas::PlanesCompose& host, const PlaneTile* tiles, const atlas::Axis* axes,
                                const kfp::KfpParams& colour, const FractalKfpPlanes* texture, uint8_t* output) {
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    const size_t pixels = static_cast<size_t>(host.width) * static_cast<size_t>(host.height);
    atlas::PlanesCompose a = host;
    auto align = [](size_t bytes) { return (bytes + 255) & ~size_t(255); };
    // inputs: per tile the field and its present planes; then the four axes
    size_t total = 0;
    for (int t = 0; t < 2; ++t) {
        if (!tiles[t].field) continue;
        const size_t n = static_cast<size_t>(tiles[t].width) * static_cast<size_t>(tiles[t].height);
        total += align(n * 4) + 9 * align(n * 8);
    }
    for (int i = 0; i < 4; ++i) total += 3 * align(axes[i].weight.size() * 4);
    char* in = static_cast<char*>(s.atlas_tiles.get(total + 256));
    size_t at = 0;
    auto put = [&](const void* src, size_t bytes) -> const void* {
        if (!src || !bytes) return nullptr;
        check(hipMemcpyAsync(in + at, src, bytes, hipMemcpyHostToDevice, s.stream), "H2D atlas planes input");
        const void* d = in + at;
        at += align(bytes);
        return d;
    };
    atlas::PlaneSource* sources[2] = {&a.primary, &a.secondary};
    for (int t = 0; t < 2; ++t) {
        atlas::PlaneSource& src = *sources[t];
        src.field = nullptr;
        for (auto& q : src.ints) q = nullptr;
        for (auto& q : src.doubles) q = nullptr;
        if (!tiles[t].field) continue;
        const size_t n = static_cast<size_t>(tiles[t].width) * static_cast<size_t>(tiles[t].height);
        src.field = static_cast<const float*>(put(tiles[t].field, n * 4));
        const FractalKfpPlanes* p = tiles[t].planes;
        const int64_t* ints[2] = {p->orbit_iteration, p->iteration};
        const double* doubles[7] = {p->bailout, p->transition, p->phase, p->de_x, p->de_y, p->test1, p->test2};
        for (int k = 0; k < 2; ++k) src.ints[k] = static_cast<const int64_t*>(put(ints[k], ints[k] ? n * 8 : 0));
        for (int k = 0; k < 7; ++k)
            src.doubles[k] = static_cast<const double*>(put(doubles[k], doubles[k] ? n * 8 : 0));
    }
    atlas::AxisDev* views[4] = {&a.primary.x, &a.primary.y, &a.secondary.x, &a.secondary.y};
    for (int i = 0; i < 4; ++i) {
        const size_t n = axes[i].weight.size();
        *views[i] = atlas::AxisDev{static_cast<const int32_t*>(put(axes[i].index0.data(), n * 4)),
                                   static_cast<const int32_t*>(put(axes[i].index1.data(), n * 4)),
                                   static_cast<const float*>(put(axes[i].weight.data(), n * 4))};
    }
    // outputs: the field and each plane that either tile carries
    char* out = static_cast<char*>(s.atlas_field.get(align(pixels * 4) + 9 * align(pixels * 8) + 256));
    a.out_field = reinterpret_cast<float*>(out);
    for (int k = 0; k < 2; ++k)
        a.out_ints[k] = (a.primary.ints[k] || a.secondary.ints[k])
            ? reinterpret_cast<int64_t*>(out + align(pixels * 4) + static_cast<size_t>(k) * align(pixels * 8)) : nullptr;
    for (int k = 0; k < 7; ++k)
        a.out_doubles[k] = (a.primary.doubles[k] || a.secondary.doubles[k])
            ? reinterpret_cast<double*>(out + align(pixels * 4) + static_cast<size_t>(k + 2) * align(pixels * 8))
            : nullptr;
    kfp::KfpParams k = colour;
    k.field = a.out_field;
    k.orbit_iteration = a.out_ints[0];
    k.iteration = a.out_ints[1];
    k.bailout = a.out_doubles[0];
    k.transition = a.out_doubles[1];
    k.phase = a.out_doubles[2];
    k.de_x = a.out_doubles[3];
    k.de_y = a.out_doubles[4];
    k.test1 = a.out_doubles[5];
    k.test2 = a.out_doubles[6];
    k.texture = nullptr;
    if (k.texture_active && texture) {
        const size_t bytes = static_cast<size_t>(texture->texture_stride) * static_cast<size_t>(texture->texture_height);
        void* t = s.kfp_planes.get(bytes + 256);
        check(hipMemcpyAsync(t, texture->texture_rgb, bytes, hipMemcpyHostToDevice, s.stream), "H2D texture");
        k.texture = static_cast<const uint8_t*>(t);
    }
    k.output = static_cast<uint8_t*>(s.kfp_rgb.get(pixels * 3));
    void* l = s.kfp_lut.get(static_cast<size_t>(colour.lut_size) * 3);
    check(hipMemcpyAsync(l, colour.lut, static_cast<size_t>(colour.lut_size) * 3, hipMemcpyHostToDevice, s.stream),
          "H2D lut");
    k.lut = static_cast<const uint8_t*>(l);
    void* scratch = s.kfp_scratch.get(64);
    const uint64_t wall0 = now_ns();
    check(hipEventRecord(s.start, s.stream), "event");
    check(launch_atlas_planes(a, s.stream), "atlas planes compose launch");
    check(launch_kfp_kernel(k, scratch, s.compute_units, s.stream), "KFP colour launch");
    check(hipEventRecord(s.stop, s.stream), "event");
    check(hipMemcpyAsync(output, k.output, pixels * 3, hipMemcpyDeviceToHost, s.stream), "D2H rgb");
    check(hipStreamSynchronize(s.stream), "atlas planes colour");
    float kernel_ms = 0.0f;
    check(hipEventElapsedTime(&kernel_ms, s.start, s.stop), "event time");
    s.timing.kernel_ms = kernel_ms;
    s.timing.total_ms = static_cast<double>(now_ns() - wall0) * 1e-6;
    s.timing.pixels = static_cast<double>(pixels);
}

void atlas_colourise_kfp_rect(const kfp::KfpParams& parent, const kfp::KfpParams& child, const kfp::RectAtlas& g,
                              const std::vector<float>& edge_x, const std::vector<float>& edge_y) {
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    auto align = [](size_t bytes) { return (bytes + 255) & ~size_t(255); };
    const size_t pn = static_cast<size_t>(parent.width) * static_cast<size_t>(parent.height);
    const size_t cn = child.field ? static_cast<size_t>(child.width) * static_cast<size_t>(child.height) : 0;
    const size_t lut_bytes = static_cast<size_t>(parent.lut_size) * 3;
    char* in = static_cast<char*>(s.atlas_tiles.get(align(pn * 4) + align(cn * 4) + align(edge_x.size() * 4)
                                                    + align(edge_y.size() * 4) + align(lut_bytes) + 256));
    size_t at = 0;
    auto put = [&](const void* src, size_t bytes) -> const void* {
        if (!src || !bytes) return nullptr;
        check(hipMemcpyAsync(in + at, src, bytes, hipMemcpyHostToDevice, s.stream), "H2D KFP atlas input");
        const void* d = in + at;
        at += align(bytes);
        return d;
    };
    kfp::KfpParams params[2] = {parent, child};
    params[0].field = static_cast<const float*>(put(parent.field, pn * 4));
    params[1].field = static_cast<const float*>(put(child.field, cn * 4));
    const uint8_t* lut = static_cast<const uint8_t*>(put(parent.lut, lut_bytes));
    params[0].lut = params[1].lut = lut;
    kfp::RectAtlas geometry = g;
    geometry.edge_x = static_cast<const float*>(put(edge_x.data(), edge_x.size() * 4));
    geometry.edge_y = static_cast<const float*>(put(edge_y.data(), edge_y.size() * 4));
    params[0].output = params[1].output = static_cast<uint8_t*>(s.kfp_rgb.get(pn * 3));
    auto* device_params = static_cast<kfp::KfpParams*>(s.kfp_planes.get(2 * sizeof(kfp::KfpParams) + 256));
    void* scratch = s.kfp_scratch.get(128);
    const uint64_t wall0 = now_ns();
    check(hipEventRecord(s.start, s.stream), "event");
    check(launch_kfp_rect_atlas(params, device_params, geometry, scratch, s.compute_units, s.stream),
          "KFP atlas launch");
    check(hipEventRecord(s.stop, s.stream), "event");
    check(hipMemcpyAsync(parent.output, params[0].output, pn * 3, hipMemcpyDeviceToHost, s.stream), "D2H rgb");
    check(hipStreamSynchronize(s.stream), "KFP atlas");
    float kernel_ms = 0.0f;
    check(hipEventElapsedTime(&kernel_ms, s.start, s.stop), "event time");
    s.timing.kernel_ms = kernel_ms;
    s.timing.total_ms = static_cast<double>(now_ns() - wall0) * 1e-6;
    s.timing.pixels = static_cast<double>(pn);
}

void aurora_run(const aurora::Params& host, const AuroraInputs& in) {
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    const size_t pixels = static_cast<size_t>(host.width) * static_cast<size_t>(host.height);
    aurora::Params p = host;
    // every input in one block, each piece 256-byte aligned
    struct Piece {
        const void* data;
        size_t bytes;
    };
    std::vector<Piece> pieces;
    auto add = [&](const void* data, size_t bytes) {
        pieces.push_back(Piece{data, bytes});
        return pieces.size() - 1;
    };
    const size_t parent_field = add(in.parent_field, in.parent_field_count * 4);
    const size_t child_field = add(in.child_field, in.child_field_count * 4);
    const size_t parent_rgb = add(in.parent_rgb, in.parent_rgb_bytes);
    const size_t child_rgb = add(in.child_rgb, in.child_rgb_bytes);
    size_t axis_piece[4][3];
    for (int a = 0; a < 4; ++a) {
        const atlas::Axis* axis = in.axes[a];
        const size_t n = axis ? axis->weight.size() : 0;
        axis_piece[a][0] = add(n ? axis->index0.data() : nullptr, n * 4);
        axis_piece[a][1] = add(n ? axis->index1.data() : nullptr, n * 4);
        axis_piece[a][2] = add(n ? axis->weight.data() : nullptr, n * 4);
    }
    const size_t edge_x = add(in.edge_x ? in.edge_x->data() : nullptr, in.edge_x ? in.edge_x->size() * 4 : 0);
    const size_t edge_y = add(in.edge_y ? in.edge_y->data() : nullptr, in.edge_y ? in.edge_y->size() * 4 : 0);
    const size_t palette = add(in.palette ? in.palette->data() : nullptr, in.palette ? in.palette->size() : 0);
    std::vector<size_t> at(pieces.size());
    size_t total = 0;
    for (size_t i = 0; i < pieces.size(); ++i) {
        at[i] = total;
        total += (pieces[i].bytes + 255) & ~size_t(255);
    }
    char* base = static_cast<char*>(s.aurora_in.get(total + 256));
    auto device = [&](size_t i) -> const void* { return pieces[i].bytes ? base + at[i] : nullptr; };
    for (size_t i = 0; i < pieces.size(); ++i)
        if (pieces[i].bytes)
            check(hipMemcpyAsync(base + at[i], pieces[i].data, pieces[i].bytes, hipMemcpyHostToDevice, s.stream),
                  "H2D compositor input");
    p.parent_field = static_cast<const float*>(device(parent_field));
    p.child_field = static_cast<const float*>(device(child_field));
    p.parent_rgb = static_cast<const uint8_t*>(device(parent_rgb));
    p.child_rgb = static_cast<const uint8_t*>(device(child_rgb));
    atlas::AxisDev* views[4] = {&p.parent_x, &p.parent_y, &p.child_x, &p.child_y};
    for (int a = 0; a < 4; ++a)
        *views[a] = atlas::AxisDev{static_cast<const int32_t*>(device(axis_piece[a][0])),
                                   static_cast<const int32_t*>(device(axis_piece[a][1])),
                                   static_cast<const float*>(device(axis_piece[a][2]))};
    p.edge_x = static_cast<const float*>(device(edge_x));
    p.edge_y = static_cast<const float*>(device(edge_y));
    p.palette = static_cast<const uint8_t*>(device(palette));
    const bool rgb_out = in.out_rgb != nullptr;
    void* out = s.aurora_out.get(pixels * (rgb_out ? 3 : 4));
    p.out_field = rgb_out ? nullptr : static_cast<float*>(out);
    p.out_rgb = rgb_out ? static_cast<uint8_t*>(out) : nullptr;
    if (p.mode == aurora::kAccents)  // in place on the caller's RGB
        check(hipMemcpyAsync(out, in.out_rgb, pixels * 3, hipMemcpyHostToDevice, s.stream), "H2D rgb");
    const uint64_t wall0 = now_ns();
    check(hipEventRecord(s.start, s.stream), "event");
    check(launch_aurora(p, s.stream), "compositor launch");
    check(hipEventRecord(s.stop, s.stream), "event");
    check(hipMemcpyAsync(rgb_out ? static_cast<void*>(in.out_rgb) : static_cast<void*>(in.out_field), out,
                         pixels * (rgb_out ? 3 : 4), hipMemcpyDeviceToHost, s.stream),
          "D2H compositor output");
    check(hipStreamSynchronize(s.stream), "compositor");
    float kernel_ms = 0.0f;
    check(hipEventElapsedTime(&kernel_ms, s.start, s.stop), "event time");
    s.timing.kernel_ms = kernel_ms;
    s.timing.total_ms = static_cast<double>(now_ns() - wall0) * 1e-6;
    s.timing.pixels = static_cast<double>(pixels);
}

void atlas_colourise_kfp_raw(const atlas::ComposeParams& host, size_t parent_count, size_t child_count,
                             const atlas::Axis* axes, const kfp::KfpParams& colour, uint8_t* output) {
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    const size_t pixels = static_cast<size_t>(host.width) * static_cast<size_t>(host.height);
    atlas::ComposeParams a = host;
    // tiles: parent then child, 256-byte aligned
    const size_t child_at = (parent_count * 4 + 255) & ~size_t(255);
    char* tiles = static_cast<char*>(s.atlas_tiles.get(child_at + child_count * 4 + 256));
    check(hipMemcpyAsync(tiles, host.parent, parent_count * 4, hipMemcpyHostToDevice, s.stream), "H2D parent");
    a.parent = reinterpret_cast<const float*>(tiles);
    a.child = nullptr;
    if (child_count) {
        check(hipMemcpyAsync(tiles + child_at, host.child, child_count * 4, hipMemcpyHostToDevice, s.stream),
              "H2D child");
        a.child = reinterpret_cast<const float*>(tiles + child_at);
    }
    // the four axes: index0, index1, weight of each, packed
    size_t bytes = 0;
    for (int i = 0; i < 4; ++i) bytes += ((axes[i].weight.size() * 4 + 255) & ~size_t(255)) * 3;
    char* ax = static_cast<char*>(s.atlas_axes.get(bytes + 256));
    atlas::AxisDev* views[4] = {&a.parent_x, &a.parent_y, &a.child_x, &a.child_y};
    size_t at = 0;
    for (int i = 0; i < 4; ++i) {
        const size_t n = axes[i].weight.size();
        const size_t stride = (n * 4 + 255) & ~size_t(255);
        *views[i] = atlas::AxisDev{nullptr, nullptr, nullptr};
        if (n == 0) continue;
        check(hipMemcpyAsync(ax + at, axes[i].index0.data(), n * 4, hipMemcpyHostToDevice, s.stream), "H2D axis");
        check(hipMemcpyAsync(ax + at + stride, axes[i].index1.data(), n * 4, hipMemcpyHostToDevice, s.stream),
              "H2D axis");
        check(hipMemcpyAsync(ax + at + 2 * stride, axes[i].weight.data(), n * 4, hipMemcpyHostToDevice, s.stream),
              "H2D axis");
        *views[i] = atlas::AxisDev{reinterpret_cast<const int32_t*>(ax + at),
                                   reinterpret_cast<const int32_t*>(ax + at + stride),
                                   reinterpret_cast<const float*>(ax + at + 2 * stride)};
        at += 3 * stride;
    }
    a.out = static_cast<float*>(s.atlas_field.get(pixels * sizeof(float)));
    kfp::KfpParams k = colour;
    k.field = a.out;
    k.output = static_cast<uint8_t*>(s.kfp_rgb.get(pixels * 3));
    void* l = s.kfp_lut.get(static_cast<size_t>(colour.lut_size) * 3);
    check(hipMemcpyAsync(l, colour.lut, static_cast<size_t>(colour.lut_size) * 3, hipMemcpyHostToDevice, s.stream),
          "H2D lut");
    k.lut = static_cast<const uint8_t*>(l);
    void* scratch = s.kfp_scratch.get(64);
    const uint64_t wall0 = now_ns();
    check(hipEventRecord(s.start, s.stream), "event");
    check(launch_atlas_compose(a, s.stream), "atlas compose launch");
    check(launch_kfp_kernel(k, scratch, s.compute_units, s.stream), "KFP colour launch");
    check(hipEventRecord(s.stop, s.stream), "event");
    check(hipMemcpyAsync(output, k.output, pixels * 3, hipMemcpyDeviceToHost, s.stream), "D2H rgb");
    check(hipStreamSynchronize(s.stream), "atlas colour");
    float kernel_ms = 0.0f;
    check(hipEventElapsedTime(&kernel_ms, s.start, s.stop), "event time");
    s.timing.kernel_ms = kernel_ms;
    s.timing.total_ms = static_cast<double>(now_ns() - wall0) * 1e-6;
    s.timing.pixels = static_cast<double>(pixels);
}

void render_field_floatexp(ReferenceContext& c, const FieldRequest& r, int32_t* counts, bool double_tail) {
    if (c.formula != 0) {
        render_alternate_floatexp(c, r, counts);
        return;
    }
    const int w = r.width, h = r.height;
    PreparedRender prep;
    prepare_render(c, r, prep);
    ensure_device_tables(c);
    const DeviceTables& t = *c.device;
    RenderParams p = prep.params;
    p.orbit = t.orbit;
    const SeriesDev scalar_series = p.series;
    p.series = t.series;
    p.series.count = scalar_series.count;
    p.series.iteration = scalar_series.iteration;
    p.series.available = scalar_series.available;
    p.series.radius2_m = scalar_series.radius2_m;
    p.series.radius2_e = scalar_series.radius2_e;
    p.linear = t.linear[prep.linear_tier];
    p.linear_dir = t.linear_dir[prep.linear_tier];
    p.budget_ticks = 0;  // no time budget on this path

    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    const size_t pixels = static_cast<size_t>(w) * static_cast<size_t>(h);
    const std::vector<double>& off_m = prep.offset_m;
    const std::vector<int32_t>& off_e = prep.offset_e;
    char* doff = static_cast<char*>(s.offsets.get(off_m.size() * 12 + 512));
    const size_t exp_offset = (off_m.size() * 8 + 255) & ~size_t(255);
    check(hipMemcpyAsync(doff, off_m.data(), off_m.size() * 8, hipMemcpyHostToDevice, s.stream), "offsets");
    check(hipMemcpyAsync(doff + exp_offset, off_e.data(), off_e.size() * 4, hipMemcpyHostToDevice, s.stream),
          "offsets");
    p.x_off_m = reinterpret_cast<const double*>(doff);
    p.y_off_m = p.x_off_m + w;
    p.x_off_e = reinterpret_cast<const int32_t*>(doff + exp_offset);
    p.y_off_e = p.x_off_e + w;
    fx::FxOut out{};
    out.field = r.device_pointers ? r.output : static_cast<float*>(s.output.get(pixels * sizeof(float)));
    out.counts = r.device_pointers ? counts : static_cast<int32_t*>(s.fx_counts.get(pixels * sizeof(int32_t)));
    out.double_tail = double_tail ? 1 : 0;
    const uint64_t wall0 = now_ns();
    check(hipEventRecord(s.start, s.stream), "event");
    check(launch_fx_kernel(p, out, s.stream), "floatexp kernel launch");
    check(hipEventRecord(s.stop, s.stream), "event");
    if (!r.device_pointers) {
        check(hipMemcpyAsync(r.output, out.field, pixels * sizeof(float), hipMemcpyDeviceToHost, s.stream), "D2H");
        check(hipMemcpyAsync(counts, out.counts, pixels * sizeof(int32_t), hipMemcpyDeviceToHost, s.stream), "D2H");
    }
    check(hipStreamSynchronize(s.stream), "floatexp render");
    float kernel_ms = 0.0f;
    check(hipEventElapsedTime(&kernel_ms, s.start, s.stop), "event time");
    s.timing.kernel_ms = kernel_ms;
    s.timing.total_ms = static_cast<double>(now_ns() - wall0) * 1e-6;
    s.timing.pixels = static_cast<double>(pixels);
}

void check_tile_order(const uint32_t* tile_order, size_t tile_count, int width, int height) {
    const size_t tiles = static_cast<size_t>((width + 31) / 32) * static_cast<size_t>((height + 31) / 32);
    if (tile_count != tiles) throw std::runtime_error("tile order must list every 32x32 tile once");
    std::vector<uint32_t> sorted(tile_order, tile_order + tile_count);
    std::sort(sorted.begin(), sorted.end());
    for (size_t i = 0; i < tiles; ++i)
        if (sorted[i] != i) throw std::runtime_error("tile order must list every 32x32 tile once");
}

namespace {

// Launch parameters of one alternate-formula render with the reference's
// device tables bound and the per-column / per-row offsets uploaded to
// `offsets` on the library stream (lock held; `prep` must stay alive until
// the stream has run the copy).
AltRenderParams bind_alternate(const ReferenceContext& c, const PreparedAlternate& prep, LaunchState& s,
                               DeviceBuffer& offsets) {
    const DeviceTables& t = *c.device;
    AltRenderParams p = prep.params;
    p.re = t.orbit.re;
    p.im = t.orbit.im;
    p.e = t.orbit.e;
    p.margin_m = t.alt_margin_m;
    p.margin_e = t.alt_margin_e;
    p.steps = t.alt;
    p.dir = t.alt_dir;
    const SeriesDev scalar_series = p.series;
    p.series = t.series;
    p.series.count = scalar_series.count;
    p.series.iteration = scalar_series.iteration;
    p.series.radius2_m = scalar_series.radius2_m;
    p.series.radius2_e = scalar_series.radius2_e;
    const std::vector<double>& off_m = prep.offset_m;
    const std::vector<int32_t>& off_e = prep.offset_e;
    char* doff = static_cast<char*>(offsets.get(off_m.size() * 12 + 512));
    const size_t exp_offset = (off_m.size() * 8 + 255) & ~size_t(255);
    check(hipMemcpyAsync(doff, off_m.data(), off_m.size() * 8, hipMemcpyHostToDevice, s.stream), "offsets");
    check(hipMemcpyAsync(doff + exp_offset, off_e.data(), off_e.size() * 4, hipMemcpyHostToDevice, s.stream),
          "offsets");
    p.x_off_m = reinterpret_cast<const double*>(doff);
    p.y_off_m = p.x_off_m + p.width;
    p.x_off_e = reinterpret_cast<const int32_t*>(doff + exp_offset);
    p.y_off_e = p.x_off_e + p.width;
    return p;
}

void bind_budget(AltRenderParams& p, LaunchState& s, int time_budget_ms) {
    void* budget = s.budget.get(256);
    p.budget_start = static_cast<const unsigned long long*>(budget);
    p.budget_flag = reinterpret_cast<int*>(static_cast<char*>(budget) + 8);
    p.budget_ticks = time_budget_ms > 0 ? static_cast<unsigned long long>(time_budget_ms * s.wall_clock_khz) : 0ULL;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, a, b), "event time");
    return ms;
}

double log10_zoom(const ReferenceContext& c, const char* zoom_text) {
    const FE z = parse_zoom(zoom_text, c.precision_bits);
    return (std::log(z.m) + static_cast<double>(z.e) * 0.69314718055994530942) / 2.30258509299404568402;
}

FieldRequest preview_request(const FieldRequest& r) {
    FieldRequest q = r;
    q.width = (r.width + 3) / 4;
    q.height = (r.height + 3) / 4;
    q.output = nullptr;
    q.planes = nullptr;
    q.stats = nullptr;
    q.device_pointers = true;
    return q;
}

// Preview render (quarter size, per-pixel pass budget) and the tile
// features of the full field `r` into s.mlp_features; lock held.  Returns
// the device features pointer.
float* launch_preview_features(const ReferenceContext& c, const FieldRequest& r, const PreparedAlternate& prep,
                               int work_limit, LaunchState& s, hipEvent_t before, hipEvent_t after) {
    const FieldRequest q = preview_request(r);
    AltRenderParams p = bind_alternate(c, prep, s, s.preview_offsets);
    const size_t pixels = static_cast<size_t>(q.width) * static_cast<size_t>(q.height);
    p.output = static_cast<float*>(s.preview_output.get(pixels * sizeof(float)));
    p.work = static_cast<int32_t*>(s.preview_work.get(pixels * 4));
    p.work_limit = work_limit;
    bind_budget(p, s, 0);
    const int tiles_x = (r.width + mlp::kTile - 1) / mlp::kTile, tiles_y = (r.height + mlp::kTile - 1) / mlp::kTile;
    float* features = static_cast<float*>(
        s.mlp_features.get(static_cast<size_t>(tiles_x) * tiles_y * mlp::kInputs * sizeof(float)));
    check(hipEventRecord(before, s.stream), "event");
    check(launch_alt_kernel(p, s.stream), "preview kernel launch");
    check(mlp::features(p.work, q.width, q.height, tiles_x, tiles_y,
                        static_cast<float>(log10_zoom(c, r.zoom_text) / 100.0), features, s.stream),
          "tile features");
    check(hipEventRecord(after, s.stream), "event");
    return features;
}

// The cost model over the device features of `tiles` tiles with the host
// `weights`, and the tiles sorted by decreasing prediction (equal or NaN
// predictions keep row-major order); lock held.  s.mark[2] is recorded
// after the model.
std::vector<uint32_t> heaviest_first(LaunchState& s, const float* features, int tiles, const float* weights) {
    float* w = static_cast<float*>(s.mlp_weights.get(mlp::kParams * sizeof(float)));
    float* y = static_cast<float*>(s.mlp_y.get(static_cast<size_t>(tiles) * sizeof(float)));
    check(hipMemcpyAsync(w, weights, mlp::kParams * sizeof(float), hipMemcpyHostToDevice, s.stream), "H2D weights");
    check(mlp::forward(w, features, tiles, y, nullptr, s.stream), "cost model forward");
    check(hipEventRecord(s.mark[2], s.stream), "event");
    std::vector<float> predicted(static_cast<size_t>(tiles));
    check(hipMemcpyAsync(predicted.data(), y, predicted.size() * sizeof(float), hipMemcpyDeviceToHost, s.stream),
          "D2H costs");
    check(hipStreamSynchronize(s.stream), "cost model");
    std::vector<uint32_t> order(static_cast<size_t>(tiles));
    for (int i = 0; i < tiles; ++i) order[static_cast<size_t>(i)] = static_cast<uint32_t>(i);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const float ca = predicted[a], cb = predicted[b];
        return (ca == ca ? ca : -INFINITY) > (cb == cb ? cb : -INFINITY);
    });
    return order;
}

// Quarter-size preview of Mandelbrot field `r` and the tile features of the
// full field into s.mlp_features (lock held).  The preview `q` runs with cap
// min(max_iter, work_limit): every loop pass advances at least one
// iteration, so no pixel takes more than work_limit passes.  Its work map
// (loop passes per pixel, from the statistics kernels) feeds the features
// exactly as an alternate-formula preview's does.  `p` is bound to the
// reference's tables for `prep` (the preview's per-frame values); `points`
// must outlive the stream work.
float* launch_field_preview(const FieldRequest& r, const FieldRequest& q, const PreparedRender& prep, RenderParams p,
                            double log10z, LaunchState& s, PointStaging& points, hipEvent_t before,
                            hipEvent_t after) {
    bind_field_frame(p, prep, q, s, s.preview_offsets, s.preview_state, points);
    const size_t pixels = static_cast<size_t>(q.width) * static_cast<size_t>(q.height);
    p.output = static_cast<float*>(s.preview_output.get(pixels * sizeof(float)));
    p.stats = static_cast<unsigned long long*>(s.preview_stats.get(kStatsFields * 8));
    p.work = static_cast<int32_t*>(s.preview_work.get(pixels * 4));
    const int tiles_x = (r.width + mlp::kTile - 1) / mlp::kTile, tiles_y = (r.height + mlp::kTile - 1) / mlp::kTile;
    float* features = static_cast<float*>(
        s.mlp_features.get(static_cast<size_t>(tiles_x) * tiles_y * mlp::kInputs * sizeof(float)));
    check(hipEventRecord(before, s.stream), "event");
    check(hipMemsetAsync(p.stats, 0, kStatsFields * 8, s.stream), "stats clear");
    check(hipMemsetAsync(p.work, 0, pixels * 4, s.stream), "work clear");
    check(launch_field_kernel(p, false, true, s.compute_units * 8, s.stream, s.side, s.fork, s.join),
          "preview kernel launch");
    check(mlp::features(p.work, q.width, q.height, tiles_x, tiles_y, static_cast<float>(log10z / 100.0), features,
                        s.stream),
          "tile features");
    check(hipEventRecord(after, s.stream), "event");
    return features;
}

FieldRequest field_preview_request(const FieldRequest& r, int work_limit) {
    FieldRequest q = preview_request(r);
    q.max_iter = std::min(r.max_iter, work_limit);
    q.work = nullptr;
    q.tile_order = nullptr;
    q.tile_count = 0;
    return q;
}

}  // namespace

void render_alternate(ReferenceContext& c, const FieldRequest& r, const uint32_t* tile_order, size_t tile_count,
                      int32_t* work, int32_t* counts) {
    PreparedAlternate prep;
    prepare_alternate(c, r, prep);
    if (tile_order) check_tile_order(tile_order, tile_count, r.width, r.height);
    ensure_device_tables(c);
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    const size_t pixels = static_cast<size_t>(r.width) * static_cast<size_t>(r.height);
    AltRenderParams p = bind_alternate(c, prep, s, s.offsets);
    p.output = r.device_pointers ? r.output : static_cast<float*>(s.output.get(pixels * sizeof(float)));
    bind_budget(p, s, r.options.time_budget_ms);
    if (tile_order) {
        void* d = s.tile_order.get(tile_count * 4);
        check(hipMemcpyAsync(d, tile_order, tile_count * 4, hipMemcpyHostToDevice, s.stream), "H2D tile order");
        p.tile_order = static_cast<const uint32_t*>(d);
        p.tile_count = static_cast<int32_t>(tile_count);
    }
    if (work) p.work = static_cast<int32_t*>(s.work.get(pixels * 4));
    if (counts) p.counts = static_cast<int32_t*>(s.fx_counts.get(pixels * 4));
    PointStaging points;
    size_t out_count = pixels;
    if (r.points) {
        upload_points(s, r, points, p.point_re, p.point_im, p.point_e);
        p.point_count = r.point_count;
        out_count = static_cast<size_t>(r.point_count);
    }
    const uint64_t wall0 = now_ns();
    check(hipEventRecord(s.start, s.stream), "event");
    check(launch_alt_kernel(p, s.stream), "alternate kernel launch");
    check(hipEventRecord(s.stop, s.stream), "event");
    if (!r.device_pointers)
        check(hipMemcpyAsync(r.output, p.output, out_count * sizeof(float), hipMemcpyDeviceToHost, s.stream),
              "D2H");
    if (work) check(hipMemcpyAsync(work, p.work, pixels * 4, hipMemcpyDeviceToHost, s.stream), "D2H work");
    if (counts) check(hipMemcpyAsync(counts, p.counts, pixels * 4, hipMemcpyDeviceToHost, s.stream), "D2H counts");
    check(hipStreamSynchronize(s.stream), "alternate render");
    const float kernel_ms = elapsed(s.start, s.stop);
    s.timing.kernel_ms = kernel_ms;
    s.timing.total_ms = static_cast<double>(now_ns() - wall0) * 1e-6;
    s.timing.pixels = static_cast<double>(pixels);
    if (r.stats) {  // the alternate renderer reports pixels and time only (renderer.cpp:12439-12446)
        std::fill(r.stats, r.stats + kStatsFields, 0ULL);
        r.stats[0] = static_cast<uint64_t>(out_count);
        r.stats[22] = static_cast<uint64_t>(static_cast<double>(kernel_ms) * 1e6);
    }
}

void render_alternate_floatexp(ReferenceContext& c, const FieldRequest& r, int32_t* counts) {
    PreparedAlternate prep;
    prepare_alternate(c, r, prep);
    ensure_device_tables(c);
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    const size_t pixels = static_cast<size_t>(r.width) * static_cast<size_t>(r.height);
    AltRenderParams p = bind_alternate(c, prep, s, s.offsets);
    p.budget_ticks = 0;  // no time budget on this path
    fx::FxOut out{};
    out.field = r.device_pointers ? r.output : static_cast<float*>(s.output.get(pixels * sizeof(float)));
    out.counts = r.device_pointers ? counts : static_cast<int32_t*>(s.fx_counts.get(pixels * sizeof(int32_t)));
    const uint64_t wall0 = now_ns();
    check(hipEventRecord(s.start, s.stream), "event");
    check(launch_fx_alt_kernel(p, out, s.stream), "floatexp kernel launch");
    check(hipEventRecord(s.stop, s.stream), "event");
    if (!r.device_pointers) {
        check(hipMemcpyAsync(r.output, out.field, pixels * sizeof(float), hipMemcpyDeviceToHost, s.stream), "D2H");
        check(hipMemcpyAsync(counts, out.counts, pixels * sizeof(int32_t), hipMemcpyDeviceToHost, s.stream), "D2H");
    }
    check(hipStreamSynchronize(s.stream), "floatexp render");
    s.timing.kernel_ms = elapsed(s.start, s.stop);
    s.timing.total_ms = static_cast<double>(now_ns() - wall0) * 1e-6;
    s.timing.pixels = static_cast<double>(pixels);
}

void alternate_preview_features(ReferenceContext& c, const FieldRequest& r, int work_limit, float* features) {
    PreparedAlternate prep;
    prepare_alternate(c, preview_request(r), prep);
    ensure_device_tables(c);
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    const size_t count = static_cast<size_t>((r.width + 31) / 32) * ((r.height + 31) / 32) * mlp::kInputs;
    const float* d = launch_preview_features(c, r, prep, work_limit, s, s.start, s.stop);
    check(hipMemcpyAsync(features, d, count * sizeof(float), hipMemcpyDeviceToHost, s.stream), "D2H features");
    check(hipStreamSynchronize(s.stream), "preview");
    const float ms = elapsed(s.start, s.stop);
    s.timing = LastTiming{ms, ms, static_cast<double>(count / mlp::kInputs)};
}

void render_alternate_learned(ReferenceContext& c, const FieldRequest& r, const float* weights, int work_limit,
                              uint32_t* order_out, double* timings) {
    PreparedAlternate preview_prep, prep;
    prepare_alternate(c, preview_request(r), preview_prep);
    prepare_alternate(c, r, prep);
    ensure_device_tables(c);
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    const uint64_t wall0 = now_ns();
    const int tiles = ((r.width + 31) / 32) * ((r.height + 31) / 32);
    // preview, features, cost model
    const float* features = launch_preview_features(c, r, preview_prep, work_limit, s, s.mark[0], s.mark[1]);
    const std::vector<uint32_t> order = heaviest_first(s, features, tiles, weights);
    if (order_out) std::copy(order.begin(), order.end(), order_out);
    const uint64_t sorted_ns = now_ns();
    // the field in that order
    const size_t pixels = static_cast<size_t>(r.width) * static_cast<size_t>(r.height);
    AltRenderParams p = bind_alternate(c, prep, s, s.offsets);
    p.output = r.device_pointers ? r.output : static_cast<float*>(s.output.get(pixels * sizeof(float)));
    bind_budget(p, s, r.options.time_budget_ms);
    void* d_order = s.tile_order.get(order.size() * 4);
    check(hipMemcpyAsync(d_order, order.data(), order.size() * 4, hipMemcpyHostToDevice, s.stream), "H2D order");
    p.tile_order = static_cast<const uint32_t*>(d_order);
    p.tile_count = tiles;
    check(hipEventRecord(s.start, s.stream), "event");
    check(launch_alt_kernel(p, s.stream), "alternate kernel launch");
    check(hipEventRecord(s.stop, s.stream), "event");
    if (!r.device_pointers)
        check(hipMemcpyAsync(r.output, p.output, pixels * sizeof(float), hipMemcpyDeviceToHost, s.stream), "D2H");
    check(hipStreamSynchronize(s.stream), "learned render");
    const float render_ms = elapsed(s.start, s.stop);
    s.timing.kernel_ms = render_ms;
    s.timing.total_ms = static_cast<double>(now_ns() - wall0) * 1e-6;
    s.timing.pixels = static_cast<double>(pixels);
    if (timings) {
        timings[0] = elapsed(s.mark[0], s.mark[1]);  // preview render + features
        timings[1] = elapsed(s.mark[1], s.mark[2]);  // weights upload + cost model
        timings[2] = render_ms;                      // ordered render
        timings[3] = static_cast<double>(sorted_ns - wall0) * 1e-6;  // wall time up to the ordered launch
        timings[4] = s.timing.total_ms;
    }
}

void field_preview_features(ReferenceContext& c, const FieldRequest& r, int work_limit, float* features) {
    const FieldRequest q = field_preview_request(r, work_limit);
    PreparedRender prep;
    prepare_render(c, q, prep);
    const RenderParams p = bind_field_tables(c, prep);
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    const size_t count = static_cast<size_t>((r.width + 31) / 32) * ((r.height + 31) / 32) * mlp::kInputs;
    PointStaging points;
    const float* d = launch_field_preview(r, q, prep, p, log10_zoom(c, r.zoom_text), s, points, s.start, s.stop);
    check(hipMemcpyAsync(features, d, count * sizeof(float), hipMemcpyDeviceToHost, s.stream), "D2H features");
    check(hipStreamSynchronize(s.stream), "preview");
    const float ms = elapsed(s.start, s.stop);
    s.timing = LastTiming{ms, ms, static_cast<double>(count / mlp::kInputs)};
}

void render_field_learned(ReferenceContext& c, const FieldRequest& r, const float* weights, int work_limit,
                          uint32_t* order_out, double* timings) {
    const FieldRequest q = field_preview_request(r, work_limit);
    PreparedRender preview_prep, prep;
    prepare_render(c, q, preview_prep);
    prepare_render(c, r, prep);
    const RenderParams preview_params = bind_field_tables(c, preview_prep);
    const RenderParams params = bind_field_tables(c, prep);
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    const uint64_t wall0 = now_ns();
    const int tiles = ((r.width + 31) / 32) * ((r.height + 31) / 32);
    PointStaging points;
    const float* features = launch_field_preview(r, q, preview_prep, preview_params, log10_zoom(c, r.zoom_text), s,
                                                 points, s.mark[0], s.mark[1]);
    const std::vector<uint32_t> order = heaviest_first(s, features, tiles, weights);
    if (order_out) std::copy(order.begin(), order.end(), order_out);
    const uint64_t sorted_ns = now_ns();
    FieldRequest ordered = r;
    ordered.tile_order = order.data();
    ordered.tile_count = order.size();
    render_field_locked(ordered, prep, params, s);
    if (timings) {
        timings[0] = elapsed(s.mark[0], s.mark[1]);  // preview render + features
        timings[1] = elapsed(s.mark[1], s.mark[2]);  // weights upload + cost model
        timings[2] = s.timing.kernel_ms;             // ordered render
        timings[3] = static_cast<double>(sorted_ns - wall0) * 1e-6;
        timings[4] = static_cast<double>(now_ns() - wall0) * 1e-6;
    }
    s.timing.total_ms = static_cast<double>(now_ns() - wall0) * 1e-6;
}

// ---- learned tile-cost model -------------------------------------------------

namespace {

mlp::Workspace mlp_workspace(LaunchState& s, int n) {
    const size_t count = static_cast<size_t>(n);
    mlp::Workspace ws{};
    ws.y = static_cast<float*>(s.mlp_y.get(count * 4));
    ws.act = static_cast<float*>(s.mlp_act.get(count * mlp::kActivations * 4));
    ws.delta = static_cast<float*>(s.mlp_delta.get(count * mlp::kDeltas * 4));
    ws.err2 = static_cast<float*>(s.mlp_err2.get(count * 4));
    ws.partial = static_cast<float*>(
        s.mlp_partial.get(static_cast<size_t>(mlp::chunk_count(n)) * mlp::kParams * 4));
    return ws;
}

}  // namespace

void mlp_forward(const float* weights, const float* x, int n, float* y) {
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    check(hipEventRecord(s.start, s.stream), "event");
    check(mlp::forward(weights, x, n, y, nullptr, s.stream), "cost model forward");
    check(hipEventRecord(s.stop, s.stream), "event");
    check(hipStreamSynchronize(s.stream), "cost model forward");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, s.start, s.stop), "event time");
    s.timing = LastTiming{ms, ms, static_cast<double>(n)};
}

void mlp_gradient(const float* weights, const float* x, const float* t, int n, float* grad, float* loss) {
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    const mlp::Workspace ws = mlp_workspace(s, n);
    check(mlp::gradient(weights, x, t, n, ws, grad, loss, s.stream), "cost model gradient");
    check(hipStreamSynchronize(s.stream), "cost model gradient");
}

void mlp_train(float* weights, float* adam_m, float* adam_v, const float* x, const float* t, int n, int steps,
               float learning_rate, int first_step, float* loss_history) {
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    const mlp::Workspace ws = mlp_workspace(s, n);
    float* grad = static_cast<float*>(s.mlp_grad.get(mlp::kParams * 4));
    check(hipEventRecord(s.start, s.stream), "event");
    for (int k = 0; k < steps; ++k) {
        check(mlp::gradient(weights, x, t, n, ws, grad, loss_history + k, s.stream), "cost model gradient");
        check(mlp::adam(weights, grad, adam_m, adam_v, learning_rate, first_step + k, s.stream), "cost model step");
    }
    check(hipEventRecord(s.stop, s.stream), "event");
    check(hipStreamSynchronize(s.stream), "cost model training");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, s.start, s.stop), "event time");
    s.timing = LastTiming{ms, ms, static_cast<double>(n) * steps};
}

// ---- batched fields --------------------------------------------------------

class PreparedBatch {
public:
    std::vector<BatchFieldRequest> fields;
    std::vector<RenderParams> params;      // host image of q.fields
    std::vector<uint32_t> field_major;     // tile codes in field-major order
    std::vector<uint32_t> order;           // tile codes in launch order
    DeviceBuffer params_d, tiles_d, offsets_d, state_d, queues_d, counters_d, direct_d, coords_d;
    BatchLaunch q{};
    int fast_blocks_per_cu = 5;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_blocks = 0;
    ~PreparedBatch() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (DeviceBuffer* b : {&params_d, &tiles_d, &offsets_d, &state_d, &queues_d, &counters_d, &direct_d, &coords_d})
            if (b->ptr) (void)hipFree(b->ptr);
    }
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
    }
};

std::shared_ptr<PreparedBatch> prepare_batch(std::vector<BatchFieldRequest> fields) {
    if (fields.empty()) throw std::runtime_error("empty batch");
    if (fields.size() >= (1U << 12)) throw std::runtime_error("too many fields in one batch");
    auto b = std::make_shared<PreparedBatch>();
    std::vector<PreparedRender> preps(fields.size());
    std::vector<direct::DirectSetup> directs(fields.size());
    std::vector<const RungDev*> rungs(fields.size(), nullptr);
    std::vector<const RefWalkDev*> walks(fields.size(), nullptr);
    std::vector<int> walk_limits(fields.size(), 0);
    size_t pixels_total = 0, offsets_total = 0, state_total = 0, coords_total = 0;
    bool any_direct = false;
    for (size_t f = 0; f < fields.size(); ++f) {
        const FieldRequest& r = fields[f].request;
        if (!r.device_pointers || r.planes || r.stats) throw std::runtime_error("batch fields render to device pointers only");
        if (fields[f].context->formula != 0) throw std::runtime_error("batched fields use the Mandelbrot formula");
        const size_t tiles = static_cast<size_t>((r.width + 15) / 16) * static_cast<size_t>((r.height + 15) / 16);
        if (tiles >= (1U << 20)) throw std::runtime_error("batch field too large");
        const size_t pixels = static_cast<size_t>(r.width) * static_cast<size_t>(r.height);
        pixels_total += pixels;
        if (fields[f].direct) {
            const ReferenceContext& c = *fields[f].context;
            const FractalRenderOptions& o = r.options;
            directs[f] = direct::make_setup_values(r.width, r.height, fields[f].direct_zoom, c.x_center, c.y_center,
                                                   r.max_iter, 0, 0.0, 0.0, o.output_bias, o.escape_radius_mode,
                                                   o.coordinate_mode, static_cast<uint32_t>(o.reserved[0]));
            coords_total += (static_cast<size_t>(r.width + r.height) * 8 + 255) & ~size_t(255);
            any_direct = true;
        } else {
            prepare_render(*fields[f].context, r, preps[f]);
            ensure_device_tables(*fields[f].context);
            const int cap = 31 - __builtin_clz(static_cast<unsigned>(preps[f].params.max_linear_length));
            ensure_plain_tables(*fields[f].context, preps[f].linear_tier, cap, preps[f].params.approximation_order >= 3,
                                rungs[f]);
            {
                // the on-reference walk, as in render_field
                const RenderParams& q = preps[f].params;
                const bool linear = q.approximation_order <= 1;
                const int level =
                    31 - __builtin_clz(static_cast<unsigned>(linear ? q.max_linear_length : q.max_bla_length));
                if (level < kMaxLevels) {
                    ensure_ref_walk(*fields[f].context, linear ? preps[f].linear_tier : 2, level, walks[f]);
                    walk_limits[f] = 1 << level;
                }
            }
            offsets_total += (preps[f].offset_m.size() * 12 + 512 + 255) & ~size_t(255);
            state_total += (pixels * kStateBytesPerPixel + 20 * 256 + 255) & ~size_t(255);
        }
        for (uint32_t t = 0; t < tiles; ++t) b->field_major.push_back((static_cast<uint32_t>(f) << 20) | t);
    }
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    char* off_base = static_cast<char*>(b->offsets_d.get(offsets_total));
    char* state_base = static_cast<char*>(b->state_d.get(state_total));
    size_t off_at = 0, state_at = 0, coords_at = 0;
    b->params.assign(fields.size(), RenderParams{});
    std::vector<direct::DirectParams> direct_params(any_direct ? fields.size() : 0);  // width 0: not direct
    char* coords_base = any_direct ? static_cast<char*>(b->coords_d.get(coords_total)) : nullptr;
    for (size_t f = 0; f < fields.size(); ++f) {
        const FieldRequest& r = fields[f].request;
        if (fields[f].direct) {
            direct::DirectParams d = directs[f].params;
            double* coords = reinterpret_cast<double*>(coords_base + coords_at);
            coords_at += (static_cast<size_t>(r.width + r.height) * 8 + 255) & ~size_t(255);
            check(hipMemcpyAsync(coords, directs[f].x.data(), static_cast<size_t>(r.width) * 8, hipMemcpyHostToDevice,
                                 s.stream), "batch direct x");
            check(hipMemcpyAsync(coords + r.width, directs[f].y.data(), static_cast<size_t>(r.height) * 8,
                                 hipMemcpyHostToDevice, s.stream), "batch direct y");
            d.x = coords;
            d.y = coords + r.width;
            d.output = r.output;
            direct_params[f] = d;
            b->params[f].width = r.width;
            b->params[f].height = r.height;
            continue;
        }
        const DeviceTables& t = *fields[f].context->device;
        const PreparedRender& prep = preps[f];
        RenderParams p = prep.params;
        p.orbit = t.orbit;
        const SeriesDev scalar_series = p.series;
        p.series = t.series;
        p.series.count = scalar_series.count;
        p.series.iteration = scalar_series.iteration;
        p.series.available = scalar_series.available;
        p.series.radius2_m = scalar_series.radius2_m;
        p.series.radius2_e = scalar_series.radius2_e;
        p.linear = t.linear[prep.linear_tier];
        p.linear_dir = t.linear_dir[prep.linear_tier];
        p.cubic = t.cubic;
        p.cubic_dir = t.cubic_dir;
        // the plain pass as in render_field; the pixel on the reference
        // (reference_pixel -1) is left to batch_fast_kernel
        p.rungs = rungs[f];
        p.plain_top = 31 - __builtin_clz(static_cast<unsigned>(p.max_linear_length));
        p.plain_pass = p.rungs != nullptr && p.fast_linear ? 1 : 0;
        p.reference_pixel = -1;
        p.ref_walk = walks[f];
        p.ref_walk_limit = walk_limits[f];
        char* doff = off_base + off_at;
        const size_t exp_offset = (prep.offset_m.size() * 8 + 255) & ~size_t(255);
        check(hipMemcpyAsync(doff, prep.offset_m.data(), prep.offset_m.size() * 8, hipMemcpyHostToDevice, s.stream),
              "batch offsets");
        check(hipMemcpyAsync(doff + exp_offset, prep.offset_e.data(), prep.offset_e.size() * 4,
                             hipMemcpyHostToDevice, s.stream), "batch offsets");
        off_at += (prep.offset_m.size() * 12 + 512 + 255) & ~size_t(255);
        p.x_off_m = reinterpret_cast<const double*>(doff);
        p.y_off_m = p.x_off_m + r.width;
        p.x_off_e = reinterpret_cast<const int32_t*>(doff + exp_offset);
        p.y_off_e = p.x_off_e + r.width;
        p.output = r.output;
        const size_t pixels = static_cast<size_t>(r.width) * static_cast<size_t>(r.height);
        bind_state_store(p, state_base + state_at, pixels);
        state_at += (pixels * kStateBytesPerPixel + 20 * 256 + 255) & ~size_t(255);
        p.budget_ticks = 0;  // no time budget in a batch
        b->params[f] = p;
    }
    void* params_d = b->params_d.get(b->params.size() * sizeof(RenderParams));
    check(hipMemcpyAsync(params_d, b->params.data(), b->params.size() * sizeof(RenderParams), hipMemcpyHostToDevice,
                         s.stream), "batch params");
    void* queues = b->queues_d.get(pixels_total * 16);
    void* counters = b->counters_d.get(256);
    b->q.direct_fields = nullptr;
    if (any_direct) {
        void* direct_d = b->direct_d.get(direct_params.size() * sizeof(direct::DirectParams));
        check(hipMemcpyAsync(direct_d, direct_params.data(), direct_params.size() * sizeof(direct::DirectParams),
                             hipMemcpyHostToDevice, s.stream), "batch direct params");
        b->q.direct_fields = static_cast<const direct::DirectParams*>(direct_d);
    }
    b->q.fields = static_cast<const RenderParams*>(params_d);
    b->q.field_count = static_cast<unsigned>(fields.size());
    b->q.general_queue = static_cast<unsigned long long*>(queues);
    b->q.tail_queue = static_cast<unsigned long long*>(queues) + pixels_total;
    b->q.counters = static_cast<unsigned*>(counters);
    b->fields = std::move(fields);
    b->order = b->field_major;
    void* tiles = b->tiles_d.get(b->order.size() * 4);
    check(hipMemcpyAsync(tiles, b->order.data(), b->order.size() * 4, hipMemcpyHostToDevice, s.stream), "tiles");
    b->q.tiles = static_cast<const uint32_t*>(tiles);
    b->q.tile_count = static_cast<unsigned>(b->order.size());
    check(hipStreamSynchronize(s.stream), "batch upload");
    return b;
}

void batch_set_order(PreparedBatch& b, int order, const float* costs, size_t cost_count) {
    std::vector<uint32_t> next;
    next.reserve(b.field_major.size());
    if (order == 0) {
        next = b.field_major;
    } else if (order == 1) {  // round-robin over fields, one tile each
        std::vector<std::vector<uint32_t>> per(b.fields.size());
        for (uint32_t code : b.field_major) per[code >> 20].push_back(code);
        for (size_t i = 0;; ++i) {
            bool any = false;
            for (auto& v : per)
                if (i < v.size()) {
                    next.push_back(v[i]);
                    any = true;
                }
            if (!any) break;
        }
    } else if (order == 2) {
        if (!costs || cost_count != b.field_major.size())
            throw std::runtime_error("cost order needs one cost per tile");
        std::vector<size_t> idx(b.field_major.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) { return costs[x] > costs[y]; });
        for (size_t i : idx) next.push_back(b.field_major[i]);
    } else {
        throw std::runtime_error("unknown batch tile order");
    }
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    b.order = std::move(next);
    check(hipMemcpyAsync(b.tiles_d.ptr, b.order.data(), b.order.size() * 4, hipMemcpyHostToDevice, s.stream),
          "tiles");
    check(hipStreamSynchronize(s.stream), "tiles");
}

size_t batch_tile_count(const PreparedBatch& b) { return b.field_major.size(); }

void batch_set_floatexp(PreparedBatch& b, bool floatexp) {
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    if ((b.q.floatexp != 0) == floatexp) return;
    b.q.floatexp = floatexp ? 1 : 0;
    b.drop_graph();  // the captured launch sequence differs
}

void batch_render(PreparedBatch& b, int fast_blocks_per_cu, bool use_graph) {
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    if (fast_blocks_per_cu > 0) b.fast_blocks_per_cu = fast_blocks_per_cu;
    const int fast_blocks = s.compute_units * b.fast_blocks_per_cu;
    const int persistent_blocks = s.compute_units * 8;
    const uint64_t wall0 = now_ns();
    if (use_graph) {
        if (!b.graph_exec || b.graph_blocks != fast_blocks) {
            b.drop_graph();
            check(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal), "graph capture");
            const hipError_t launched = launch_batch(b.q, fast_blocks, persistent_blocks, s.stream);
            const hipError_t ended = hipStreamEndCapture(s.stream, &b.graph);
            check(launched, "batch launch (capture)");
            check(ended, "graph capture end");
            check(hipGraphInstantiate(&b.graph_exec, b.graph, nullptr, nullptr, 0), "graph instantiate");
            b.graph_blocks = fast_blocks;
        }
        check(hipEventRecord(s.start, s.stream), "event");
        check(hipGraphLaunch(b.graph_exec, s.stream), "graph launch");
        check(hipEventRecord(s.stop, s.stream), "event");
    } else {
        check(hipEventRecord(s.start, s.stream), "event");
        check(launch_batch(b.q, fast_blocks, persistent_blocks, s.stream), "batch launch");
        check(hipEventRecord(s.stop, s.stream), "event");
    }
    check(hipStreamSynchronize(s.stream), "batch render");
    float kernel_ms = 0.0f;
    check(hipEventElapsedTime(&kernel_ms, s.start, s.stop), "event time");
    size_t pixels = 0;
    for (const auto& f : b.fields) pixels += static_cast<size_t>(f.request.width) * f.request.height;
    s.timing.kernel_ms = kernel_ms;
    s.timing.total_ms = static_cast<double>(now_ns() - wall0) * 1e-6;
    s.timing.pixels = static_cast<double>(pixels);
}

void render_direct_on_device(const direct::DirectParams& host, const double* x, const double* y, bool planes,
                             bool device_pointers) {
    LaunchState& s = state();
    std::lock_guard<std::mutex> lock(s.mutex);
    init_locked(s);
    const size_t pixels = static_cast<size_t>(host.width) * static_cast<size_t>(host.height);
    direct::DirectParams p = host;
    double* coords = static_cast<double*>(s.direct_coords.get((static_cast<size_t>(host.width) + host.height) * 8));
    check(hipMemcpyAsync(coords, x, static_cast<size_t>(host.width) * 8, hipMemcpyHostToDevice, s.stream), "H2D x");
    check(hipMemcpyAsync(coords + host.width, y, static_cast<size_t>(host.height) * 8, hipMemcpyHostToDevice,
                         s.stream), "H2D y");
    p.x = coords;
    p.y = coords + host.width;
    if (!device_pointers) {
        p.output = static_cast<float*>(s.output.get(pixels * sizeof(float)));
        if (planes) {
            char* base = static_cast<char*>(s.planes.get(pixels * 48));
            p.plane_iteration = reinterpret_cast<int64_t*>(base);
            p.plane_phase = reinterpret_cast<double*>(base + pixels * 8);
            p.plane_de_x = p.plane_phase + pixels;
            p.plane_de_y = p.plane_de_x + pixels;
            p.plane_test1 = p.plane_de_y + pixels;
            p.plane_test2 = p.plane_test1 + pixels;
        }
    }
    const uint64_t wall0 = now_ns();
    check(hipEventRecord(s.start, s.stream), "event");
    check(launch_direct_kernel(p, planes, s.stream), "direct kernel launch");
    check(hipEventRecord(s.stop, s.stream), "event");
    if (!device_pointers) {
        check(hipMemcpyAsync(host.output, p.output, pixels * sizeof(float), hipMemcpyDeviceToHost, s.stream), "D2H");
        if (planes) {
            check(hipMemcpyAsync(host.plane_iteration, p.plane_iteration, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
            check(hipMemcpyAsync(host.plane_phase, p.plane_phase, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
            check(hipMemcpyAsync(host.plane_de_x, p.plane_de_x, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
            check(hipMemcpyAsync(host.plane_de_y, p.plane_de_y, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
            check(hipMemcpyAsync(host.plane_test1, p.plane_test1, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
            check(hipMemcpyAsync(host.plane_test2, p.plane_test2, pixels * 8, hipMemcpyDeviceToHost, s.stream), "D2H");
        }
    }
    check(hipStreamSynchronize(s.stream), "direct render");
    float kernel_ms = 0.0f;
    check(hipEventElapsedTime(&kernel_ms, s.start, s.stop), "event time");
    s.timing.kernel_ms = kernel_ms;
    s.timing.total_ms = static_cast<double>(now_ns() - wall0) * 1e-6;
    s.timing.pixels = static_cast<double>(pixels);
}

// ---- pipelined segment --------------------------------------------------------

namespace {

// One field's reference and staged tables, built by a worker and rendered
// by the calling thread.
struct SegmentSlot {
    std::unique_ptr<ReferenceContext> context;
    HostTables host;
    StagedTables staged;
    PrebuiltTables tables;
    bool direct = false;
    long double direct_zoom = 0.0L;
};

// The handle render's shallow test (renderer.cpp:19920-19960).
bool shallow_zoom(const char* zoom, long precision_bits) {
    const FE z = parse_zoom(zoom, precision_bits);
    const long double log10_zoom =
        (std::log(static_cast<long double>(z.m)) + static_cast<long double>(z.e) * 0.693147180559945309417232121458176568L)
        / std::log(10.0L);
    return log10_zoom < 6.0L;
}

// Tables staged by a worker: copies on the side stream, and the compute
// stream waits for them before the render's launches (no host wait).
void upload_prepared_tables(ReferenceContext& c, const HostTables& h, StagedTables& staged) {
    std::lock_guard<std::mutex> lock(c.device_mutex);
    LaunchState& s = state();
    {
        std::lock_guard<std::mutex> slock(s.mutex);
        init_locked(s);
    }
    auto t = upload_staged(c, h, staged, s.device, s.side);
    check(hipStreamWaitEvent(s.stream, staged.done, 0), "wait for table upload");
    c.device = std::move(t);
}

}  // namespace

void render_segment(const std::vector<SegmentField>& fields, const FractalRenderOptions& o, int threads,
                    bool pipelined, SegmentTiming& timing) {
    const size_t n = fields.size();
    const uint64_t wall0 = now_ns();
    std::vector<SegmentSlot> slots(n);
    std::mutex m;
    std::condition_variable cv;
    std::deque<size_t> ready;
    std::exception_ptr failure;
    std::atomic<size_t> next{0};
    std::atomic<bool> stop{false};
    uint64_t build_ns = 0, built_ns = 0;
    // References are built longest first (cap x precision, the MPFR orbit's
    // cost), so that the deep fields at the end of a segment do not start
    // last and leave the pool idle but for them; each is rendered as it is
    // ready, so the order changes no result.
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        return static_cast<double>(fields[a].max_iter) * fields[a].precision_bits
            > static_cast<double>(fields[b].max_iter) * fields[b].precision_bits;
    });
    auto work = [&]() {
        for (;;) {
            const size_t at = next.fetch_add(1);
            if (at >= n || stop.load()) return;
            const size_t i = order[at];
            try {
                const uint64_t t0 = now_ns();
                const SegmentField& f = fields[i];
                SegmentSlot& slot = slots[i];
                slot.context = build_reference(f.x.c_str(), f.y.c_str(), f.zoom.c_str(), f.max_iter, f.precision_bits,
                                               8, false, 0, o.escape_radius_mode, o.coordinate_mode);
                slot.direct = shallow_zoom(f.zoom.c_str(), slot.context->precision_bits);
                if (slot.direct) {
                    slot.direct_zoom = direct::parse_zoom_long(f.zoom.c_str());
                } else {
                    build_host_tables(*slot.context, slot.host);
                    stage_tables(slot.host, slot.staged);
                    const FieldRequest request{f.output, f.width, f.height, f.zoom.c_str(), f.max_iter,
                                               f.series_order, f.series_block, o, nullptr, nullptr, true};
                    PreparedRender prep;
                    prepare_render(*slot.context, request, prep);
                    prebuild_field_tables(*slot.context, prep, slot.tables);
                }
                const uint64_t t1 = now_ns();
                std::lock_guard<std::mutex> lock(m);
                build_ns += t1 - t0;
                built_ns = t1 - wall0;
                ready.push_back(i);
            } catch (...) {
                std::lock_guard<std::mutex> lock(m);
                if (!failure) failure = std::current_exception();
                stop = true;
            }
            cv.notify_all();
        }
    };
    std::vector<std::thread> pool;
    const int workers = static_cast<int>(std::max<size_t>(1, std::min<size_t>(static_cast<size_t>(std::max(threads, 1)), n)));
    for (int w = 0; w < workers; ++w) pool.emplace_back(work);
    std::exception_ptr render_failure;
    try {
        for (size_t done = 0; done < n; ++done) {
            size_t i = done;
            {
                std::unique_lock<std::mutex> lock(m);
                cv.wait(lock, [&] { return failure != nullptr || (pipelined ? !ready.empty() : ready.size() == n); });
                if (failure) break;
                if (pipelined) {
                    i = ready.front();
                    ready.pop_front();
                }
            }
            const SegmentField& f = fields[i];
            SegmentSlot& slot = slots[i];
            if (slot.direct) {
                direct::DirectSetup d = direct::make_setup_values(
                    f.width, f.height, slot.direct_zoom, slot.context->x_center, slot.context->y_center, f.max_iter, 0,
                    0.0, 0.0, o.output_bias, o.escape_radius_mode, o.coordinate_mode, static_cast<uint32_t>(o.reserved[0]));
                d.params.output = f.output;
                render_direct_on_device(d.params, d.x.data(), d.y.data(), false, true);
            } else {
                upload_prepared_tables(*slot.context, slot.host, slot.staged);
                install_prebuilt(*slot.context, slot.tables);
                const FieldRequest request{f.output, f.width, f.height, f.zoom.c_str(), f.max_iter, f.series_order,
                                           f.series_block, o, nullptr, nullptr, true};
                render_field(*slot.context, request);
            }
            timing.render_ms += last_timing().kernel_ms;
            // the render has synchronised: the reference, its tables and the staging go
            slot.staged.release();
            slot.host = HostTables{};
            slot.tables = PrebuiltTables{};
            slot.context.reset();
        }
    } catch (...) {
        render_failure = std::current_exception();
        stop = true;
    }
    for (std::thread& t : pool) t.join();
    if (render_failure) std::rethrow_exception(render_failure);
    if (failure) std::rethrow_exception(failure);
    timing.build_ms = static_cast<double>(build_ns) * 1e-6;
    timing.built_ms = static_cast<double>(built_ns) * 1e-6;
    timing.wall_ms = static_cast<double>(now_ns() - wall0) * 1e-6;
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
