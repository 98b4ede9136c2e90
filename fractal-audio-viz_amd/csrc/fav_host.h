// fav_host.h -- host-side reference context of the HIP field backend.
//
// A reference is built once on the CPU (MPFR orbit, image series and BLA
// hierarchies, renderer.cpp:10640-11868), then its render tables are copied
// into HBM and stay resident until the handle is destroyed.
#pragma once

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fav_device.h"
#include "fav_numeric.h"

namespace fav {

// Compact orbit shared by a root reference and its radius tier clones
// (ReferenceOrbitData, renderer.cpp:8881-8890) plus the per-index values the
// kernel reads instead of recomputing.
struct OrbitData {
    std::vector<SC> scaled;
    std::vector<double> real_double, imag_double;
    std::vector<double> zr_m, zi_m, margin_m, glitch_m;
    std::vector<int32_t> zr_e, zi_e, margin_e, glitch_e;
    // Values of the loop top for a perturbation negligible against Z_n
    // (see negligible_delta_exponent in reference_build.cpp): |Z_n|^2 as the
    // norm R^2 + margin, the largest delta exponent for which Z_n + delta,
    // the margin and that norm round to the orbit-only values, and the
    // glitch / escape outcomes of the loop-top tests for them.
    std::vector<double> zn_m;
    std::vector<int32_t> zn_e, small_exp, flags;
    // alternate formulas: the MPFR escape margin |Z_n|^2 - R^2
    // (renderer.cpp:10852-10860), read instead of the FloatExp one
    std::vector<double> alt_margin_m;
    std::vector<int32_t> alt_margin_e;
};

struct AltLevels {
    std::vector<std::vector<AltStepDev>> levels;
};

struct CubicLevels {
    std::vector<std::vector<CubicStepDev>> levels;
};

struct LinearLevels {
    std::vector<std::vector<LinearStepDev>> levels;
};

// Device-resident copy of one reference's tables.
struct DeviceTables {
    int device = -1;
    void* orbit_block = nullptr;   // all orbit arrays in one allocation
    void* linear_block[2] = {nullptr, nullptr};
    void* cubic_block = nullptr;
    void* series_block = nullptr;
    OrbitDev orbit{};
    SeriesDev series{};
    const LinearStepDev* linear[2] = {nullptr, nullptr};
    LevelDir linear_dir[2]{};
    const CubicStepDev* cubic = nullptr;
    LevelDir cubic_dir{};
    void* alt_block = nullptr;
    void* fast_block = nullptr;
    // plain-double block loop tables per linear tier and level cap, built on
    // first use (ensure_plain_tables); null until then
    void* rung_block[2][kMaxLevels][2] = {};
    const RungDev* rungs[2][kMaxLevels][2] = {};
    bool plain_built[2][kMaxLevels][2] = {};
    // the walk of the pixel on the reference per table (linear tiers 0, 1;
    // cubic 2) and floor(log2) of the block-length limit, built on first use
    // (ensure_ref_walk); null until then
    void* walk_block[3][kMaxLevels] = {};
    const RefWalkDev* walks[3][kMaxLevels] = {};
    bool walk_built[3][kMaxLevels] = {};
    const AltStepDev* alt = nullptr;
    const double* alt_margin_m = nullptr;  // inside orbit_block
    const int32_t* alt_margin_e = nullptr;
    LevelDir alt_dir{};
    size_t bytes = 0;
    ~DeviceTables();
};

struct ReferenceContext {
    int formula = 0;
    int escape_radius_mode = 0;
    int coordinate_mode = 0;
    // BLA validity radius eps = 2^bla_epsilon_exponent * |Z| (the
    // reference's -38, renderer.cpp:11644-11645)
    int bla_epsilon_exponent = -38;
    long precision_bits = 0;
    int requested_max_iter = 0;
    int requested_series_order = 8;
    bool reusable = false;
    std::string x_text, y_text;
    // builder orbit (FloatExp pairs), retained for reusable roots only
    std::vector<FEC> fast_orbit;
    std::shared_ptr<const OrbitData> orbit;
    int map_end = 0;
    FE input_radius{1.0, 0};
    FE deep_input_radius{0.0, 0};
    // ImageSeries, renderer.cpp:8591-8597
    bool series_enabled = false;
    int series_order = 0;
    int series_iteration = 1;
    FE series_radius_squared{0.0, 0};
    std::vector<SC> series_coefficients;
    CubicLevels cubic;
    LinearLevels linear[2];  // [0] ordinary, [1] deep (input radius * 1e-20)
    // alternate formulas (renderer.cpp:8527-8531, 8881-8905)
    SC parameter{0.0, 0.0, 0};
    AltLevels alt;
    FE alt_input_radius_squared{0.0, 0};
    int alt_start = 1;
    // centre and Julia constant as doubles, for the direct renderer below 1e6
    long double x_center = 0.0L, y_center = 0.0L;
    double julia_re = 0.0, julia_im = 0.0;
    uint64_t reference_build_ns = 0, series_build_ns = 0, bla_build_ns = 0;
    uint64_t upload_ns = 0;
    std::mutex device_mutex;
    std::shared_ptr<DeviceTables> device;  // deleter bound where the tables are uploaded
};

// Builders (reference_build.cpp).  They throw std::runtime_error.
std::unique_ptr<ReferenceContext> build_reference(
    const char* x, const char* y, const char* zoom, int max_iter, int precision_bits,
    int series_order, bool reusable, int formula, int escape_radius_mode, int coordinate_mode,
    const char* julia_re = "0", const char* julia_im = "0", int bla_epsilon_exponent = -38);

// Alternate formulas (alt_build.cpp): the real 2x2 linear BLA hierarchy and
// the settled-cycle repair of the projected orbit.
void build_alternate_bla(ReferenceContext& context);
void stabilize_alternate_cycle(ReferenceContext& context, OrbitData& orbit);
std::unique_ptr<ReferenceContext> clone_reference(const ReferenceContext& source, const char* zoom);

// parse_zoom_float_exp, renderer.cpp:11892-11915.
FE parse_zoom(const char* text, long precision_bits);

// Device side (device_tables.cpp).
void ensure_device_tables(ReferenceContext& context);

}  // namespace fav
