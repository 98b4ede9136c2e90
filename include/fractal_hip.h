/*
 * fractal_hip.h -- C ABI of the MI355X (gfx950) deep-zoom field backend.
 *
 * The library built from fractal-audio-viz_amd/csrc (libfractal_hip.so)
 * exports the hot-path subset of the reference's renderer.h ABI, version 10,
 * with identical names, argument meaning, struct layouts, return codes and
 * thread-local error text.  Every render runs on the HIP device; the
 * reference's CPU backends (0 = scalar, 1 = AVX2) are not part of this
 * library and are rejected with an error instead of being emulated.
 *
 * Struct layouts below must stay byte-identical to renderer.h: they are the
 * wire format that visualizer.py's ctypes mirrors fill in
 * (visualizer.py:824-944, 986-1112, 1117-1215, 1237-1256).
 *
 * Conventions shared by every entry point (renderer.cpp:4596-4599,
 * 15922-15924, 19961-19969):
 *   - int results: 0 = success, 1 = failure; creators return NULL on failure;
 *   - the failure message is read with fractal_last_error() on the same
 *     thread;
 *   - the caller owns every buffer: float32 fields are row-major W*H, planes
 *     are W*H int64/double arrays, RGB is interleaved W*H*3 bytes;
 *   - reference handles are opaque increasing tokens (never pointers).
 */
#ifndef FRACTAL_HIP_H
#define FRACTAL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FRACTAL_ABI_VERSION 10
#define FRACTAL_RENDER_OPTIONS_VERSION 2
#define FRACTAL_KFP_OPTIONS_VERSION 3
#define FRACTAL_KFP_PLANES_VERSION 1
#define FRACTAL_RENDER_PLANES_VERSION 1
#define FRACTAL_KFP_MAX_MULTI_COLORS 256

#define FRACTAL_FORMULA_MANDELBROT 0
#define FRACTAL_FORMULA_JULIA 1
#define FRACTAL_FORMULA_BURNING_SHIP 2
#define FRACTAL_FORMULA_TRICORN 3

#define FRACTAL_COORDINATE_MODE_PROJECT 0
#define FRACTAL_COORDINATE_MODE_KALLES 1

#define FRACTAL_RENDER_HINT_SKIP_PHASE (1u << 0)
#define FRACTAL_RENDER_HINT_SKIP_ANALYTIC_DE (1u << 1)

/* Same layout as renderer.h:43-67. */
typedef struct FractalRenderOptions {
    uint32_t struct_size;
    uint32_t version;
    int32_t strict;
    int32_t allow_recovery;
    int32_t time_budget_ms;
    int32_t disable_bla;
    int32_t disable_cycle;
    int32_t strict_cycle;
    int32_t series_min_terms;
    int32_t series_max_terms;
    int32_t max_bla_length;
    int32_t max_linear_bla_length;
    int32_t backend;
    int32_t escape_radius_mode;
    int32_t coordinate_mode;
    int32_t reserved[1];
    double output_bias;
} FractalRenderOptions;

/* Same layout as renderer.h:75-119. */
typedef struct FractalKfpOptions {
    uint32_t struct_size;
    uint32_t version;
    double iter_div;
    double color_offset;
    double ratio;
    int32_t color_method;
    int32_t smooth_method;
    int32_t smooth;
    int32_t flat;
    int32_t inverse_transition;
    double phase_color_strength;
    int32_t multi_color;
    int32_t blend_multi_color;
    uint32_t multi_color_count;
    double multi_color_period[FRACTAL_KFP_MAX_MULTI_COLORS];
    int32_t multi_color_start[FRACTAL_KFP_MAX_MULTI_COLORS];
    int32_t multi_color_type[FRACTAL_KFP_MAX_MULTI_COLORS];
    double power;
    int32_t slopes;
    double slope_power;
    double slope_ratio;
    double slope_angle;
    int32_t differences;
    int32_t interior_color[3];
    double field_bias;
    int32_t bailout_radius_preset;
    double bailout_radius_custom;
    int32_t bailout_norm_preset;
    double bailout_norm_custom;
    int32_t texture_enabled;
    double texture_merge;
    double texture_power;
    double texture_ratio;
    int32_t texture_resize;
    int32_t use_opengl;
    int32_t use_srgb;
    int32_t show_glitches;
} FractalKfpOptions;

/* Same layout as renderer.h:135-153. */
typedef struct FractalKfpPlanes {
    uint32_t struct_size;
    uint32_t version;
    const int64_t *orbit_iteration;
    const double *bailout;
    const int64_t *iteration;
    const double *transition;
    const double *phase;
    const double *de_x;
    const double *de_y;
    const double *test1;
    const double *test2;
    const uint8_t *texture_rgb;
    int32_t texture_width;
    int32_t texture_height;
    int32_t texture_stride;
} FractalKfpPlanes;

/* Same layout as renderer.h:161-170. */
typedef struct FractalRenderPlanes {
    uint32_t struct_size;
    uint32_t version;
    int64_t *orbit_iteration;
    double *phase;
    double *de_x;
    double *de_y;
    double *test1;
    double *test2;
} FractalRenderPlanes;

/* ---- Library identity and diagnostics --------------------------------- */

/* renderer.h:172 / renderer.cpp:15872.  Returns 10. */
int fractal_abi_version(void);
/* renderer.h:173 / renderer.cpp:15886.  Returns 2. */
int fractal_render_options_version(void);
/* renderer.h:174 / renderer.cpp:15888-15890 (defaults: renderer.cpp:4708-4727),
 * except that backend defaults to 2, the only backend this library runs. */
void fractal_render_options_default(FractalRenderOptions *options);
/* renderer.h:177 / renderer.cpp:15892-15920.  Bit 4 (double-precision GPU
 * compute) and bit 8 (physical GPU) when a gfx950 device is visible; 0
 * otherwise.  The CPU bits 1/2 are never set by this library. */
int fractal_backend_capabilities(void);
/* renderer.h:178 / renderer.cpp:15922-15924. */
const char *fractal_last_error(void);

/* renderer.h:180-182 / renderer.cpp:15874-15884.  Counter order follows
 * RenderStats (renderer.cpp:4615-4706): 16 basic or 39 extended fields. */
void fractal_set_stats_enabled(int enabled);
int fractal_get_last_stats(uint64_t *values, int capacity);
int fractal_get_last_stats_ex(uint64_t *values, int capacity);

/* ---- Reference orbits (host MPFR + BLA build, uploaded to HBM) -------- */

/* renderer.h:184-191 / renderer.cpp:19568-19594. */
void *fractal_create_reference(
    const char *x_center, const char *y_center, const char *viewport_zoom,
    int max_iter, int precision_bits, int series_order);
/* renderer.h:193-200 / renderer.cpp:19596-19622. */
void *fractal_create_reference_reusable(
    const char *x_center, const char *y_center, const char *viewport_zoom,
    int max_iter, int precision_bits, int series_order);
/* renderer.h:203-211 / renderer.cpp:19624-19653. */
void *fractal_create_reference_reusable_options(
    const char *x_center, const char *y_center, const char *viewport_zoom,
    int max_iter, int precision_bits, int series_order,
    const FractalRenderOptions *options);
/* renderer.h:215-225 / renderer.cpp:19655-19685.  Formula 0-3 (Mandelbrot,
 * Julia, burning ship, tricorn). */
void *fractal_create_reference_ex(
    const char *x_center, const char *y_center, const char *viewport_zoom,
    int max_iter, int precision_bits, int series_order, int formula,
    const char *julia_real, const char *julia_imag);
/* renderer.h:226-237 / renderer.cpp:19687-19722. */
void *fractal_create_reference_ex_options(
    const char *x_center, const char *y_center, const char *viewport_zoom,
    int max_iter, int precision_bits, int series_order, int formula,
    const char *julia_real, const char *julia_imag,
    const FractalRenderOptions *options);
/* renderer.h:239 / renderer.cpp:19724-19750 (retargeted BLA radii,
 * renderer.cpp:11776-11868). */
void *fractal_clone_reference(void *source_handle, const char *viewport_zoom);
/* renderer.h:240 / renderer.cpp:19752-19772.  Idempotent on NULL; a stale
 * handle sets an error. */
void fractal_destroy_reference(void *handle);
/* renderer.h:241 / renderer.cpp:19779-19812: build ns (orbit, series, BLA),
 * series iteration, series order. */
int fractal_get_reference_stats(void *handle, uint64_t *values, int capacity);

/* ---- Field renders (HIP) ---------------------------------------------- */

/* renderer.h:243-254. */
int fractal_render_mandelbrot_reference_ex(
    float *output, int width, int height, const char *zoom_text, void *handle,
    int max_iter, int threads, int series_order, int series_block,
    const FractalRenderOptions *options);
/* renderer.h:257-268. */
int fractal_render_reference_ex(
    float *output, int width, int height, const char *zoom_text, void *handle,
    int max_iter, int threads, int series_order, int series_block,
    const FractalRenderOptions *options);
/* renderer.h:273-285 / renderer.cpp:19814-19970. */
int fractal_render_reference_ex_planes(
    float *output, int width, int height, const char *zoom_text, void *handle,
    int max_iter, int threads, int series_order, int series_block,
    const FractalRenderOptions *options, FractalRenderPlanes *planes);
/* renderer.h:376-386 / renderer.cpp:20250-20273. */
int render_mandelbrot_reference(
    float *output, int width, int height, const char *zoom_text, void *handle,
    int max_iter, int threads, int series_order, int series_block);
/* renderer.h:338-353 / renderer.cpp:20279-20358.  use_perturbation=0 runs
 * the direct double-precision renderer (formulas 0-3, render_direct,
 * renderer.cpp:10138-10278); use_perturbation=1 builds a reference and runs
 * the perturbation path (Mandelbrot only). */
int render_fractal_ex(
    float *output, int width, int height, const char *zoom_text,
    const char *x_center, const char *y_center, int max_iter,
    int precision_bits, int use_perturbation, int threads, int formula,
    double julia_real, double julia_imag,
    const FractalRenderOptions *options);
/* renderer.h:323-334 / renderer.cpp:20478-20506: render_fractal_ex with
 * formula 0. */
int render_mandelbrot_ex(
    float *output, int width, int height, const char *zoom_text,
    const char *x_center, const char *y_center, int max_iter,
    int precision_bits, int use_perturbation, int threads,
    const FractalRenderOptions *options);
/* renderer.h:387-398 / renderer.cpp:20508-20533: default options. */
int render_mandelbrot(
    float *output, int width, int height, const char *zoom_text,
    const char *x_center, const char *y_center, int max_iter,
    int precision_bits, int use_perturbation, int threads);
/* renderer.h:357-373 / renderer.cpp:20360-20474.  Planes with
 * use_perturbation=0 follow render_direct_with_planes
 * (renderer.cpp:10280-10426).  Planes are accepted for Mandelbrot, and for
 * Julia when use_perturbation=0 (the reference's GPU backend takes
 * Mandelbrot only); burning ship / tricorn planes fail as there. */
int render_fractal_ex_planes(
    float *output, int width, int height, const char *zoom_text,
    const char *x_center, const char *y_center, int max_iter,
    int precision_bits, int use_perturbation, int threads, int formula,
    double julia_real, double julia_imag,
    const FractalRenderOptions *options, FractalRenderPlanes *planes);

/* ---- KFP colour transfer (HIP) ------------------------------------------ */

/* renderer.h:452-465 / renderer.cpp:16547-16565 (fractal_colourise_kfp_impl,
 * renderer.cpp:16314-16545).  Kalles SetColor of the scalar field with a
 * lut_size-entry RGB LUT; phase/vocal/instrumental/pitch are validated and
 * otherwise unused, as in the reference.  Output is W*H*3 bytes. */
int fractal_colourise_kfp(
    const float *field, uint8_t *output, int width, int height, int max_iter,
    double phase, double vocal, double instrumental, double pitch,
    const FractalKfpOptions *options, const uint8_t *lut, int lut_size,
    int threads);
/* renderer.h:468-481 / renderer.cpp:16567-16600: the reference's optional
 * GPU colour pass; here the same HIP pass. */
int fractal_colourise_kfp_opencl(
    const float *field, uint8_t *output, int width, int height, int max_iter,
    double phase, double vocal, double instrumental, double pitch,
    const FractalKfpOptions *options, const uint8_t *lut, int lut_size,
    int threads);
/* renderer.h:483-498 / renderer.cpp:16729-16748, 16314-16545: with
 * Kalles planes (orbit or stored iteration, bailout, transition, phase, DE,
 * |z|^2 tests, texture; each may be NULL) the colour follows the planes as
 * the reference's SetColor does (renderer.cpp:5655-5976, 6696-7289). */
int fractal_colourise_kfp_planes(
    const float *field, uint8_t *output, int width, int height, int max_iter,
    double phase, double vocal, double instrumental, double pitch,
    const FractalKfpOptions *options, const FractalKfpPlanes *planes,
    const uint8_t *lut, int lut_size, int threads);
/* renderer.h:502-515 / renderer.cpp:18252-18271. */
int fractal_colourise_kfp_precise(
    const float *field, uint8_t *output, int width, int height, int max_iter,
    double phase, double vocal, double instrumental, double pitch,
    const FractalKfpOptions *options, const uint8_t *lut, int lut_size,
    int threads);

/* Raw KFP atlas colour pass: the parent tile (zoomed by parent_zoom) and
 * the child tile (its centre child_fraction of the frame, zoomed by
 * child_zoom, feathered in over up to 64 pixels) reprojected into one
 * output-sized scalar field with interior-preserving bilinear sampling,
 * then one KFP pass.  Tiles stored relative to their own caps when
 * options->field_bias > 0.  renderer.h:663-686 / renderer.cpp:17969-18265. */
int fractal_atlas_colourise_kfp_raw(
    const float *parent, int parent_width, int parent_height,
    int parent_max_iter, const float *child, int child_width, int child_height,
    int child_max_iter, uint8_t *output, int output_width, int output_height,
    double parent_zoom, double child_fraction, double child_zoom, int max_iter,
    double phase, double vocal, double instrumental, double pitch,
    const FractalKfpOptions *options, const uint8_t *lut, int lut_size,
    int threads);
/* renderer.h:688-711 / renderer.cpp:18366-18395: the same pass. */
int fractal_atlas_colourise_kfp_raw_precise(
    const float *parent, int parent_width, int parent_height,
    int parent_max_iter, const float *child, int child_width, int child_height,
    int child_max_iter, uint8_t *output, int output_width, int output_height,
    double parent_zoom, double child_fraction, double child_zoom, int max_iter,
    double phase, double vocal, double instrumental, double pitch,
    const FractalKfpOptions *options, const uint8_t *lut, int lut_size,
    int threads);

/* renderer.h:554-576 / renderer.cpp:16750-17321: an output-sized parent
 * field and a child field placed at (child_left, child_top), each coloured
 * by the KFP pass from its own samples and transfer range, blended across
 * the feather; on the device. */
int fractal_atlas_colourise_kfp(
    const float *parent, int parent_width, int parent_height,
    const float *child, int child_width, int child_height, uint8_t *output,
    int output_width, int output_height, int max_iter, int child_left,
    int child_top, int feather, double phase, double vocal,
    double instrumental, double pitch, const FractalKfpOptions *options,
    const uint8_t *lut, int lut_size, int threads);
/* renderer.h:577-599 / renderer.cpp:18298-18327: the same pass. */
int fractal_atlas_colourise_kfp_precise(
    const float *parent, int parent_width, int parent_height,
    const float *child, int child_width, int child_height, uint8_t *output,
    int output_width, int output_height, int max_iter, int child_left,
    int child_top, int feather, double phase, double vocal,
    double instrumental, double pitch, const FractalKfpOptions *options,
    const uint8_t *lut, int lut_size, int threads);
/* renderer.h:518-535 / renderer.cpp:17803-17968: the centred crop of one
 * scalar tile (interior-preserving) coloured by the KFP pass, on the
 * device. */
int fractal_crop_colourise_kfp(
    const float *source, int source_width, int source_height,
    uint8_t *output, int output_width, int output_height, double zoom_factor,
    int max_iter, double phase, double vocal, double instrumental,
    double pitch, const FractalKfpOptions *options, const uint8_t *lut,
    int lut_size, int threads);
/* renderer.h:536-553 / renderer.cpp:18273-18296: the same pass. */
int fractal_crop_colourise_kfp_precise(
    const float *source, int source_width, int source_height,
    uint8_t *output, int output_width, int output_height, double zoom_factor,
    int max_iter, double phase, double vocal, double instrumental,
    double pitch, const FractalKfpOptions *options, const uint8_t *lut,
    int lut_size, int threads);
/* renderer.h:604-631 / renderer.cpp:17323-17801: the parent and centred
 * child scalar tiles and their Kalles planes reprojected into one frame
 * (field interior-preserving, integer planes nearest, double planes
 * bilinear with the sign-carrying transition kept), then one KFP pass over
 * the composed planes; compose and colour run on the device. */
int fractal_atlas_colourise_kfp_planes(
    const float *parent, int parent_width, int parent_height,
    int parent_max_iter, const FractalKfpPlanes *parent_planes,
    const float *child, int child_width, int child_height, int child_max_iter,
    const FractalKfpPlanes *child_planes, uint8_t *output, int output_width,
    int output_height, double parent_zoom, double child_fraction,
    double child_zoom, int max_iter, int centered_input, double phase,
    double vocal, double instrumental, double pitch,
    const FractalKfpOptions *options, const uint8_t *lut, int lut_size,
    int threads);
/* renderer.h:632-661 / renderer.cpp:18329-18364: the same pass. */
int fractal_atlas_colourise_kfp_planes_precise(
    const float *parent, int parent_width, int parent_height,
    int parent_max_iter, const FractalKfpPlanes *parent_planes,
    const float *child, int child_width, int child_height, int child_max_iter,
    const FractalKfpPlanes *child_planes, uint8_t *output, int output_width,
    int output_height, double parent_zoom, double child_fraction,
    double child_zoom, int max_iter, int centered_input, double phase,
    double vocal, double instrumental, double pitch,
    const FractalKfpOptions *options, const uint8_t *lut, int lut_size,
    int threads);

/* ---- Ordinary (Aurora) palette and per-frame compositors --------------
 * The palette (renderer.cpp:4782-4923) is built on the host; every pixel is
 * computed on the device.  Byte and float results equal the reference's
 * CPU functions. */

/* renderer.h:400-411 / renderer.cpp:15926-15951, 13771-13825. */
int fractal_colourise(
    const float *field, uint8_t *output, int width, int height, int max_iter,
    double phase, double vocal, double instrumental, double pitch, int threads);
/* renderer.h:426-441 / renderer.cpp:16210-16248: the palette projected
 * through three RGB accents, with an interior colour. */
int fractal_colourise_accents(
    const float *field, uint8_t *output, int width, int height, int max_iter,
    double phase, double vocal, double instrumental, double pitch,
    const uint8_t *accents, int interior_red, int interior_green,
    int interior_blue, int threads);
/* renderer.h:442-449 / renderer.cpp:16250-16300 (in place). */
int fractal_apply_aurora_accents(
    uint8_t *output, int width, int height, const uint8_t *accents,
    double pitch, int threads);
/* renderer.h:716-725 / renderer.cpp:18399-18465. */
int fractal_crop_rgb(
    const uint8_t *source, int source_width, int source_height,
    uint8_t *output, int output_width, int output_height, double zoom_factor,
    int threads);
/* renderer.h:726-741 / renderer.cpp:18467-18627. */
int fractal_atlas_composite_rgb(
    const uint8_t *parent, int parent_width, int parent_height,
    const uint8_t *child, int child_width, int child_height, uint8_t *output,
    int output_width, int output_height, double parent_zoom,
    double child_fraction, double child_zoom, int feather, int threads);
/* renderer.h:786-795 / renderer.cpp:18631-18689. */
int fractal_crop_field(
    const float *source, int source_width, int source_height, float *output,
    int output_width, int output_height, double zoom_factor, int threads);
/* renderer.h:797-813 / renderer.cpp:18890-18927. */
int fractal_atlas_field(
    const float *parent, int parent_width, int parent_height,
    int parent_max_iter, const float *child, int child_width,
    int child_height, int child_max_iter, float *output, int output_width,
    int output_height, double parent_zoom, double child_fraction,
    int palette_max_iter, int threads);
/* renderer.h:815-835 / renderer.cpp:18697-18888. */
int fractal_atlas_field_ex(
    const float *parent, int parent_width, int parent_height,
    int parent_max_iter, const float *child, int child_width,
    int child_height, int child_max_iter, float *output, int output_width,
    int output_height, double parent_zoom, double child_fraction,
    double child_zoom, double parent_field_bias, double child_field_bias,
    double output_field_bias, int palette_max_iter, int feather, int threads);
/* renderer.h:943-957 / renderer.cpp:18930-19053. */
int fractal_crop_colourise(
    const float *source, int source_width, int source_height,
    uint8_t *output, int output_width, int output_height, double zoom_factor,
    int max_iter, double phase, double vocal, double instrumental,
    double pitch, int threads);
/* renderer.h:959-976 / renderer.cpp:19055-19094. */
int fractal_crop_colourise_interior(
    const float *source, int source_width, int source_height,
    uint8_t *output, int output_width, int output_height, double zoom_factor,
    int max_iter, double phase, double vocal, double instrumental,
    double pitch, int interior_red, int interior_green, int interior_blue,
    int threads);
/* renderer.h:978-996 / renderer.cpp:19096-19136. */
int fractal_crop_colourise_accents(
    const float *source, int source_width, int source_height,
    uint8_t *output, int output_width, int output_height, double zoom_factor,
    int max_iter, double phase, double vocal, double instrumental,
    double pitch, const uint8_t *accents, int interior_red,
    int interior_green, int interior_blue, int threads);
/* renderer.h:997-1017 / renderer.cpp:19137-19455. */
int fractal_atlas_colourise(
    const float *parent, int parent_width, int parent_height,
    int parent_max_iter, const float *child, int child_width,
    int child_height, int child_max_iter, uint8_t *output, int output_width,
    int output_height, double parent_zoom, double child_fraction,
    int palette_max_iter, double phase, double vocal, double instrumental,
    double pitch, int threads);
/* renderer.h:1019-1042 / renderer.cpp:19457-19508. */
int fractal_atlas_colourise_interior(
    const float *parent, int parent_width, int parent_height,
    int parent_max_iter, const float *child, int child_width,
    int child_height, int child_max_iter, uint8_t *output, int output_width,
    int output_height, double parent_zoom, double child_fraction,
    int palette_max_iter, double phase, double vocal, double instrumental,
    double pitch, int interior_red, int interior_green, int interior_blue,
    int threads);
/* renderer.h:1044-1068 / renderer.cpp:19510-19566. */
int fractal_atlas_colourise_accents(
    const float *parent, int parent_width, int parent_height,
    int parent_max_iter, const float *child, int child_width,
    int child_height, int child_max_iter, uint8_t *output, int output_width,
    int output_height, double parent_zoom, double child_fraction,
    int palette_max_iter, double phase, double vocal, double instrumental,
    double pitch, const uint8_t *accents, int interior_red,
    int interior_green, int interior_blue, int threads);

/* Deep render of caller-supplied dc offsets (mantissas and exponents, one
 * value per point) against a reference handle: the field loop of
 * fractal_render_reference_ex with the points in place of a pixel grid and
 * the farthest point's radius in place of the viewport's, for Mandelbrot
 * and alternate-formula handles.  On this backend: project coordinates, no
 * planes (as the reference's GPU point path).
 * renderer.h:287-300 / renderer.cpp:20019-20216. */
int fractal_render_points(
    float *output, int point_count, const char *zoom_text,
    const double *real_mantissa, const double *imag_mantissa,
    const int32_t *exponents, void *handle, int max_iter, int threads,
    int series_order, int series_block, const FractalRenderOptions *options);
/* renderer.h:307-321 / renderer.cpp:20218-20247: the same; non-NULL planes
 * are refused on this backend ("point metadata requires the native scalar
 * backend"). */
int fractal_render_points_ex_planes(
    float *output, int point_count, const char *zoom_text,
    const double *real_mantissa, const double *imag_mantissa,
    const int32_t *exponents, void *handle, int max_iter, int threads,
    int series_order, int series_block, const FractalRenderOptions *options,
    FractalRenderPlanes *planes);

/* ---- Extensions of the HIP backend (no counterpart in renderer.h) ------ */

/* fractal_create_reference_ex_options with the BLA validity radius
 * eps = 2^bla_epsilon_exponent * |Z| instead of the reference's fixed
 * 2^-38 (renderer.cpp:11644-11645; linear, cubic and alternate tables;
 * -60..-20).  A smaller eps gives shorter BLA blocks whose truncation
 * error is smaller.  No counterpart in the reference. */
void *fractal_hip_create_reference_tuned(
    const char *x_center, const char *y_center, const char *viewport_zoom,
    int max_iter, int precision_bits, int image_series_order, int formula,
    const char *julia_real, const char *julia_imag,
    const FractalRenderOptions *options, int bla_epsilon_exponent);

/* The FP32 floatexp perturbation path of a handle (no counterpart in the
 * reference): deltas in float mantissas with an int32 exponent, FP64
 * orbit/BLA tables, FP64 escape/rebase tests and smoothing.  Mandelbrot map:
 * pixels finish in the FP64 plain-double tail where the FP64 path enters it.
 * Julia, burning ship, tricorn (zoom >= 1e6): the map's exact perturbation
 * step (the burning ship's sign branches and axis-crossing reconstruction)
 * and the real 2x2 BLA blocks in float mantissas.  Writes the float field
 * and per-pixel escape counts (max_iter inside, -1 unresolved);
 * device_pointers = 1 when both already live in HBM. */
int fractal_hip_render_reference_floatexp(
    float *output, int32_t *counts, int width, int height,
    const char *zoom_text, void *handle, int max_iter, int series_order,
    int series_block, const FractalRenderOptions *options,
    int device_pointers);

/* A perturbation field of any formula (Mandelbrot, Julia, burning ship,
 * tricorn; zoom >= 1e6; no planes) with its 32x32 tiles launched in
 * tile_order (a permutation of the ceil(w/32)*ceil(h/32) row-major tiles;
 * NULL = the plain grid), and, when work is not NULL, each pixel's loop
 * passes (BLA blocks, exact, replayed and tail steps) written to it.  Host
 * pointers; the field is identical for every order.  RenderStats are
 * published for Mandelbrot fields when statistics are enabled.  No
 * counterpart in the reference: the schedule hook of the learned tile-cost
 * model. */
int fractal_hip_render_reference_scheduled(
    float *output, int width, int height, const char *zoom_text, void *handle,
    int max_iter, int series_order, int series_block,
    const FractalRenderOptions *options, const uint32_t *tile_order,
    int tile_count, int32_t *work);

/* The FP64 field of a handle (zoom >= 1e6, no planes) and each pixel's
 * escape iteration (max_iter inside, -1 where the field is NaN), host
 * pointers: the FP64 counts the floatexp path is measured against.  No
 * counterpart in the reference (the reference reports counts only through
 * the Mandelbrot iteration plane). */
int fractal_hip_render_reference_counts(
    float *output, int32_t *counts, int width, int height,
    const char *zoom_text, void *handle, int max_iter, int series_order,
    int series_block, const FractalRenderOptions *options);

/* Tile features of a field for the cost model below: renders the
 * quarter-size preview ((w+3)/4 x (h+3)/4, same view) with at most
 * work_limit loop passes per pixel (alternate formulas) or with cap
 * min(max_iter, work_limit) (Mandelbrot map; every pass advances at least
 * one iteration, so passes stay within the limit too) and writes, per 32x32 tile of the
 * w x h field (row-major), its 8x8 patch of log(1 + preview passes) / 8
 * (edge pixels repeated past the preview's border) and log10(zoom) / 100:
 * ceil(w/32) * ceil(h/32) * 65 floats to the host buffer `features`.
 * No counterpart in the reference. */
int fractal_hip_preview_features(
    float *features, int width, int height, const char *zoom_text,
    void *handle, int max_iter, int series_order, int series_block,
    const FractalRenderOptions *options, int work_limit);

/* The learned schedule end to end: preview and features as above, the cost
 * model with `weights` (host, fractal_hip_cost_model_parameter_count()
 * floats), 32x32 tiles sorted by decreasing predicted cost (ties and NaN
 * in row-major order), and the field rendered in that order into the host
 * buffer `output` -- identical to fractal_hip_render_reference_scheduled
 * with any order.  tile_order (optional) receives the order; timings
 * (optional, 5 doubles): preview + features ms, cost model ms, ordered
 * render ms (HIP events), wall ms up to the ordered launch, total wall ms.
 * No counterpart in the reference. */
int fractal_hip_render_reference_learned(
    float *output, int width, int height, const char *zoom_text, void *handle,
    int max_iter, int series_order, int series_block,
    const FractalRenderOptions *options, const float *weights, int work_limit,
    uint32_t *tile_order, double *timings);

/* ---- Learned tile-cost model (no counterpart in the reference) ---------
 * A 65 -> 64 -> 64 -> 1 ReLU network with bf16 weights and inputs and fp32
 * sums that predicts log(1 + cost) of a 32x32 tile from an 8x8 patch of a
 * quarter-resolution preview's per-pixel work (log(1 + passes)) plus
 * log10(zoom) / 100.  It only orders tiles, so renders stay identical.
 * weights: fractal_hip_cost_model_parameter_count() fp32 values, laid out
 * W1[65][64], b1[64], W2[64][64], b2[64], W3[64], b3 and rounded to bf16
 * on use.  Every pointer below is device memory; features hold count * 65
 * floats, row-major.  Calls on one process are serialised. */
int fractal_hip_cost_model_parameter_count(void);

/* costs[i] = network(features[i]).  count in [1, 2^22]. */
int fractal_hip_cost_model_forward(const float *weights, const float *features,
                                   int count, float *costs);

/* gradient (parameter_count floats) of loss = mean((network(x) - t)^2)
 * with respect to the bf16 weights, and loss (one float). */
int fractal_hip_cost_model_gradient(const float *weights,
                                    const float *features,
                                    const float *targets, int count,
                                    float *gradient, float *loss);

/* steps full-batch Adam updates (beta 0.9 / 0.999, eps 1e-8) of the fp32
 * master weights, Adam step numbers first_step, first_step + 1, ...;
 * adam_m / adam_v are the moment buffers (zero before the first step);
 * loss_history[k] receives the loss before update k.  Deterministic. */
int fractal_hip_cost_model_train(float *weights, float *adam_m, float *adam_v,
                                 const float *features, const float *targets,
                                 int count, int steps, float learning_rate,
                                 int first_step, float *loss_history);

/* ---- Batched fields (no counterpart in the reference) ------------------
 * Several Mandelbrot fields with their own reference handles, sizes, zooms
 * and caps rendered by one launch sequence over a shared device tile queue
 * (the keyframe fields of one atlas segment).  Each field's output is
 * bit-identical to its single-field render; outputs are device pointers.
 * A field below zoom 1e6 is rendered, as its single-field render is, by the
 * direct double renderer at its handle's centre (renderer.cpp:19920-19960),
 * from the same tile queue. */
typedef struct FractalHipBatchField {
    void *handle;
    const char *zoom_text;
    float *device_output;
    int32_t width;
    int32_t height;
    int32_t max_iter;
    int32_t series_order;
    int32_t series_block;
    int32_t reserved;
} FractalHipBatchField;

/* Uploads every field's launch parameters, offsets and state storage; the
 * handles stay referenced until the batch is destroyed. */
void *fractal_hip_batch_create(const FractalHipBatchField *fields, int count,
                               const FractalRenderOptions *options);
/* Number of 16x16 tiles over all fields (field-major order). */
int fractal_hip_batch_tile_count(void *batch);
/* Tile order: 0 field-major, 1 fields interleaved tile by tile, 2 heaviest
 * first by tile_costs (cost_count = tile count, field-major order). */
int fractal_hip_batch_set_order(void *batch, int order, const float *tile_costs,
                                int cost_count);
/* One render of the batch; fast_blocks_per_cu 0 keeps the current grid,
 * use_graph 1 replays a captured HIP graph of the launch sequence. */
/* Precision of a batch's perturbation fields: 0 = the FP64 launch sequence
 * (default, each field bit-identical to its single-field render), 1 = the
 * FP32 floatexp pixel of fractal_hip_render_reference_floatexp in one
 * persistent kernel over the tile queue (fast_blocks_per_cu blocks per CU);
 * shallow fields stay on the direct renderer.  No counterpart in the
 * reference. */
int fractal_hip_batch_set_precision(void *batch, int floatexp);
int fractal_hip_batch_render(void *batch, int fast_blocks_per_cu, int use_graph);
void fractal_hip_batch_destroy(void *batch);

/* ---- Pipelined segment (no counterpart in the reference) ---------------
 * The keyframe fields of a segment with their references built here:
 * build_threads host threads build each field's MPFR reference orbit, image
 * series and BLA tables (fractal_create_reference_ex_options with image
 * series order 8, Mandelbrot) and stage the tables in pinned memory; the
 * calling thread copies each staged reference to HBM with
 * hipMemcpyAsync on a side stream and renders the field on the compute
 * stream while later references are still being built (the reference's
 * pipeline builds every tier on a pool first, visualizer.py:16834-16990,
 * then uploads and renders synchronously, renderer.cpp:9079-9252).
 * pipelined 0 builds every reference before the first render.  Each field
 * equals its single-field handle render (a field below zoom 1e6 goes
 * through the direct renderer at its centre).  Outputs are device
 * pointers.  timings, when set, receives 4 doubles: wall ms of the call,
 * summed host build ms, summed field-launch ms, wall ms until the last
 * reference was staged.  Returns 0 or 1 (fractal_last_error). */
typedef struct FractalHipSegmentField {
    const char *x_center;
    const char *y_center;
    const char *zoom_text;
    float *device_output;
    int32_t width;
    int32_t height;
    int32_t max_iter;
    int32_t precision_bits;
    int32_t series_order;
    int32_t series_block;
} FractalHipSegmentField;

int fractal_hip_render_segment(const FractalHipSegmentField *fields, int count,
                               const FractalRenderOptions *options, int build_threads,
                               int pipelined, double *timings);

/* fractal_colourise_kfp on device pointers: device_field (W*H floats) and
 * device_rgb (W*H*3 bytes) already live in HBM; lut is a host pointer. */
int fractal_hip_colourise_kfp_device(
    const float *device_field, uint8_t *device_rgb, int width, int height,
    int max_iter, const FractalKfpOptions *options, const uint8_t *lut,
    int lut_size);

/* fractal_hip_colourise_kfp_device with no device allocation.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results; the LUT (a host pointer,
 * at most 65536 entries: more is an error) is staged into, and the colour
 * method 4 range reduction works in, a fixed 196,864-byte scratch array of
 * the library that each device holds until the process ends.  Every call on
 * a device shares that scratch: calls that overlap on one device race. */
int fractal_hip_colourise_kfp_device_noalloc(
    const float *device_field, uint8_t *device_rgb, int width, int height,
    int max_iter, const FractalKfpOptions *options, const uint8_t *lut,
    int lut_size);

/* Same contract as fractal_render_reference_ex_planes, but `device_output`
 * (width*height floats) and the plane arrays of `device_planes` (may be
 * NULL) are device pointers already allocated in HBM.  Used to time the
 * field kernel with its inputs and outputs resident on the GPU. */
int fractal_hip_render_reference_device(
    float *device_output, int width, int height, const char *zoom_text,
    void *handle, int max_iter, int series_order, int series_block,
    const FractalRenderOptions *options, FractalRenderPlanes *device_planes);

/* Timing of the most recent field launch: values[0] = field kernel time in
 * ms from HIP events on the library stream, values[1] = wall time of the
 * call in ms (launch to results on the host), values[2] = pixels, and, when
 * capacity >= 4, values[3] = the plain pass's own kernel time in ms from
 * HIP events around it on the same stream (0 when it did not run).
 * Returns the number of values written (3 or 4), or -1 if capacity < 3. */
int fractal_hip_last_timing(double *values, int capacity);

/* Copies a reference's orbit and BLA tables into HBM now instead of at its
 * first render.  Returns 0 on success, 1 on failure. */
int fractal_hip_reference_upload(void *handle);

#ifdef __cplusplus
}
#endif

#endif /* FRACTAL_HIP_H */
