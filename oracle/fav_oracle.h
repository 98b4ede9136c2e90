/*
 * fav_oracle.h -- CPU oracle for the deep-zoom field path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg compare the HIP
 * product against.  The product (libfractal_hip.so) never links, loads or
 * calls it.
 *
 * It is a plain-C restatement of the reference's scalar FP64 path in
 * renderer.cpp (citations on each function in fav_oracle.c), compiled with
 * -ffp-contract=off so every product and sum is rounded separately, exactly
 * as the reference build used for the golden fixtures.
 *
 * Parity of this restatement is pinned by tests/test_oracle_*.py:
 *   - bit-exact against golden vectors written by the reference's own
 *     renderer.cpp compiled from source (oracle/_ref, tests/golden/);
 *   - against closed forms and brute force (direct MPFR iteration of single
 *     pixels, the analytic viewport mapping, the radius-2 escape set).
 */
#ifndef FAV_ORACLE_H
#define FAV_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/fractal_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fo_reference fo_reference;

/* Number of RenderStats counters (renderer.cpp:4615-4643, 4651). */
#define FO_STATS_FIELDS 39

fo_reference *fo_reference_create(
    const char *x_center, const char *y_center, const char *viewport_zoom,
    int max_iter, int precision_bits, int series_order,
    int escape_radius_mode, int coordinate_mode,
    char *error, int error_capacity);

/* fractal_create_reference_ex_options with a formula (0 Mandelbrot, 1 Julia,
 * 2 burning ship, 3 tricorn; renderer.cpp:13650-13700).  julia_re/julia_im
 * are the Julia constant's decimal text (ignored otherwise). */
fo_reference *fo_reference_create_formula(
    const char *x, const char *y, const char *zoom, int max_iter, int precision_bits,
    int series_order, int escape_radius_mode, int coordinate_mode, int formula,
    const char *julia_re, const char *julia_im, char *error, int error_capacity);

/* The same with the BLA validity radius eps = 2^bla_epsilon_exponent
 * (the reference's -38, renderer.cpp:11644-11645; -60..-20).
 * escape_radius_mode 2 (every function here) is radius 1e6 (R^2 = 1e12):
 * not a mode of the reference or of the HIP library; the oracle carries it
 * to show what the reference's arithmetic does with that radius. */
fo_reference *fo_reference_create_tuned(
    const char *x, const char *y, const char *zoom, int max_iter, int precision_bits,
    int series_order, int escape_radius_mode, int coordinate_mode, int formula,
    const char *julia_re, const char *julia_im, int bla_epsilon_exponent,
    char *error, int error_capacity);

fo_reference *fo_reference_clone(
    const fo_reference *source, const char *viewport_zoom,
    char *error, int error_capacity);

void fo_reference_destroy(fo_reference *reference);

/* values[0..] = orbit length, map_end, series enabled, series iteration,
 * series order, cubic level count, linear level count, deep linear level
 * count, cubic level-0 size.  Returns the number written. */
int fo_reference_info(const fo_reference *reference, int64_t *values, int capacity);

/* Copies the compact orbit (mantissa pairs and shared exponents). */
int fo_reference_orbit(const fo_reference *reference, double *real, double *imag,
                       int32_t *exponent, int capacity);

/* Deep Mandelbrot field render for log10(zoom) >= 6 (renderer.cpp:12672-13583).
 * stats may be NULL or hold FO_STATS_FIELDS counters.  threads <= 0 uses
 * the OpenMP default.  Returns 0 on success, 1 on failure. */
int fo_render(
    float *output, int width, int height, const char *zoom_text,
    const fo_reference *reference, int max_iter, int series_order,
    int series_block, const FractalRenderOptions *options,
    FractalRenderPlanes *planes, uint64_t *stats, int threads,
    char *error, int error_capacity);

/* fo_render restricted to the crop_width x crop_height rectangle at
 * (crop_x, crop_y) of the width x height field: output (and planes) hold
 * the rectangle row-major; each pixel's offset and every field-wide
 * decision are the full field's, so the values equal the matching pixels of
 * the full render.  stats count the rectangle's pixels. */
int fo_render_crop(
    float *output, int width, int height, int crop_x, int crop_y,
    int crop_width, int crop_height, const char *zoom_text,
    const fo_reference *reference, int max_iter, int series_order,
    int series_block, const FractalRenderOptions *options,
    FractalRenderPlanes *planes, uint64_t *stats, int threads,
    char *error, int error_capacity);

/* Deep render (Mandelbrot or alternate formula) of caller-supplied dc offsets (mantissa pairs and
 * exponents, one value per point; fractal_render_points, renderer.cpp:
 * 20019-20186).  Glitched points stay NaN in strict mode (no MPFR repair).
 * Returns 0 on success, 1 on failure. */
int fo_render_points(
    float *output, int point_count, const char *zoom_text,
    const double *real_mantissa, const double *imag_mantissa,
    const int32_t *exponents, const fo_reference *reference, int max_iter,
    int series_order, int series_block, const FractalRenderOptions *options,
    uint64_t *stats, int threads, char *error, int error_capacity);

/* Shallow direct render in plain doubles (render_direct / render_direct_with_planes,
 * renderer.cpp:10009-10426): zoom and centre are decimal text parsed as
 * long double and rounded to double; formula 0-3 (Mandelbrot, Julia,
 * burning ship, tricorn); planes may be NULL; plane_flags are the
 * FRACTAL_RENDER_HINT_* bits. */
int fo_render_direct(
    float *output, int width, int height, const char *zoom, const char *x_center,
    const char *y_center, int max_iter, int formula, double julia_real, double julia_imag,
    double output_bias, int escape_radius_mode, int coordinate_mode, uint32_t plane_flags,
    FractalRenderPlanes *planes, int threads, char *error, int error_capacity);

/* KFP SetColor transfer of a scalar field (renderer.cpp:16314-16545). */
int fo_colourise_kfp(
    const float *field, uint8_t *output, int width, int height, int max_iter,
    const FractalKfpOptions *options, const FractalKfpPlanes *planes,
    const uint8_t *lut, int lut_size, int threads,
    char *error, int error_capacity);

/* Raw KFP atlas colour pass (renderer.cpp:17969-18265): parent and child
 * scalar tiles reprojected into one frame, the child feathered into the
 * centre, then the scalar-field KFP transfer. */
int fo_atlas_colourise_kfp_raw(
    const float *parent, int parent_width, int parent_height, int parent_max_iter,
    const float *child, int child_width, int child_height, int child_max_iter,
    uint8_t *output, int output_width, int output_height, double parent_zoom,
    double child_fraction, double child_zoom, int max_iter,
    const FractalKfpOptions *options, const uint8_t *lut, int lut_size, int threads,
    char *error, int error_capacity);

/* 1024-entry cyclic sine-interpolated palette (visualizer.py:13492-13521). */
int fo_kfp_palette_lut(const uint8_t *stops, int stop_count, uint8_t *lut,
                       int lut_size);

/* Ordinary palette and per-frame compositors (fav_oracle_aurora.c): each
 * fo_X takes the arguments of the renderer.h entry point fractal_X
 * (renderer.h:400-449, 713-835, 943-1068) and returns 0, or 1 for invalid
 * arguments. */
int fo_colourise(const float *field, uint8_t *output, int width, int height, int max_iter, double phase,
                 double vocal, double instrumental, double pitch, int threads);
int fo_colourise_accents(const float *field, uint8_t *output, int width, int height, int max_iter, double phase,
                         double vocal, double instrumental, double pitch, const uint8_t *accents, int interior_red,
                         int interior_green, int interior_blue, int threads);
int fo_apply_aurora_accents(uint8_t *output, int width, int height, const uint8_t *accents, double pitch,
                            int threads);
int fo_crop_rgb(const uint8_t *source, int source_width, int source_height, uint8_t *output, int output_width,
                int output_height, double zoom_factor, int threads);
int fo_atlas_composite_rgb(const uint8_t *parent, int parent_width, int parent_height, const uint8_t *child,
                           int child_width, int child_height, uint8_t *output, int output_width, int output_height,
                           double parent_zoom, double child_fraction, double child_zoom, int feather, int threads);
int fo_crop_field(const float *source, int source_width, int source_height, float *output, int output_width,
                  int output_height, double zoom_factor, int threads);
int fo_atlas_field(const float *parent, int parent_width, int parent_height, int parent_max_iter, const float *child,
                   int child_width, int child_height, int child_max_iter, float *output, int output_width,
                   int output_height, double parent_zoom, double child_fraction, int palette_max_iter, int threads);
int fo_atlas_field_ex(const float *parent, int parent_width, int parent_height, int parent_max_iter,
                      const float *child, int child_width, int child_height, int child_max_iter, float *output,
                      int output_width, int output_height, double parent_zoom, double child_fraction,
                      double child_zoom, double parent_field_bias, double child_field_bias, double output_field_bias,
                      int palette_max_iter, int feather, int threads);
int fo_crop_colourise(const float *source, int source_width, int source_height, uint8_t *output, int output_width,
                      int output_height, double zoom_factor, int max_iter, double phase, double vocal,
                      double instrumental, double pitch, int threads);
int fo_crop_colourise_interior(const float *source, int source_width, int source_height, uint8_t *output,
                               int output_width, int output_height, double zoom_factor, int max_iter, double phase,
                               double vocal, double instrumental, double pitch, int interior_red, int interior_green,
                               int interior_blue, int threads);
int fo_crop_colourise_accents(const float *source, int source_width, int source_height, uint8_t *output,
                              int output_width, int output_height, double zoom_factor, int max_iter, double phase,
                              double vocal, double instrumental, double pitch, const uint8_t *accents,
                              int interior_red, int interior_green, int interior_blue, int threads);
int fo_atlas_colourise(const float *parent, int parent_width, int parent_height, int parent_max_iter,
                       const float *child, int child_width, int child_height, int child_max_iter, uint8_t *output,
                       int output_width, int output_height, double parent_zoom, double child_fraction,
                       int palette_max_iter, double phase, double vocal, double instrumental, double pitch,
                       int threads);
int fo_atlas_colourise_interior(const float *parent, int parent_width, int parent_height, int parent_max_iter,
                                const float *child, int child_width, int child_height, int child_max_iter,
                                uint8_t *output, int output_width, int output_height, double parent_zoom,
                                double child_fraction, int palette_max_iter, double phase, double vocal,
                                double instrumental, double pitch, int interior_red, int interior_green,
                                int interior_blue, int threads);
int fo_atlas_colourise_accents(const float *parent, int parent_width, int parent_height, int parent_max_iter,
                               const float *child, int child_width, int child_height, int child_max_iter,
                               uint8_t *output, int output_width, int output_height, double parent_zoom,
                               double child_fraction, int palette_max_iter, double phase, double vocal,
                               double instrumental, double pitch, const uint8_t *accents, int interior_red,
                               int interior_green, int interior_blue, int threads);

#ifdef __cplusplus
}
#endif

#endif
