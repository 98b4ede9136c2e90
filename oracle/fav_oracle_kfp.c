/*
 * fav_oracle_kfp.c -- placeholder translation unit for the KFP colour and
 * direct-render restatements.  TEST INFRASTRUCTURE ONLY.
 */
#include "fav_oracle.h"

int fo_render_direct(float *output, int width, int height, double zoom, double x_center,
                     double y_center, int max_iter, double output_bias, int escape_radius_mode,
                     int coordinate_mode, FractalRenderPlanes *planes, int threads) {
    (void)output; (void)width; (void)height; (void)zoom; (void)x_center; (void)y_center;
    (void)max_iter; (void)output_bias; (void)escape_radius_mode; (void)coordinate_mode;
    (void)planes; (void)threads;
    return 1;
}
