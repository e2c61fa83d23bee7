/*
 * scanline_oracle.h -- C API of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * The oracle is a literal single-threaded C++ restatement of the reference's
 * comb-scanline rasterizer (gpu-scanline/src/rasterizer/R_cut_A_mask_comb_scanline/
 * ras_cut_mask_comb_scanline.cu and gpu-scanline/src/rasterizer/shared/{ras_cut.h,
 * ras_qm_mask.h, ras_qm_mask.cu, ras_base.cpp}), compiled with -ffp-contract=off and
 * IEEE division/sqrt.  Only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark may load it, and only as the checker -- never as the renderer.  The same
 * source built with -fopenmp is the reported CPU baseline.
 */
#ifndef SCANLINE_ORACLE_H
#define SCANLINE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_result oracle_result;

/* Run the whole pipeline.  Arrays follow include/scanline.h's sl_scene (gradient ramp and
 * table textures last; NULL with size 0 when absent).  `m16` is the column-major input
 * transform (NULL = identity).  flags bit 0: keep per-stage intermediates (needed by
 * oracle_get for everything but "image"). */
oracle_result *oracle_run(uint32_t n_paths, uint32_t n_curves, uint32_t n_vertices,
                          const float *vertex, const uint32_t *curve_vertex_pos,
                          const uint8_t *curve_type, const uint32_t *curve_path_id,
                          const float *curve_arc_w, const uint8_t *fill_rule,
                          const uint32_t *fill_color, int width, int height, int samples,
                          int srgb, int ms_output, const float *m16, int flags,
                          const int32_t *ramp, uint32_t n_ramp, const int32_t *table,
                          uint32_t n_table);

/* Borrow one named array of a result (see oracle_names for the list). */
int oracle_get(const oracle_result *r, const char *name, const void **ptr, size_t *nbytes);
const char *oracle_names(void);

/* Wall time (seconds) of the render part of oracle_run, excluding upload and tables. */
double oracle_render_seconds(const oracle_result *r);
void oracle_free(oracle_result *r);
int oracle_threads(void);

/* ---- pieces exposed for closed-form pinning tests ---- */
/* Sample positions of one pixel (n = 8, 16, 32), as x,y pairs in [0,1)^2. */
int oracle_sample_positions(int n, float *xy);
/* 256x256 line->mask table: words_per_texel = 1 (8 spp), 2 (16 spp), 4 (32 spp). */
int oracle_pixel_table(int n, uint32_t *out);
void oracle_fetch_pixel(int n, float p0x, float p0y, float p1x, float p1y, uint32_t out[4]);
void oracle_fetch_pixel_batch(int n, int64_t count, const float *pts, uint32_t *out);
void oracle_vertical_mask(int n, float y, uint32_t out[4]);
void oracle_vertical_mask2(int n, float y0, float y1, uint32_t out[4]);
void oracle_solve_quad(float a, float b, float c, float *r0, float *r1);
float oracle_curve_eval(int type, float w, const float *v, float t);
/* sRGB8 / UNORM8 tables used by the compositor: dec[256] floats, thr[256] floats. */
void oracle_color_tables(int srgb, float *dec, float *thr);

#ifdef __cplusplus
}
#endif

#endif
