/*
 * CPU oracle for the screen-space GI path (TEST INFRASTRUCTURE ONLY).
 *
 * A scalar C++17 / OpenMP restatement of the reference GLSL passes of
 * LegitEngine's SSVGI renderer.  It is the checker for the HIP path and the
 * CPU baseline timed beside it.  Nothing in the product links or calls it:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the reference repository root).  Pinning status of each function (what
 * known-answer test fixes it) is listed in oracle.cpp's header.
 *
 * Conventions: images are row-major float arrays with C interleaved channels;
 * storage precisions (FP16, RGBA8) are widened exactly to float by the caller;
 * `round_fp16` arguments re-round a pass's result to FP16 (round to nearest
 * even) where the product stores that pass in FP16.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gather parameters (indirectLighting.frag:4-9 uniform block + the constants
 * the shader hard-codes, promoted to parameters as the Hi-Z point-fetch mode needs). */
typedef struct {
  float view[16];   /* world -> camera, column-major (glm layout) */
  float proj[16];   /* camera -> ndc, column-major */
  int32_t width, height;
  int32_t dirs;      /* D: directions per pixel (reference: 4, indirectLighting.frag:166) */
  int32_t max_steps; /* S: step cap; <= 0 means the reference rule only (:212) */
  int32_t grid;      /* interleave pattern size (reference: 4, :141) */
  int32_t sampler;   /* 0 = trilinear over avg+blur pyramids (reference), 1 = Hi-Z point fetch at floor(lod) */
  int32_t n_levels;  /* pyramid levels present */
  int32_t footprint_dirs; /* direction count in the step growth (:212, :217) and lod (:234); <= 0 = dirs */
  int32_t grid_y;    /* pattern height, <= 0 = grid: index (x % grid) + (y % grid_y) * grid */
} orc_gather_params;

uint32_t orc_version(void);

/* Host maths */
void orc_hammersley(int32_t i, int32_t n, float out[2]);
void orc_mat4_inverse(const float m[16], float out[16]);
void orc_perspective_rh_no(float fovy, float aspect, float znear, float zfar, float out[16]);
void orc_camera_transform(const float pos[3], float hor, float vert, float out[16]);
void orc_mat4_mul(const float a[16], const float b[16], float out[16]);
float orc_horizon_contribution(const float eye[3], const float tangent[3], const float n[3], float tmin, float tmax);
double orc_horizon_contribution_f64(const double eye[3], const double tangent[3], const double n[3], double tmin, double tmax);
void orc_oct_decode(float ex, float ey, float out[3]);
uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);

/* Pyramid passes */
int orc_mip_avg(const float *src, int32_t w, int32_t h, int32_t c, float *dst, int32_t round_fp16);
int orc_mip_max(const float *src, int32_t w, int32_t h, float *dst);
int orc_box_blur(const float *src, int32_t w, int32_t h, int32_t c, int32_t radius, float *dst, int32_t round_fp16);

/* Permutations (element = elem_bytes opaque bytes) */
int orc_deinterleave(const void *src, int32_t w, int32_t h, int32_t elem_bytes, int32_t gx, int32_t gy, void *dst);
int orc_interleave(const void *src, int32_t w, int32_t h, int32_t elem_bytes, int32_t gx, int32_t gy, void *dst);

/* Samplers (exposed for known-answer tests) */
void orc_sample_trilinear(const float *const *levels, const int32_t *lw, const int32_t *lh, int32_t n_levels, int32_t c,
                          float sx, float sy, float lod, float *out);
void orc_sample_point(const float *const *levels, const int32_t *lw, const int32_t *lh, int32_t n_levels, int32_t c,
                      float sx, float sy, float lod, float *out);

/* Gather: depth = radial distance (W*H), oct = octahedral normal (W*H*2),
 * pyr_depth[l] (1 channel), pyr_rad[l] (4 channels RGBA).  out = RGBA (W*H*4,
 * alpha 1).  steps (optional, W*H int32) = on-screen steps summed over dirs.
 * threads <= 0 uses all OpenMP threads. */
int orc_gather_f32(const orc_gather_params *p, const float *depth, const float *oct,
                   const float *const *pyr_depth, const float *const *pyr_rad,
                   const int32_t *lw, const int32_t *lh, float *out, int32_t *steps,
                   int32_t x0, int32_t y0, int32_t cw, int32_t ch, int32_t threads);
int orc_gather_f64(const orc_gather_params *p, const float *depth, const float *oct,
                   const float *const *pyr_depth, const float *const *pyr_rad,
                   const int32_t *lw, const int32_t *lh, double *out, int32_t *steps,
                   int32_t x0, int32_t y0, int32_t cw, int32_t ch, int32_t threads);

/* The FP32 gather on whole rows rows[0..n_rows) (any order, repeats allowed):
 * out = n_rows x width RGBA in list order; steps and hits (optional)
 * likewise, hits = on-screen steps whose sample rose above the running
 * horizon (indirectLighting.frag:254) and so fetched radiance.  For
 * stratified samples of a frame: OpenMP spreads 64-pixel row segments. */
int orc_gather_rows_f32(const orc_gather_params *p, const float *depth, const float *oct,
                        const float *const *pyr_depth, const float *const *pyr_rad,
                        const int32_t *lw, const int32_t *lh, const int32_t *rows, int32_t n_rows,
                        float *out, int32_t *steps, int32_t *hits, int32_t threads);

/* Geometry-aware bilateral (denoiser.frag:87-110 weights), taps [-r, r)^2. */
int orc_bilateral(const float *in, const float *depth, const float *oct, int32_t w, int32_t h,
                  int32_t radius, float depth_sigma, float normal_sigma, float *out, int32_t threads);
/* Live regression denoiser (denoiser.frag:72-185). */
int orc_regression_denoise(const float *in, const float *depth, int32_t w, int32_t h, int32_t radius, float *out);

/* Final composite (finalGatherer.frag:42-58). */
int orc_composite(const float *direct, const float *albedo, const float *indirect, int32_t w, int32_t h, float *out);

double orc_rel_l2(const float *x, const float *y, int64_t n, int32_t stride, int32_t channels);
double orc_rel_l2_f64(const float *x, const double *y, int64_t n, int32_t stride, int32_t channels);
int32_t orc_max_threads(void);

#ifdef __cplusplus
}
#endif
