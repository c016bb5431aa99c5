/*
 * lgx — MI355X-native screen-space indirect lighting (C ABI).
 *
 * Drop-in for the compute passes of LegitEngine's SSVGI renderer:
 *   mip pyramid builder      MipBuilder::BuildMips        (src/Render/Common/MipBuilder.h:142-181)
 *                            + mipLevelBuilder.frag:17-43
 *   per-mip box pre-blur     BlurBuilder::ApplyBlur       (src/Render/Common/BlurBuilder.h:14-46)
 *                            + blurLayerBuilder.frag:17-35
 *   (de)interleave           InterleaveBuilder::Deinterleave / Interleave
 *                            (src/Render/Common/InterleaveBuilder.h:14-80)
 *                            + deinterleave.frag:16-28, interleave.frag:15-27
 *   horizon gather           IndirectLightPass (src/Render/Renderers/SSVGIRenderer.h:223-263)
 *                            + SSVGI/indirectLighting.frag:114-272
 *   denoise                  DenoiserPass (SSVGIRenderer.h:265-302) + denoiser.frag:72-185
 *                            (live regression filter, and the geometry-aware bilateral of
 *                            its commented weights at denoiser.frag:87-110)
 *   composite                final gathering pass (SSVGIRenderer.h:304-342) + finalGatherer.frag:42-58
 *
 * The reference binds these passes through LegitVulkan's RenderGraph
 * (AddPass + record funcs, proxy image ids); the C++ RenderGraph in
 * include/lgx/render_graph.h keeps that shape and calls the entry points
 * below from its record funcs.
 *
 * Conventions
 *  - All pointers in lgx_image are device pointers (hipMalloc'd memory);
 *    entry points never allocate device memory.  Host pointers are marked.
 *  - Streams are hipStream_t (declared here without HIP headers).
 *  - Return value: 0 on success, a negative LGX_E* code for a rejected
 *    argument (nothing is launched), or a positive hipError_t from the launch.
 *    Nothing throws or aborts across the ABI; lgx_strerror names the code.
 *  - Calls are asynchronous on the given stream unless documented otherwise.
 */
#ifndef LGX_H
#define LGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *lgx_stream_t; /* identical to hipStream_t */

#define LGX_MAX_LEVELS 10 /* MippedProxy::mipsCount (MipBuilder.h:21) */
#define LGX_MAX_DIRS 256
#define LGX_MAX_STEPS 1024

/* Pixel formats (vk::Format counterparts in SSVGIRenderer.h:394-404). */
enum {
  LGX_R32F = 1,    /* eR32Sfloat */
  LGX_RG32F = 2,   /* eR32G32Sfloat */
  LGX_RGBA32F = 3, /* eR32G32B32A32Sfloat */
  LGX_RG16F = 4,   /* eR16G16Sfloat */
  LGX_RGBA16F = 5, /* eR16G16B16A16Sfloat */
  LGX_RGBA8 = 6,   /* eR8G8B8A8Unorm */
  LGX_RGBA64F = 7, /* FP64 accumulation targets only */
  LGX_R16F = 8
};

/* Error codes */
enum {
  LGX_OK = 0,
  LGX_EINVAL = -1,       /* null pointer or bad enum */
  LGX_EFORMAT = -2,      /* format not accepted by this entry point */
  LGX_ESIZE = -3,        /* extents inconsistent (e.g. floor-halving, grid divisibility) */
  LGX_EUNSUPPORTED = -4, /* parameter outside the supported range (dirs, steps, levels) */
  LGX_EWORKSPACE = -5,   /* workspace missing or not prepared for these parameters */
  LGX_ENODEVICE = -6     /* no HIP device */
};

typedef struct {
  void *ptr;        /* device pointer to texel (0,0) */
  int32_t w, h;     /* extent in texels */
  int32_t pitch;    /* row pitch in texels (>= w) */
  int32_t format;   /* LGX_* format */
  int32_t reserved; /* must be 0 */
} lgx_image;

/* G-buffer consumed by the gather (gBufferBuilder.frag:28-38 contract):
 * depth = radial distance ||P - cam|| (R32F), normal = octahedral world
 * normal (RG16F or RG32F), albedo (RGBA8 or RGBA32F, linear),
 * radiance = direct light (RGBA16F or RGBA32F). */
typedef struct {
  lgx_image depth, normal, albedo, radiance;
} lgx_gbuffer;

/* Pyramids sampled by the gather: level 0 is full resolution. */
typedef struct {
  int32_t n_levels;
  lgx_image depth[LGX_MAX_LEVELS];    /* R32F: max-Z Hi-Z, or avg + box pre-blur (reference mode) */
  lgx_image radiance[LGX_MAX_LEVELS]; /* RGBA16F / RGBA32F avg + box pre-blur */
} lgx_pyramid;

enum { LGX_MIP_AVG = 0, LGX_MIP_MAX = 1 };                         /* MipBuilder::FilterTypes (MipBuilder.h:136-140) */
enum { LGX_SAMPLER_TRILINEAR = 0, LGX_SAMPLER_HIZ_POINT = 1 };    /* depth/radiance fetch mode */
/* F32: float math and sums; F64: double math and sums; F32_ACC64: float
 * per-step math, double accumulation of the contributions (output RGBA64F). */
enum { LGX_PREC_F32 = 0, LGX_PREC_F64 = 1, LGX_PREC_F32_ACC64 = 2 };
enum { LGX_OUT_INTERLEAVED = 0, LGX_OUT_SLICES = 1 };

typedef struct {
  float view[16]; /* world -> camera, column-major (IndirectLightingData, SSVGIRenderer.h:476-481) */
  float proj[16]; /* camera -> ndc, column-major */
  int32_t width, height; /* viewport (viewportSize.xy) */
  int32_t dirs;      /* directions per pixel (reference: 4) */
  int32_t max_steps; /* step cap S; <= 0 = reference step rule only */
  int32_t grid;      /* interleave pattern width (reference: 4; square unless grid_y is set) */
  int32_t sampler;   /* LGX_SAMPLER_* */
  int32_t precision; /* LGX_PREC_*: F64 computes and accumulates in double (output RGBA64F) */
  int32_t out_layout;/* LGX_OUT_*: SLICES writes the output deinterleaved (grid x grid slices) */
  int32_t *step_counts; /* optional device int32[W*H]: on-screen steps per pixel (sum over dirs), NULL = off */
  int32_t tile;      /* launch shape: an index below lgx_gather_tiles() (0 = default) */
  int32_t roi[4];    /* x0, y0, w, h: gather only these pixels (w = 0: the whole frame); output stays full-frame */
  int32_t footprint_dirs; /* direction count in the step growth and lod (indirectLighting.frag:212,217,234);
                             <= 0 = dirs, as in the shader.  A few directions with a finer footprint feed the
                             learned denoiser. */
  int32_t grid_y;    /* interleave pattern height, <= 0 = grid (the shader's square pattern).  Pixel (x, y)
                        takes pattern index (x % grid) + (y % grid_y) * grid of grid * grid_y Hammersley
                        offsets (:155-164); grid * grid_y <= 64.  LGX_OUT_SLICES needs the pattern to divide
                        the frame. */
  int32_t *hit_counts; /* optional device int32[W*H], with step_counts: the steps whose sample rose above the
                          running horizon (:254), i.e. the ones that fetch radiance, summed over dirs; NULL = off */
} lgx_gather_params;

/* Workspace for the gather: per-(pattern, direction) screen directions and
 * per-(pattern, step) offsets and levels, computed on the host in the
 * shader's own arithmetic.  Depends only on width/height/dirs/max_steps/
 * grid/precision, not on the camera: prepare once per resolution. */
size_t lgx_gather_workspace_bytes(const lgx_gather_params *p);
/* Fills the workspace (device memory of lgx_gather_workspace_bytes).
 * Synchronous: returns after the upload has completed. */
int lgx_gather_prepare(const lgx_gather_params *p, void *workspace);

/* IndirectLightPass (SSVGIRenderer.h:223-263; indirectLighting.frag:114-272).
 * out: RGBA32F (F32) or RGBA64F (F64), width x height.  A launch shape that
 * swizzles (lgx_gather_tile_info) first writes column-swizzled copies of the
 * fine pyramid levels into the workspace, after the tables: calls that
 * overlap in time on one such workspace race. */
int lgx_gather(const lgx_gather_params *p, void *workspace, const lgx_gbuffer *g, const lgx_pyramid *pyr,
               lgx_image *out, lgx_stream_t stream);

/* IndirectLightPass without allocations: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Computes
 * exactly what lgx_gather computes, with the workspace tables held in a
 * static device array of LGX_GATHER_SCRATCH_BYTES per GPU (kept until the
 * process ends) instead of caller memory.  Rejects parameters whose tables
 * exceed it with LGX_EUNSUPPORTED (only large FP64 table sets do), and the
 * swizzling launch shapes (their level copies would not fit).  The
 * first call with a given set of table parameters (extent, dirs, max_steps,
 * pattern, footprint, FP64 or not) on a device uploads the tables
 * synchronously on `stream`; later calls with the same set launch only the
 * kernel.  Calls that overlap in time on one GPU with different table
 * parameters race on the shared array. */
#define LGX_GATHER_SCRATCH_BYTES (512 * 1024)
int lgx_gather_noalloc(const lgx_gather_params *p, const lgx_gbuffer *g, const lgx_pyramid *pyr, lgx_image *out,
                       lgx_stream_t stream);

/* MipBuilder::BuildMips (MipBuilder.h:142-181) + mipLevelBuilder.frag.
 * levels[0] is the source; levels[1..n_levels-1] are written, each
 * floor(prev/2) in both extents.  AVG: mean of the 2x2 children in the
 * order (0,0),(1,0),(0,1),(1,1); MAX: max of them (R32F only). Output is
 * re-rounded to the level's storage format (FP16 round-to-nearest-even). */
int lgx_mip_build(const lgx_image *levels, int32_t n_levels, int32_t mode, lgx_stream_t stream);

/* BlurBuilder::ApplyBlur (BlurBuilder.h:14-46) + blurLayerBuilder.frag:
 * box mean over [-r, r)^2, clamp to edge; radius 0 copies.  Batched: n
 * (src[i], dst[i], radius[i]) triples in one launch (host arrays of
 * descriptors).  src[i] and dst[i] must have equal extents and formats. */
int lgx_box_blur(const lgx_image *src, const lgx_image *dst, const int32_t *radius, int32_t n, lgx_stream_t stream);

/* InterleaveBuilder::Deinterleave (InterleaveBuilder.h:14-46): dst pixel d
 * = src pixel (d % cell) * grid + d / cell, cell = extent / grid. */
int lgx_deinterleave(const lgx_image *in, lgx_image *out, int32_t gx, int32_t gy, lgx_stream_t stream);
/* InterleaveBuilder::Interleave (InterleaveBuilder.h:48-80, with the
 * interleave program that :60-62 meant to bind): dst pixel i = src pixel
 * (i % grid) * cell + i / grid. */
int lgx_interleave(const lgx_image *in, lgx_image *out, int32_t gx, int32_t gy, lgx_stream_t stream);

typedef struct {
  int32_t radius;       /* taps [-r, r)^2; 0 = passthrough (denoiser.frag:82-86) */
  float depth_sharpness;  /* 10.5 (denoiser.frag:103) */
  float normal_sharpness; /* 20.0 (denoiser.frag:105) */
  int32_t reserved;
} lgx_bilateral_params;

/* Geometry-aware bilateral (DenoiserPass with the weights of
 * denoiser.frag:87-110).  in/out RGBA32F; g supplies depth and normal. */
int lgx_bilateral(const lgx_bilateral_params *p, const lgx_image *in, const lgx_gbuffer *g, lgx_image *out,
                  lgx_stream_t stream);

/* Live DenoiserPass regression filter (denoiser.frag:72-185): per-pixel
 * least squares colour = a*depth + b over the 4x4 taps [-2,2)^2. */
int lgx_regression_denoise(const lgx_image *in, const lgx_gbuffer *g, int32_t radius, lgx_image *out,
                           lgx_stream_t stream);

/* Final gathering (finalGatherer.frag:42-58): direct + indirect * albedo. */
int lgx_composite(const lgx_gbuffer *g, const lgx_image *indirect, lgx_image *out, lgx_stream_t stream);

/* ---- seeded analytic scene (replaces ShadowPass/GBufferPass/LightPass,
 * SSVGIRenderer.h:61-205; the producer is not on the hot path) ---- */
typedef struct {
  float wall_albedo[6][3];
  int32_t n_boxes, n_spheres;
  float box_min[32][3], box_max[32][3], box_albedo[32][3];
  float sph_center[8][3], sph_radius[8], sph_albedo[8][3];
  float light_dir[3]; /* unit, towards the light */
  float light_intensity;
  float cam_pos[3];
  float cam_hor, cam_vert; /* Camera (Scene.h:20-34) */
  float fovy;              /* radians */
  float znear, zfar;
} lgx_scene;

/* Host only: scene from a seed (SplitMix64), boxes <= 32, spheres <= 8. */
int lgx_scene_make(uint64_t seed, int32_t n_boxes, int32_t n_spheres, lgx_scene *out);
/* Host only: re-samples the camera (free position, aim) from pose_seed,
 * keeping the room, boxes, spheres and light. */
int lgx_scene_pose(lgx_scene *s, uint64_t pose_seed);
/* Host only: view and projection matrices for the scene camera:
 * view = inverse(translate(pos) * rotY(hor) * rotX(vert)),
 * proj = perspectiveRH_NO(fovy, w/h, znear, zfar) * scale(1,-1,-1)
 * (SSVGIRenderer.h:54-58). */
int lgx_scene_matrices(const lgx_scene *s, int32_t width, int32_t height, float view[16], float proj[16]);
/* Render the G-buffer.  *_cpu: host pointers (the g images' ptr fields are
 * host memory), computed on host threads; *_gpu: device pointers, async. */
int lgx_scene_render_cpu(const lgx_scene *s, const lgx_gbuffer *g);
int lgx_scene_render_gpu(const lgx_scene *s, const lgx_gbuffer *g, lgx_stream_t stream);

/* ---- the SSVGI renderer on the C++ RenderGraph (include/lgx/ssvgi.h) ----
 * Replaces SSVGIRenderer's GI passes (SSVGIRenderer.h:207-342): the mips and
 * Hi-Z, per-mip pre-blur, IndirectLightPass, interleave, DenoiserPass and
 * the final gathering, declared per frame through RenderGraph::AddPass and
 * executed on the given stream.  Creation allocates every viewport resource
 * (RecreateSwapchainResources, SSVGIRenderer.h:24-29); frames allocate
 * nothing.  The G-buffer is caller memory, rebound each frame. */
typedef struct lgx_renderer lgx_renderer;
enum { LGX_DENOISER_NONE = 0, LGX_DENOISER_BILATERAL = 1, LGX_DENOISER_REGRESSION = 2, LGX_DENOISER_UNET = 3 };
typedef struct {
  int32_t width, height;
  int32_t dirs, max_steps, levels, grid, sampler;
  int32_t denoiser, denoise_passes, denoise_radius;
  float depth_sharpness, normal_sharpness;
  int32_t normal_format, albedo_format, radiance_format;
  int32_t tile;
  int32_t composite; /* 1: add the final gathering pass */
  int32_t profile;   /* 1: per-pass hipEvent timings */
  int32_t validate;  /* 1: passes may only touch the views and buffers they declare */
  int32_t footprint_dirs; /* gather step growth / lod direction count (lgx_gather_params), <= 0 = dirs */
  int32_t grid_y;    /* interleave pattern height (lgx_gather_params), <= 0 = grid.  A pattern that does not
                        divide the frame has the gather write the image directly (no InterleavePass). */
} lgx_renderer_config;

int lgx_renderer_create(const lgx_renderer_config *cfg, lgx_renderer **out);
/* Declares and runs one frame. */
int lgx_renderer_render(lgx_renderer *r, const lgx_gbuffer *g, const float view[16], const float proj[16],
                        lgx_stream_t stream);
/* Declares one frame and captures it as a hipGraph (G-buffer pointers and
 * camera are baked in); lgx_renderer_replay relaunches it. */
int lgx_renderer_capture(lgx_renderer *r, const lgx_gbuffer *g, const float view[16], const float proj[16],
                         lgx_stream_t stream);
int lgx_renderer_replay(lgx_renderer *r, lgx_stream_t stream);
/* LGX_DENOISER_UNET: the network the DenoiserPass runs (created with max n >= 1
 * and the viewport extent; owned by the caller, must outlive the renderer). */
struct lgx_unet;
int lgx_renderer_attach_unet(lgx_renderer *r, struct lgx_unet *u);
/* Descriptor of a named intermediate ("depth_pyramid", "radiance_pyramid",
 * "blurred_radiance", "blurred_depth", "indirect_slices", "indirect",
 * "denoised", "final") at mip level `mip`. */
int lgx_renderer_intermediate(lgx_renderer *r, const char *name, int32_t mip, lgx_image *out);
/* Per-pass GPU times (hipEvents on the frame's stream, needs profile=1) as
 * JSON [{"name":..,"ms":..},..]; waits for the frames concerned.
 * mode 0: the last frame; 1: mean per pass over the frames since the last
 * reset; 2: reset the accumulation (buf receives ""). */
int lgx_renderer_pass_times(lgx_renderer *r, int32_t mode, char *buf, size_t cap);
const char *lgx_renderer_last_error(const lgx_renderer *r);
void lgx_renderer_destroy(lgx_renderer *r);

/* ---- learned denoiser: BF16 U-Net on the matrix cores ----
 * Replaces the regression DenoiserPass (SSVGIRenderer.h:265-302,
 * denoiser.frag:112-184).  Three levels, 3x3 convolutions with 32/64/128
 * channels, ReLU, average-pool down / nearest-upsample up with concatenated
 * skips, 1x1 head predicting a correction of the noisy irradiance.  BF16
 * activations and weights, FP32 accumulation, FP32 master weights + Adam.
 * Input record: 12 FP32 per pixel = noisy irradiance rgb, radial depth,
 * octahedral normal xy, albedo rgb, 3 unused.  Images are [n][h][w] with h
 * and w multiples of 4.  Create allocates every buffer for the largest
 * n x h x w; the other calls allocate nothing. */
typedef struct lgx_unet lgx_unet;
int64_t lgx_unet_param_count(void);
int lgx_unet_create(int32_t max_n, int32_t max_h, int32_t max_w, uint64_t seed, lgx_unet **out);
/* out: device RGBA32F [n][h][w] (alpha 1). */
int lgx_unet_forward(lgx_unet *u, const float *x12, int32_t n, int32_t h, int32_t w, float *out, lgx_stream_t stream);
/* Forward + loss backward against target (device RGBA32F [n][h][w]);
 * gradients kept inside; loss_out (device float, optional) = the mean over
 * rgb of |out - target| (LGX_UNET_LOSS_L1, the default) or of
 * (out - target)^2 (LGX_UNET_LOSS_L2, see lgx_unet_set_loss). */
int lgx_unet_backward(lgx_unet *u, const float *x12, const float *target, int32_t n, int32_t h, int32_t w,
                      float *loss_out, lgx_stream_t stream);
int lgx_unet_adam_step(lgx_unet *u, float lr, lgx_stream_t stream);
/* Training loss of later backward calls: mean absolute error (the default)
 * or mean squared error, whose minimiser is the conditional mean that a
 * relative-L2 metric scores. */
#define LGX_UNET_LOSS_L1 0
#define LGX_UNET_LOSS_L2 1
int lgx_unet_set_loss(lgx_unet *u, int32_t loss);
int lgx_unet_train_step(lgx_unet *u, const float *x12, const float *target, int32_t n, int32_t h, int32_t w, float lr,
                        float *loss_out, lgx_stream_t stream);
/* Device pointer of the last forward's output inside the network (RGBA32F). */
int lgx_unet_output(lgx_unet *u, float **out);
/* Host copies of the state, synchronous. which: 0 parameters, 1 gradients
 * (get only), 2 Adam first moment, 3 Adam second moment; flat FP32 in the
 * layout [conv l: W[co][ci][3][3], b[co]] x 10, head W[3][32], b[3]. */
int lgx_unet_get_state(lgx_unet *u, int32_t which, float *host, int64_t n, int32_t *adam_t);
int lgx_unet_set_state(lgx_unet *u, int32_t which, const float *host, int64_t n, int32_t adam_t);
/* Inference from a G-buffer and the noisy irradiance (RGBA32F); out RGBA32F (any pitch). */
int lgx_unet_denoise(lgx_unet *u, const lgx_gbuffer *g, const lgx_image *noisy, lgx_image *out, lgx_stream_t stream);
/* Times the register tiles of every convolution at n x h x w (forward, and
 * the input gradients when backward != 0) on the network's own buffers,
 * keeps the fastest per layer and writes the table as JSON to buf.  The
 * network's activations are overwritten. */
int lgx_unet_autotune(lgx_unet *u, int32_t n, int32_t h, int32_t w, int32_t backward, lgx_stream_t stream,
                      char *buf, size_t cap);
/* Register tiles offered for a convolution with cout output channels, and
 * the tile index of each of the 10 convolutions (forward fwd[l], input
 * gradient dx[l]; dx[0] is unused); every tile gives bit-identical results. */
int lgx_unet_conv_tiles(int32_t cout);
int lgx_unet_set_tiles(lgx_unet *u, const int32_t *fwd, const int32_t *dx);
int lgx_unet_get_tiles(const lgx_unet *u, int32_t *fwd, int32_t *dx);
/* One wave computes D(16x16) = A(16x32, row-major) * B(32x16, row-major) on the
 * matrix core through the operand layout the convolutions use (device FP32
 * pointers; operands rounded to BF16). */
int lgx_unet_mfma_probe(const float *A, const float *B, float *D, lgx_stream_t stream);
void lgx_unet_destroy(lgx_unet *u);

/* Pitch-aware device-to-device copy of equal-extent, equal-format images (async). */
int lgx_copy_image(const lgx_image *src, lgx_image *dst, lgx_stream_t stream);

/* Number of gather launch shapes (lgx_gather_params.tile indexes them). */
int lgx_gather_tiles(void);
/* A launch shape: info[0..4] = slice pixels per workgroup in x and y, waves
 * per SIMD the registers are held to (1 = the compiler's choice), pipelined
 * step loop (0/1), swizzled levels (0/1).  A swizzling shape re-lays the
 * pyramid levels whose texel columns the lanes of a wave share modulo
 * F = grid >> level (F >= 2, power-of-two grid) into column-swizzled copies
 * before the gather (FP32 Hi-Z gathers; step counting and the other
 * samplers and precisions read the levels as they are).  All shapes give
 * identical results. */
int lgx_gather_tile_info(int32_t tile, int32_t *info);

/* ---- several views of uneven resolution in one persistent launch ----
 * No counterpart in the reference (its frame loop, main.cpp:269-312, renders
 * one view per frame); built for serving many views per submission.
 *
 * lgx_gather_batch: the IndirectLightPass of n views as ONE launch of a
 * persistent kernel whose workgroups walk the views' tiles.  Each item is
 * what lgx_gather takes, checked the same way; all share LGX_PREC_F32, the
 * sampler, the tile and the radiance / normal formats (RGBA16F + RG16F or
 * RGBA32F + RG32F); step_counts must be NULL; at most 2978 views (the
 * swizzled copies' re-layout launch has one grid row per view level).  The
 * descriptors and cameras are uploaded at creation (synchronous); the images
 * must stay allocated. */
typedef struct lgx_gather_batch lgx_gather_batch;
int lgx_gather_batch_create(const lgx_gather_params *p, const void *const *workspaces, const lgx_gbuffer *g,
                            const lgx_pyramid *pyr, lgx_image *out, int32_t n, lgx_gather_batch **res);
/* Launch shape for every view (synchronous upload; not during a capture).
 * The first swizzling shape allocates the views' swizzled copies (held by
 * the batch until it is destroyed); the workspaces only hold tables. */
int lgx_gather_batch_set_tile(lgx_gather_batch *b, int32_t tile);
/* Logical blocks (tiles x slices) of all views together. */
int64_t lgx_gather_batch_blocks(const lgx_gather_batch *b);
/* workgroups: persistent workgroups (rounded up to a multiple of 8);
 * 0 = one per logical block.  Workgroups take blocks from per-XCD counters
 * that the launch zeroes first (with a kernel: capturable), so launches of
 * ONE batch must not overlap in time (different batches may). */
int lgx_gather_batch_launch(const lgx_gather_batch *b, int32_t workgroups, lgx_stream_t stream);
void lgx_gather_batch_destroy(lgx_gather_batch *b);

/* lgx_batch: the denoised path for n views (own G-buffers and cameras).
 * The views, in ascending pixel count, form `chunks` chunks; per chunk the
 * mips, Hi-Z and pre-blur of its views and ONE persistent lgx_gather_batch
 * launch over them, then the U-Net of each of its views, smallest first,
 * and a device timestamp as each view completes, all on the caller's stream.  The whole batch can be
 * captured into ONE hipGraph and replayed. */
typedef struct lgx_batch lgx_batch;
typedef struct {
  int32_t width, height;
  lgx_gbuffer gbuffer; /* caller's device images, width x height */
  float view[16], proj[16];
} lgx_batch_view;
typedef struct {
  int32_t dirs, max_steps, levels, grid, sampler, footprint_dirs;
  int32_t tile;       /* gather launch shape */
  int32_t workgroups; /* persistent gather workgroups, 0 = one per block */
  int32_t grid_y;     /* interleave pattern height (lgx_gather_params), <= 0 = grid */
  int32_t chunks;     /* views, in ascending pixel count, are gathered in this many persistent launches
                         (<= 0: 4); each chunk's U-Nets run before the next chunk's gather */
} lgx_batch_config;
/* net: created with max n >= 1 and max h x w covering every view; NULL =
 * no denoiser (the views' outputs are the noisy gathers). */
int lgx_batch_create(const lgx_batch_config *c, const lgx_batch_view *views, int32_t n, struct lgx_unet *net,
                     lgx_batch **out);
/* Times every (tile, workgroups) pair on the batch's own gather launch,
 * keeps the fastest and writes a JSON table to buf. */
int lgx_batch_autotune(lgx_batch *b, lgx_stream_t stream, char *buf, size_t cap);
int lgx_batch_run(lgx_batch *b, lgx_stream_t stream);
int lgx_batch_capture(lgx_batch *b, lgx_stream_t stream); /* a non-default stream */
int lgx_batch_replay(lgx_batch *b, lgx_stream_t stream);
/* Per view: ms from the batch's start to that view's completion in the last
 * run or replay (synchronises the device). */
int lgx_batch_latencies(lgx_batch *b, double *ms, int32_t n);
/* which: 0 = noisy gather, 1 = denoised (RGBA32F, width x height). */
int lgx_batch_output(lgx_batch *b, int32_t i, int32_t which, lgx_image *out);
/* The launch shape in use: tile and workgroups. */
int lgx_batch_shape(const lgx_batch *b, int32_t *tile, int32_t *workgroups);
const char *lgx_batch_last_error(const lgx_batch *b);
void lgx_batch_destroy(lgx_batch *b);

const char *lgx_strerror(int code);
int32_t lgx_version(void);

#ifdef __cplusplus
}
#endif
#endif
