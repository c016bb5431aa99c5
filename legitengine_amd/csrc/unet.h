// U-Net denoiser: kernel arguments and launchers (unet_kernels.hip), used by
// the layer graph in unet.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "lgx.h"

namespace lgx {
namespace unet {

typedef __bf16 bf16;

struct ConvArgs {
  const bf16 *a;  // input A, NHWC, ca channels (half resolution when a_up)
  int ca, a_up;
  const bf16 *b;  // input B, NHWC, cb channels (concatenated after A), or null
  int cb;
  int n, h, w;    // output extent
  const bf16 *wt; // packed [9][(ca+cb)/32][cout][32]
  const float *bias;
  int cout, relu;
  bf16 *out0;     // channels [0, split)
  int split;
  bf16 *out1;     // channels [split, cout)
  const bf16 *mask0, *mask1;
  int tile;       // register tile: an index below conv3x3_tiles(cout), 0 = default
  // Inference head fused into the epilogue (cout = 32, an LDS tile): instead
  // of storing the activation, out = noisy + head_b + head_w * act per pixel.
  const float *head_w, *head_b;  // [3][32], [3]
  const float4 *head_noisy;      // residual, row pitch head_noisy_pitch
  float4 *head_out;              // row pitch head_out_pitch; null = no head
  int head_noisy_pitch, head_out_pitch;
  // First convolution at inference (cout = 32, an LDS tile, n = 1): the
  // input records are packed from the noisy image and the G-buffer while
  // the tile is staged, channels (noisy rgb, depth, normal xy, albedo rgb,
  // zeros), the values pack_input gives for those records; a and b are
  // unused.
  int gin_fmt;  // 0: off; 1: RG16F normal + RGBA8 albedo; 2: RG32F + RGBA32F; 3: records (conv_first_pair)
  const float4 *gin_noisy;
  const float *gin_depth;
  const void *gin_normal, *gin_albedo;
  int gin_noisy_pitch, gin_depth_pitch, gin_normal_pitch, gin_albedo_pitch;
  const float *gin_x12;  // conv_first_pair with gin_fmt 3: the input records [n][h][w][12]
  bf16 *rec_out;         // conv_first_pair: optional, the packed 32-channel records of every pixel
  // Also write the 2x2 average pool of the output (NHWC, h/2 x w/2, cout
  // channels), summed as avgpool_kernel sums it; null = off.  Needs an LDS
  // tile whose waves hold an even number of rows (conv3x3_tile_pool).
  bf16 *pool_out;
  // Input gradient of a layer whose input A is a nearest x2 upsample: for
  // output channels [0, split) store, instead of the full-resolution
  // gradient, its 2x2 sums at half resolution masked by up_mask > 0 (the
  // gradient of the upsampled tensor, as upsample_bwd_kernel forms it from
  // the stored values); [split, cout) are stored as usual.  Same tiles as
  // pool_out; null = off.
  bf16 *up_out;
  const bf16 *up_mask;  // half resolution, split channels
};

struct WgradArgs {
  const bf16 *a;
  int ca, a_up;
  const bf16 *b;
  int cb;
  const bf16 *dy;  // NHWC cout
  int cout;
  int n, h, w;
  int rows;        // rows per workgroup (set by conv3x3_wgrad)
  int cin_log;     // input channels of the master weights (the rest of ca + cb is zero padding)
  float *dw;       // [cout][cin_log][9] (master layout), overwritten
  float *db;       // [cout] bias gradient, overwritten
  float *part;     // scratch for the per-workgroup partial sums
  float *part_b;   // (set by conv3x3_wgrad) bias partials inside part
  size_t part_bytes;
};

hipError_t conv3x3(const ConvArgs &a, hipStream_t s);
// The first convolution from the G-buffer (a0: gin_* set, as for conv3x3,
// n = 1) or from input records (gin_fmt 3, gin_x12, any n) and the second
// one with its 2x2 pool (a1: out0, pool_out) in one kernel; the first
// layer's output stays in LDS and also goes to a0.out0 when set (the
// training forward keeps it for the backward pass).
hipError_t conv_first_pair(const ConvArgs &a0, const ConvArgs &a1, hipStream_t s);
int conv3x3_tiles(int cout);  // register tiles offered for this output channel count
int conv3x3_tile_lds(int cout, int tile);  // 1 if that tile is an LDS tile covering all cout channels
int conv3x3_tile_pool(int cout, int tile);  // 1 if that tile can write the 2x2 pooled output (pool_out)
hipError_t conv3x3_wgrad(WgradArgs a, hipStream_t s);
size_t conv3x3_wgrad_scratch_bytes();
hipError_t pack_input(const float *x12, size_t npix, bf16 *x32, hipStream_t s);
hipError_t head_fwd_img(const bf16 *d1, const lgx_image &noisy, const float *wh, const float *bh, int out_pitch,
                        float4 *out, hipStream_t s);
hipError_t avgpool(const bf16 *x, int n, int h, int w, int c, bf16 *y, hipStream_t s);
hipError_t pool_bwd_add(const bf16 *dx_in, const bf16 *dpool, const bf16 *act, int n, int h, int w, int c, bf16 *dx,
                        hipStream_t s);
hipError_t upsample_bwd(const bf16 *dfull, const bf16 *act, int n, int hh, int ww, int c, bf16 *dhalf, hipStream_t s);
// out row pitch in pixels: rows of w pixels (over all n images) start every out_pitch
hipError_t head_fwd(const bf16 *d1, const float *x12, const float *wh, const float *bh, size_t npix, int w,
                    int out_pitch, float4 *out, hipStream_t s);
// the head's forward, the loss and the head's backward in one pass (training;
// the head's output is not stored)
hipError_t head_train(const bf16 *d1, const float *x12, const float4 *target, const float *wh, const float *bh,
                      size_t npix, float inv_count, int l2, bf16 *dd1, float *dwh, float *dbh, float *loss,
                      hipStream_t s);
// The convolutions' weights inside the flat parameter vector and their BF16
// packed copies (pack_weights' layouts), for adam_pack.
struct PackLayer {
  size_t off;  // first weight [co][ci][tap] in the parameter vector
  int cout, cin_log, cin_pad;
  bf16 *wf, *wt;  // wt may be null
};
struct PackTable {
  PackLayer l[10];
  int n;
};
// Adam (bias-corrected) on the FP32 parameters; the updated convolution
// weights are also written to their packed BF16 copies.
hipError_t adam_pack(float *p, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
                     float bc1, float bc2, const PackTable &pk, hipStream_t s);
hipError_t fill(float *p, size_t n, float v, hipStream_t s);
hipError_t pack_weights(const float *w, int cout, int cin_log, int cin_pad, bf16 *wf, bf16 *wt, hipStream_t s);
hipError_t mfma_probe(const float *A, const float *B, float *D, hipStream_t s);

}  // namespace unet
}  // namespace lgx
