// U-Net denoiser kernels for gfx950: BF16 activations and weights, FP32
// accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Replaces the reference's regression DenoiserPass (SSVGIRenderer.h:265-302,
// denoiser.frag:112-184) with a learned filter; the layer graph lives in
// unet.cpp.
//
// Layouts: activations NHWC bf16 with channel counts that are multiples of
// 32 (convolution inputs) or 16 (outputs).  Convolution weights are packed
// per K-chunk: [tap 9][Cin/32][Cout][32], so one lane's MFMA operand (8
// consecutive input channels of one output channel) is a single 16-byte load.
//
// MFMA 16x16x32 operand layout (gfx950, wave64, lane l, item i of 8 / 4):
//   A (16 x 32, row m, col k):  m = l % 16, k = 8 * (l / 16) + i
//   B (32 x 16, row k, col n):  n = l % 16, k = 8 * (l / 16) + i
//   D (16 x 16):                n = l % 16, m = 4 * (l / 16) + i
// The convolution maps M = output channels, N = pixels, K = (tap, Cin):
// A is read from the packed weights, B straight from the NHWC activations
// (16 bytes = 8 channels of one pixel per lane), and each lane's result is 4
// consecutive output channels of one pixel (one 8-byte store).
#include <math.h>

#include <algorithm>

#include "lgx_internal.h"
#include "unet.h"

namespace lgx {
namespace unet {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

LGX_D float bf2f(bf16 v) { return (float)v; }
LGX_D bf16 f2bf(float v) { return (bf16)v; }

// ------------------------------------------------------------ convolution

// Input addressing shared by forward, input-gradient and weight-gradient
// kernels: channels [0, ca) come from A (optionally nearest-upsampled from
// half resolution), [ca, ca+cb) from B.
struct ConvInput {
  const bf16 *a;
  int ca, a_up;
  const bf16 *b;
  int cb;
};

// 16 bytes (8 channels starting at c, c % 8 == 0, within one 32-chunk) of
// pixel (n, y, x) of the logical input; zero outside the image.
LGX_D bf16x8 load8(const ConvInput &in, int n, int y, int x, int h, int w, int c) {
  bf16x8 z = {};
  if (y < 0 || y >= h || x < 0 || x >= w) return z;
  if (c < in.ca) {
    if (in.a_up) {
      const int hh = h >> 1, ww = w >> 1;
      return *(const bf16x8 *)(in.a + (((size_t)n * hh + (y >> 1)) * ww + (x >> 1)) * in.ca + c);
    }
    return *(const bf16x8 *)(in.a + (((size_t)n * h + y) * w + x) * in.ca + c);
  }
  return *(const bf16x8 *)(in.b + (((size_t)n * h + y) * w + x) * in.cb + (c - in.ca));
}

// The same 16 bytes for a pixel clamped into the image (always a valid
// read; the caller zeroes what lies outside).
LGX_D bf16x8 load8_clamped(const ConvInput &in, int n, int y, int x, int h, int w, int c) {
  y = min(max(y, 0), h - 1);
  x = min(max(x, 0), w - 1);
  if (c < in.ca) {
    if (in.a_up) {
      const int hh = h >> 1, ww = w >> 1;
      return *(const bf16x8 *)(in.a + (((size_t)n * hh + (y >> 1)) * ww + (x >> 1)) * in.ca + c);
    }
    return *(const bf16x8 *)(in.a + (((size_t)n * h + y) * w + x) * in.ca + c);
  }
  return *(const bf16x8 *)(in.b + (((size_t)n * h + y) * w + x) * in.cb + (c - in.ca));
}

// Stores one output row of a wave's register tile: PB blocks of 16 pixels
// (x0 + pb * 16 + col) x MB blocks of 16 channels (co0 + mb * 16 + kg * 4,
// 4 per lane): bias (bias_r, this lane's channels), ReLU, the optional
// mask (four channels in one read), the split into out0 / out1.
template <int PB, int MB, bool UP = false>
LGX_D void conv_store_row(const ConvArgs &a, const f32x4 (&acc)[PB][MB], const float (&bias_r)[MB][4], int n, int y,
                          int x0, int col, int kg, int co0) {
  typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
#pragma unroll
  for (int pb = 0; pb < PB; pb++) {
    const int x = x0 + pb * 16 + col;
    if (x >= a.w) continue;
    const size_t pix = ((size_t)n * a.h + y) * a.w + x;
#pragma unroll
    for (int mb = 0; mb < MB; mb++) {
      const int co = co0 + mb * 16 + kg * 4;
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        v[i] = acc[pb][mb][i];
        if (a.bias) v[i] += bias_r[mb][i];
        if (a.relu) v[i] = v[i] > 0.f ? v[i] : 0.f;
      }
      // split is a multiple of 16 (checked on the host), so the half is the
      // same for the whole 16-channel block: a uniform choice of pointers
      // (a per-lane choice would become a per-lane read of the arguments)
      const bool lo = co0 + mb * 16 < a.split;
      if (UP && lo) continue;  // stored as 2x2 sums at half resolution
      bf16 *const dst = lo ? a.out0 : a.out1;
      const bf16 *const mask = lo ? a.mask0 : a.mask1;
      const int stride = lo ? a.split : a.cout - a.split, cch = lo ? co : co - a.split;
      if (mask) {
        const bf16x4 m = *(const bf16x4 *)(mask + pix * stride + cch);
#pragma unroll
        for (int i = 0; i < 4; i++)
          if (!(bf2f(m[i]) > 0.f)) v[i] = 0.f;
      }
      bf16x4 o = {f2bf(v[0]), f2bf(v[1]), f2bf(v[2]), f2bf(v[3])};
      *(bf16x4 *)(dst + pix * stride + cch) = o;
    }
  }
}

// Forward 3x3 convolution (stride 1, zero padding 1) + bias + ReLU, and the
// same kernel with flipped/transposed weights for the input gradient.
// A wave computes R output rows x PB blocks of 16 pixels x all output
// channels (MB blocks of 16); the 4 waves of a workgroup sit side by side
// along x and share the weight fragments through L1.  Per 32-channel input
// chunk and horizontal tap offset dx, the wave loads the R + 2 input rows it
// needs once and feeds each of them to the three vertical taps, so an
// activation fragment serves up to 3 * MB MFMAs and a weight fragment R * PB.
// Output channels [0, split) go to out0, [split, cout) to out1 (the two
// halves of a concatenated input when computing its gradient); an optional
// mask zeroes outputs whose mask activation is <= 0 (the ReLU derivative of
// the tensor the gradient flows into).
template <int MB, int PB, int R> __global__ __launch_bounds__(256) void conv3x3_kernel(const ConvArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = blockIdx.z, y0 = blockIdx.y * R, x0 = (blockIdx.x * 4 + wave) * (PB * 16);
  // the bias, in LDS for the epilogue (a global read placed after the
  // epilogue's stores would wait for their round trip: they might alias it)
  __shared__ __attribute__((aligned(16))) float s_bias[MB * 16];
  if ((int)threadIdx.x < MB * 16) s_bias[threadIdx.x] = a.bias ? a.bias[threadIdx.x] : 0.f;
  __syncthreads();
  if (x0 >= a.w) return;
  const int col = lane & 15, kg = lane >> 4;
  const int cin = a.ca + a.cb, chunks = cin >> 5;
  f32x4 acc[R][PB][MB];
#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int pb = 0; pb < PB; pb++)
#pragma unroll
      for (int mb = 0; mb < MB; mb++) acc[r][pb][mb] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int cc = 0; cc < chunks; cc++) {
    // the chunk's source is uniform: A (optionally at half resolution) or B
      const bool from_a = cc * 32 < a.ca;
      const bool up = from_a && a.a_up;
    const int cs = from_a ? a.ca : a.cb;
    const int sw = up ? (a.w >> 1) : a.w, sh = up ? (a.h >> 1) : a.h;
    const bf16 *img = (from_a ? a.a : a.b) + (size_t)n * sh * sw * cs + (from_a ? cc * 32 : cc * 32 - a.ca) + kg * 8;
    const bf16 *wchunk = a.wt + ((size_t)cc * a.cout + col) * 32 + kg * 8;
    const size_t wtap = (size_t)chunks * a.cout * 32;
#pragma unroll 1
    for (int dx = -1; dx <= 1; dx++) {
      uint32_t xoff[PB];
      bool xok[PB];
#pragma unroll
      for (int pb = 0; pb < PB; pb++) {
        const int x = x0 + pb * 16 + col + dx;
        xok[pb] = x >= 0 && x < a.w;
        xoff[pb] = (uint32_t)(up ? (x >> 1) : x) * (uint32_t)cs;
      }
      bf16x8 bx[R + 2][PB];
#pragma unroll
      for (int rr = 0; rr < R + 2; rr++) {
        const int yy = y0 - 1 + rr;
        const bool yok = yy >= 0 && yy < a.h;
        const uint32_t roff = (uint32_t)((up ? (yy >> 1) : yy) * sw) * (uint32_t)cs;
#pragma unroll
        for (int pb = 0; pb < PB; pb++) {
          bf16x8 z = {};
          bx[rr][pb] = (yok && xok[pb]) ? *(const bf16x8 *)(img + roff + xoff[pb]) : z;
        }
      }
#pragma unroll
      for (int dy = -1; dy <= 1; dy++) {
        const bf16 *wrow = wchunk + (size_t)((dy + 1) * 3 + (dx + 1)) * wtap;
        bf16x8 aw[MB];
#pragma unroll
        for (int mb = 0; mb < MB; mb++) aw[mb] = *(const bf16x8 *)(wrow + (size_t)mb * 16 * 32);
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
          for (int pb = 0; pb < PB; pb++)
#pragma unroll
            for (int mb = 0; mb < MB; mb++)
              acc[r][pb][mb] =
                  __builtin_amdgcn_mfma_f32_16x16x32_bf16(aw[mb], bx[r + 1 + dy][pb], acc[r][pb][mb], 0, 0, 0);
      }
    }
  }
  // Epilogue.  Masks are read into registers before the stores they feed:
  // for all the compiler knows a store may alias them, so a read placed
  // after a store would wait for that store's memory round trip.
  float bias_r[MB][4];
#pragma unroll
  for (int mb = 0; mb < MB; mb++) {
    const float4 b4 = *(const float4 *)(&s_bias[mb * 16 + kg * 4]);
    bias_r[mb][0] = b4.x;
    bias_r[mb][1] = b4.y;
    bias_r[mb][2] = b4.z;
    bias_r[mb][3] = b4.w;
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int y = y0 + r;
    if (y >= a.h) break;
    conv_store_row<PB, MB>(a, acc[r], bias_r, n, y, x0, col, kg, 0);
  }
}

// The same convolution with the operands staged in LDS.  A workgroup
// computes a TH = 4*WR row x TW = 16*WP pixel tile for MB*16 output channels
// (a wide layer's channels are split over `splits` tiles), its 4 waves
// stacked vertically.  Per 32-channel input chunk it stages the
// (TH+2) x (TW+2) pixel halo tile (zeros outside the image) and the chunk's
// weights of all 9 taps in LDS; the MFMAs read only LDS while the global
// loads of the next (tile, chunk) are in flight in registers, so global
// latency overlaps the matrix work also across tiles: a workgroup walks
// tiles blockIdx.x, blockIdx.x + gridDim.x, ...  Every output is
// accumulated in the same order as conv3x3_kernel (chunk, dx, dy), so the
// two kernels give bit-identical results.
// HEAD: the inference head fused into the epilogue (its own instantiation;
// needs MB = 2, i.e. all 32 channels of a pixel in one workgroup).
// With one input chunk and no channel split every tile uses the same
// weights: they stay in LDS from the workgroup's first tile on.
// Raw G-buffer texels of one pixel for the G-buffer-fed first convolution
// (GIN = 1: RG16F normal, RGBA8 albedo; GIN = 2: RG32F, RGBA32F), kept raw
// in registers while in flight and converted when staged.
template <int GIN> struct GinRaw;
template <> struct GinRaw<1> {
  float4 c;
  float d;
  uint32_t n, al;
};
template <> struct GinRaw<2> {
  float4 c;
  float d;
  float2 n;
  float4 al;
};
template <> struct GinRaw<3> {  // the first 9 floats of an input record
  float4 a, b;
  float c;
};
// (GIN = 3 reads image n's records; the G-buffer formats have one image)
template <int GIN> LGX_D GinRaw<GIN> gin_load(const ConvArgs &a, int y, int x, int n = 0) {
  GinRaw<GIN> r;
  if constexpr (GIN == 3) {
    const float *rec = a.gin_x12 + (((size_t)n * a.h + y) * a.w + x) * 12;
    r.a = *(const float4 *)rec;
    r.b = *(const float4 *)(rec + 4);
    r.c = rec[8];
    return r;
  } else {
  r.c = a.gin_noisy[(size_t)y * a.gin_noisy_pitch + x];
  r.d = a.gin_depth[(size_t)y * a.gin_depth_pitch + x];
  if constexpr (GIN == 1) {
    r.n = ((const uint32_t *)a.gin_normal)[(size_t)y * a.gin_normal_pitch + x];
    r.al = ((const uint32_t *)a.gin_albedo)[(size_t)y * a.gin_albedo_pitch + x];
  } else {
    r.n = ((const float2 *)a.gin_normal)[(size_t)y * a.gin_normal_pitch + x];
    r.al = ((const float4 *)a.gin_albedo)[(size_t)y * a.gin_albedo_pitch + x];
  }
  return r;
  }
}
// The first 16 channels of the pixel's input record (pack_input_kernel's
// values for the record the G-buffer and the noisy image give).
template <int GIN> LGX_D void gin_pack(const GinRaw<GIN> &r, bf16x8 &v0, bf16x8 &v1) {
  if constexpr (GIN == 3) {
    v0 = bf16x8{f2bf(r.a.x), f2bf(r.a.y), f2bf(r.a.z), f2bf(r.a.w), f2bf(r.b.x), f2bf(r.b.y), f2bf(r.b.z), f2bf(r.b.w)};
    v1 = bf16x8{};
    v1[0] = f2bf(r.c);
    return;
  } else {
  float nx, ny, ar, ag, ab;
  if constexpr (GIN == 1) {
    nx = h2f((uint16_t)(r.n & 0xffffu));
    ny = h2f((uint16_t)(r.n >> 16));
    const float k = 1.0f / 255.0f;
    ar = (float)(r.al & 0xffu) * k;
    ag = (float)((r.al >> 8) & 0xffu) * k;
    ab = (float)((r.al >> 16) & 0xffu) * k;
  } else {
    nx = r.n.x;
    ny = r.n.y;
    ar = r.al.x;
    ag = r.al.y;
    ab = r.al.z;
  }
  v0 = bf16x8{f2bf(r.c.x), f2bf(r.c.y), f2bf(r.c.z), f2bf(r.d), f2bf(nx), f2bf(ny), f2bf(ar), f2bf(ag)};
  v1 = bf16x8{};
  v1[0] = f2bf(ab);
  }
}

// The value of the neighbouring lane (lane ^ 1): a DPP quad permutation
// [1, 0, 3, 2], a modifier of one VALU move (no LDS round trip as a shuffle
// through ds_bpermute would take).
LGX_D float lane_xor1(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, false));
}

// LDS bytes of conv3x3_lds_kernel<MB, WR, WP> (the layout described in it),
// and the workgroups per CU that LDS allows, at most 3 (168 VGPRs; 4 would
// leave 128, fewer than the tiles' accumulators and operands need): the
// register allocation is held to that occupancy.
constexpr int conv_lds_bytes(int MB, int WR, int WP) {
  return (4 * (((4 * WR + 2) * (16 * WP + 2) * 8 + 127) / 128 * 128) + 4 * 9 * MB * 16 * 8) * 2;
}
constexpr int conv_lds_blocks(int MB, int WR, int WP) {
  return 160 * 1024 / conv_lds_bytes(MB, WR, WP) < 3 ? 160 * 1024 / conv_lds_bytes(MB, WR, WP) : 3;
}

// UP: the epilogue of an input gradient through an upsample (ConvArgs::up_out),
// a separate instantiation so the other convolutions keep their registers.
template <int MB, int WR, int WP, bool HEAD, int GIN = 0, bool UP = false>
__global__ __launch_bounds__(256, conv_lds_blocks(MB, WR, WP)) void conv3x3_lds_kernel(const ConvArgs a, int tiles_x,
                                                                                      int tiles_y, int splits) {
  constexpr int TH = 4 * WR, TW = 16 * WP, IW = TW + 2, IH = TH + 2;
  constexpr int NPIX = IH * IW;                       // input tile pixels (with the halo)
  constexpr int IN_PIECES = (NPIX + 15) / 16 * 16 * 4;  // 16-byte pieces, pixels rounded up to 16
  constexpr int W_PIECES = 9 * MB * 16 * 4;            // of the chunk's weights
  constexpr int IN_PER = (IN_PIECES + 255) / 256, W_PER = (W_PIECES + 255) / 256;
  // LDS layout [channel group kg][pixel or weight row][8 channels], every
  // plane a multiple of 256 bytes (the 64 banks).  A 128-bit LDS access is
  // served 16 lanes at a time, each 8-lane group of a wave with one from
  // the other 16-lane half of its 32 (counters: 2-way conflicts when planes
  // started 16 banks apart, none when 8-lane group g covers bank half
  // g & 1).  Matrix-operand reads (lanes: kg = lane / 16, pixel = lane % 16)
  // therefore put 8-lane groups of different planes on opposite bank halves,
  // and the staging stores use the same rule (see piece()).
  constexpr int PLANE_IN = (NPIX * 8 + 127) / 128 * 128, PLANE_W = 9 * MB * 16 * 8;
  __shared__ __attribute__((aligned(16))) bf16 s_in[4 * PLANE_IN];
  __shared__ __attribute__((aligned(16))) bf16 s_w[4 * PLANE_W];
  static_assert(sizeof(s_in) + sizeof(s_w) == conv_lds_bytes(MB, WR, WP), "LDS layout");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // 16-byte piece q of a tile -> (pixel or weight row, part = channel group):
  // a wave's 64 pieces are 16 consecutive pixels x 4 parts, lane l taking
  // pixel l % 16 of part l / 16 -- the lane pattern of the operand reads
  auto piece = [](int q, int &pix, int &part) {
    pix = ((q >> 6) << 4) + (q & 15);
    part = (q >> 4) & 3;
  };
  const int col = lane & 15, kg = lane >> 4;
  const int chunks = (a.ca + a.cb) >> 5;
  const size_t wtap = (size_t)chunks * a.cout * 32;
  const int T = tiles_x * tiles_y * a.n * splits;
  struct Tile {
    int n, y0, x0, co0;
  };
  // consecutive tile indices: the channel splits of one pixel tile, then x, y, image
  auto tile_of = [&](int t) {
    Tile r;
    r.co0 = (t % splits) * MB * 16;
    int u = t / splits;
    r.x0 = (u % tiles_x) * TW;
    u /= tiles_x;
    r.y0 = (u % tiles_y) * TH;
    r.n = u / tiles_y;
    return r;
  };
  // Loads are unconditional (clamped to an in-image pixel / a valid piece)
  // and padding is applied when staging: a conditional load would make the
  // compiler merge values right after it and wait for the load there,
  // serialising the prefetch with the matrix work.
  bf16x8 pin[IN_PER], pw[W_PER];
  uint32_t valid = 0;  // bit i: piece i of pin lies in the image
  const bool w_resident = !HEAD && chunks == 1 && splits == 1;
  // G-buffer-fed input: one thread per halo pixel, raw texels in flight
  constexpr int GIN_PER = GIN ? (IH * IW + 255) / 256 : 1;
  GinRaw<GIN ? GIN : 1> graw[GIN_PER];
  auto fetch = [&](const Tile &tl, int cc, bool with_w) {
    if constexpr (GIN) {
      valid = 0;
#pragma unroll
      for (int i = 0; i < GIN_PER; i++) {
        const int p = min(tid + i * 256, IH * IW - 1);
        const int y = tl.y0 - 1 + p / IW, x = tl.x0 - 1 + p % IW;
        const bool ok = y >= 0 && y < a.h && x >= 0 && x < a.w;
        graw[i] = gin_load<GIN>(a, min(max(y, 0), a.h - 1), min(max(x, 0), a.w - 1));
        valid |= (uint32_t)ok << i;
      }
    } else {
      const bool from_a = cc * 32 < a.ca;
      const bool up = from_a && a.a_up;
      const int cs = from_a ? a.ca : a.cb;
      const int sw = up ? (a.w >> 1) : a.w, sh = up ? (a.h >> 1) : a.h;
      const bf16 *img = (from_a ? a.a : a.b) + (size_t)tl.n * sh * sw * cs + (from_a ? cc * 32 : cc * 32 - a.ca);
      valid = 0;
#pragma unroll
      for (int i = 0; i < IN_PER; i++) {
        int p, part;
        piece(min(tid + i * 256, IN_PIECES - 1), p, part);
        p = min(p, NPIX - 1);  // a rounding piece: loaded in bounds, never staged
        const int y = tl.y0 - 1 + p / IW, x = tl.x0 - 1 + p % IW;
        const bool ok = y >= 0 && y < a.h && x >= 0 && x < a.w;
        const int yc = min(max(y, 0), a.h - 1), xc = min(max(x, 0), a.w - 1);
        const uint32_t off = (uint32_t)((up ? (yc >> 1) : yc) * sw + (up ? (xc >> 1) : xc)) * (uint32_t)cs + part * 8;
        pin[i] = *(const bf16x8 *)(img + off);
        valid |= (uint32_t)ok << i;
      }
    }
    if (!with_w) return;
    const bf16 *wsrc = a.wt + ((size_t)cc * a.cout + tl.co0) * 32;
#pragma unroll
    for (int i = 0; i < W_PER; i++) {
      int row, part;
      piece(min(tid + i * 256, W_PIECES - 1), row, part);
      const int tap = row / (MB * 16), co = row % (MB * 16);
      pw[i] = *(const bf16x8 *)(wsrc + tap * wtap + (co * 4 + part) * 8);
    }
  };
  int t = blockIdx.x;
  if (t >= T) return;  // the whole workgroup, before any barrier
  Tile cur = tile_of(t);
  fetch(cur, 0, true);
  bool stage_w = true;  // false once resident weights are staged
  if constexpr (GIN) {  // record channels 16-31 are zero on every tile: staged once
    const bf16x8 z = {};
    for (int p = tid; p < 2 * IH * IW; p += 256) *(bf16x8 *)(s_in + (2 + p / (IH * IW)) * PLANE_IN + (p % (IH * IW)) * 8) = z;
  }
  for (;;) {
    f32x4 acc[WR][WP][MB];
#pragma unroll
    for (int r = 0; r < WR; r++)
#pragma unroll
      for (int pb = 0; pb < WP; pb++)
#pragma unroll
        for (int mb = 0; mb < MB; mb++) acc[r][pb][mb] = f32x4{0.f, 0.f, 0.f, 0.f};
    // The tile's bias, read before the matrix work and kept in registers: a
    // global read placed after the tile's stores would wait for their memory
    // round trip (for all the compiler knows they may alias it).  The read is
    // unconditional like the tile loads (a layer without bias reads its
    // weights instead and never adds the values).
    float bias_r[MB][4];
    if constexpr (!HEAD) {  // (the head's epilogue stores to LDS: it reads the bias there)
      const float *bsrc = a.bias ? a.bias : (const float *)a.wt;
#pragma unroll
      for (int mb = 0; mb < MB; mb++) {
        const float4 b4 = *(const float4 *)(bsrc + cur.co0 + mb * 16 + kg * 4);
        bias_r[mb][0] = b4.x;
        bias_r[mb][1] = b4.y;
        bias_r[mb][2] = b4.z;
        bias_r[mb][3] = b4.w;
      }
    }
    for (int cc = 0; cc < chunks; cc++) {
      if constexpr (GIN) {
#pragma unroll
        for (int i = 0; i < GIN_PER; i++) {
          const int p = tid + i * 256;
          bf16x8 v0 = {}, v1 = {};
          if ((valid >> i) & 1u) gin_pack<GIN>(graw[i], v0, v1);
          if (p < IH * IW) {
            *(bf16x8 *)(s_in + p * 8) = v0;
            *(bf16x8 *)(s_in + PLANE_IN + p * 8) = v1;
          }
        }
      } else {
#pragma unroll
        for (int i = 0; i < IN_PER; i++) {
          const int q = tid + i * 256;
          const bf16x8 z = {};
          int p, part;
          piece(q, p, part);
          if (q < IN_PIECES && p < NPIX) *(bf16x8 *)(s_in + part * PLANE_IN + p * 8) = ((valid >> i) & 1u) ? pin[i] : z;
        }
      }
      if (stage_w) {
#pragma unroll
        for (int i = 0; i < W_PER; i++) {
          const int q = tid + i * 256;
          int row, part;
          piece(q, row, part);
          if (q < W_PIECES) *(bf16x8 *)(s_w + part * PLANE_W + row * 8) = pw[i];
        }
        stage_w = !w_resident;
      }
      __syncthreads();
      // in flight during the matrix work below: the next chunk, or the next tile's first
      if (cc + 1 < chunks)
        fetch(cur, cc + 1, true);
      else if (!HEAD && t + (int)gridDim.x < T)  // a head launch has one tile per workgroup
        fetch(tile_of(t + (int)gridDim.x), 0, !w_resident);
#pragma unroll
      for (int dx = 0; dx < 3; dx++) {
        bf16x8 bx[WR + 2][WP];
#pragma unroll
        for (int rr = 0; rr < WR + 2; rr++)
#pragma unroll
          for (int pb = 0; pb < WP; pb++)
            bx[rr][pb] = *(const bf16x8 *)(s_in + kg * PLANE_IN + ((wave * WR + rr) * IW + pb * 16 + col + dx) * 8);
#pragma unroll
        for (int dy = 0; dy < 3; dy++) {
          const int tap = dy * 3 + dx;
          bf16x8 aw[MB];
#pragma unroll
          for (int mb = 0; mb < MB; mb++)
            aw[mb] = *(const bf16x8 *)(s_w + kg * PLANE_W + ((tap * MB + mb) * 16 + col) * 8);
#pragma unroll
          for (int r = 0; r < WR; r++)
#pragma unroll
            for (int pb = 0; pb < WP; pb++)
#pragma unroll
              for (int mb = 0; mb < MB; mb++)
                acc[r][pb][mb] =
                    __builtin_amdgcn_mfma_f32_16x16x32_bf16(aw[mb], bx[r + dy][pb], acc[r][pb][mb], 0, 0, 0);
        }
      }
      __syncthreads();  // the next stash overwrites the tiles
    }
    if constexpr (HEAD) {
      // fused inference head: the tile's activations (BF16, as the unfused
      // path stores them) go to LDS, then one thread per pixel forms
      // noisy + b + W * act in the unfused head's order (bit-identical)
      bf16 *act = s_in;  // [TH][TW][32]; the input tile is no longer read
      float *hw = (float *)s_w;  // head weights [3][32] and bias [3]; the chunk weights are no longer read
      if (tid < 99) hw[tid] = tid < 96 ? a.head_w[tid] : a.head_b[tid - 96];
#pragma unroll
      for (int r = 0; r < WR; r++)
#pragma unroll
        for (int pb = 0; pb < WP; pb++)
#pragma unroll
          for (int mb = 0; mb < MB; mb++) {
            const int co = mb * 16 + kg * 4;
            typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
            bf16x4 o;
#pragma unroll
            for (int i = 0; i < 4; i++) {
              float v = acc[r][pb][mb][i] + a.bias[co + i];
              o[i] = f2bf(v > 0.f ? v : 0.f);
            }
            *(bf16x4 *)(act + ((wave * WR + r) * TW + pb * 16 + col) * 32 + co) = o;
          }
      __syncthreads();
      for (int p = tid; p < TH * TW; p += 256) {
        const int y = cur.y0 + p / TW, x = cur.x0 + p % TW;
        if (y < a.h && x < a.w) {
          const float4 c = a.head_noisy[(size_t)y * a.head_noisy_pitch + x];
          float o[3] = {c.x + hw[96], c.y + hw[97], c.z + hw[98]};
#pragma unroll 1
          for (int q = 0; q < 4; q++) {
            const bf16x8 v = *(const bf16x8 *)(act + p * 32 + q * 8);
#pragma unroll
            for (int i = 0; i < 8; i++) {
              const float f = bf2f(v[i]);
#pragma unroll
              for (int k = 0; k < 3; k++) o[k] += hw[k * 32 + q * 8 + i] * f;
            }
          }
          a.head_out[(size_t)y * a.head_out_pitch + x] = make_float4(o[0], o[1], o[2], 1.f);
        }
      }
      __syncthreads();  // the next tile's stash overwrites the activations
    } else {
    if (a.pool_out) {
      // 2x2 average of the stored (BF16) outputs: rows r, r + 1 are this
      // lane's, columns x, x + 1 are lanes col, col + 1 (same channel group);
      // summed in avgpool_kernel's order (y, x), (y, x+1), (y+1, x), (y+1, x+1)
      const int hh = a.h >> 1, ww = a.w >> 1;
#pragma unroll
      for (int r = 0; r + 1 < WR; r += 2) {
        const int y = cur.y0 + wave * WR + r;  // even
#pragma unroll
        for (int pb = 0; pb < WP; pb++) {
          const int x = cur.x0 + pb * 16 + col;
#pragma unroll
          for (int mb = 0; mb < MB; mb++) {
            const int co = cur.co0 + mb * 16 + kg * 4;
            float t[4], b[4], tn[4], bn[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
              float v0 = acc[r][pb][mb][i], v1 = acc[r + 1][pb][mb][i];
              if (a.bias) {
                v0 += bias_r[mb][i];
                v1 += bias_r[mb][i];
              }
              if (a.relu) {
                v0 = v0 > 0.f ? v0 : 0.f;
                v1 = v1 > 0.f ? v1 : 0.f;
              }
              t[i] = bf2f(f2bf(v0));
              b[i] = bf2f(f2bf(v1));
              tn[i] = lane_xor1(t[i]);
              bn[i] = lane_xor1(b[i]);
            }
            if (!(col & 1) && (y >> 1) < hh && (x >> 1) < ww) {
              typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
              bf16x4 o;
#pragma unroll
              for (int i = 0; i < 4; i++) o[i] = f2bf((((t[i] + tn[i]) + b[i]) + bn[i]) * 0.25f);
              *(bf16x4 *)(a.pool_out + (((size_t)cur.n * hh + (y >> 1)) * ww + (x >> 1)) * a.cout + co) = o;
            }
          }
        }
      }
    }
    if constexpr (UP) {
      // channels [0, split): 2x2 sums of the stored (BF16) values in
      // upsample_bwd_kernel's order (y, x), (y, x+1), (y+1, x), (y+1, x+1),
      // masked at half resolution; rows r, r + 1 are this lane's, columns
      // x, x + 1 are lanes col, col + 1
      const int hh = a.h >> 1, ww = a.w >> 1;
#pragma unroll
      for (int r = 0; r + 1 < WR; r += 2) {
        const int y = cur.y0 + wave * WR + r;  // even
#pragma unroll
        for (int pb = 0; pb < WP; pb++) {
          const int x = cur.x0 + pb * 16 + col;
#pragma unroll
          for (int mb = 0; mb < MB; mb++) {
            if (!(cur.co0 + mb * 16 < a.split)) continue;  // uniform
            const int co = cur.co0 + mb * 16 + kg * 4;
            float t[4], b[4], tn[4], bn[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
              float v0 = acc[r][pb][mb][i], v1 = acc[r + 1][pb][mb][i];
              if (a.bias) {
                v0 += bias_r[mb][i];
                v1 += bias_r[mb][i];
              }
              if (a.relu) {
                v0 = v0 > 0.f ? v0 : 0.f;
                v1 = v1 > 0.f ? v1 : 0.f;
              }
              t[i] = bf2f(f2bf(v0));
              b[i] = bf2f(f2bf(v1));
              tn[i] = lane_xor1(t[i]);
              bn[i] = lane_xor1(b[i]);
            }
            if (!(col & 1) && (y >> 1) < hh && (x >> 1) < ww) {
              typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
              const size_t q = (((size_t)cur.n * hh + (y >> 1)) * ww + (x >> 1)) * a.split + co;
              const bf16x4 m = *(const bf16x4 *)(a.up_mask + q);
              bf16x4 o;
#pragma unroll
              for (int i = 0; i < 4; i++) o[i] = bf2f(m[i]) > 0.f ? f2bf(((t[i] + tn[i]) + b[i]) + bn[i]) : f2bf(0.f);
              *(bf16x4 *)(a.up_out + q) = o;
            }
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < WR; r++) {
      const int y = cur.y0 + wave * WR + r;
      if (y >= a.h) break;
      conv_store_row<WP, MB, UP>(a, acc[r], bias_r, cur.n, y, cur.x0, col, kg, cur.co0);
    }
    }
    t += gridDim.x;
    if (t >= T) break;
    cur = tile_of(t);
  }
}

// The first two convolutions in one kernel (inference and the training
// forward).  A workgroup's tile is 8 x 32 pixels of the second convolution's
// output.  Its input, the first convolution's output over the tile and a
// 1-pixel halo (10 x 34), is computed into LDS from the input records of the
// tile and a 2-pixel halo (12 x 36, packed from the G-buffer or the FP32
// records while staging, as pack_input_kernel packs them) and goes to HBM
// only when the backward pass needs it: about half of the two layers' HBM
// traffic, for 1.33x the first layer's matrix work (the halo).  The first
// layer's records have 9 channels, padded to 16, so one tap per 32-deep MFMA
// would leave half of K zero; here taps 2p and 2p + 1 share an MFMA
// (k = 16 * (tap & 1) + channel; the tenth tap is zero): 5 MFMAs per output
// block instead of 9.  (conv3x3 on the packed records, one tap per MFMA,
// rounds the partial sums differently.)  The second convolution is
// conv3x3_lds_kernel<2, 2, 2> on that LDS tile: the same operand order and
// the same epilogue, with the 2x2 pool.  Persistent: a workgroup walks the
// tiles t, t + gridDim.x, ...; the next tile's raw texels are in flight
// during the current tile's matrix work.
template <int GIN>
__global__ __launch_bounds__(256, 2) void conv_first_pair_kernel(const ConvArgs a0, const ConvArgs a1, int tiles_x,
                                                                 int tiles_y) {
  constexpr int MB = 2, WR = 2, WP = 2;
  constexpr int TH = 4 * WR, TW = 16 * WP;  // output tile
  constexpr int IH = TH + 2, IW = TW + 2;   // first-layer outputs (the second layer's input tile)
  constexpr int AH = TH + 4, AW = TW + 4;   // input records
  constexpr int NA = AH * AW, NB = IH * IW;
  constexpr int PLANE_A = NA * 8, PLANE_B = (NB * 8 + 127) / 128 * 128, PLANE_W = 9 * MB * 16 * 8;
  constexpr int A_BLOCKS = (NB + 15) / 16;   // 16-pixel blocks of first-layer outputs
  constexpr int A_PER = (A_BLOCKS + 3) / 4;  // per wave (blocks wave, wave + 4, ...)
  constexpr int A_GROUP = 3;                 // blocks per pass (shared weight registers)
  constexpr int GIN_PER = (NA + 255) / 256;
  constexpr int W_PIECES = 9 * MB * 16 * 4, W_PER = (W_PIECES + 255) / 256;
  static_assert(A_PER % A_GROUP == 0, "block groups");
  __shared__ __attribute__((aligned(16))) bf16 s_a[2 * PLANE_A];  // records: channels 0-7, 8-15
  __shared__ __attribute__((aligned(16))) bf16 s_b[4 * PLANE_B];  // first-layer outputs, 4 x 8 channels
  __shared__ __attribute__((aligned(16))) bf16 s_w[4 * PLANE_W];  // second-layer weights (resident)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int col = lane & 15, kg = lane >> 4;
  const int T = tiles_x * tiles_y * a0.n;  // image-major
  int t = blockIdx.x;
  if (t >= T) return;  // the whole workgroup, before any barrier
  const bf16x8 zero8 = {};
  // second-layer weights -> LDS (conv3x3_lds_kernel's layout and pieces)
  {
    bf16x8 pw[W_PER];
#pragma unroll
    for (int i = 0; i < W_PER; i++) {
      const int q = min(tid + i * 256, W_PIECES - 1);
      const int row = ((q >> 6) << 4) + (q & 15), part = (q >> 4) & 3;
      const int tap = row / (MB * 16), co = row % (MB * 16);
      pw[i] = *(const bf16x8 *)(a1.wt + (size_t)tap * 32 * 32 + (co * 4 + part) * 8);
    }
#pragma unroll
    for (int i = 0; i < W_PER; i++) {
      const int q = tid + i * 256;
      const int row = ((q >> 6) << 4) + (q & 15), part = (q >> 4) & 3;
      if (q < W_PIECES) *(bf16x8 *)(s_w + part * PLANE_W + row * 8) = pw[i];
    }
  }
  // first-layer weights in registers: the A operand of pair p, block mb
  // (output channel mb * 16 + col; k group kg = tap 2p + (kg >> 1),
  // channels 8 * (kg & 1) .. + 7 of the packed [tap][cout][32] weights)
  bf16x8 w0[5][MB];
  int toff[5];  // this lane's tap offset (elements) in the record tile, per pair
#pragma unroll
  for (int p = 0; p < 5; p++) {
    const int tap = 2 * p + (kg >> 1);
    toff[p] = tap < 9 ? ((tap / 3) * AW + tap % 3) * 8 : 0;
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
      w0[p][mb] = tap < 9 ? *(const bf16x8 *)(a0.wt + ((size_t)tap * 32 + mb * 16 + col) * 32 + (kg & 1) * 8) : zero8;
  }
  float bias0[MB][4], bias1[MB][4];
#pragma unroll
  for (int mb = 0; mb < MB; mb++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      bias0[mb][i] = a0.bias[mb * 16 + kg * 4 + i];
      bias1[mb][i] = a1.bias[mb * 16 + kg * 4 + i];
    }
  GinRaw<GIN> graw[GIN_PER];
  uint32_t valid = 0;
  auto fetch = [&](int tt) {
    const int u = tt % (tiles_x * tiles_y), n = tt / (tiles_x * tiles_y);
    const int x0 = (u % tiles_x) * TW, y0 = (u / tiles_x) * TH;
    valid = 0;
#pragma unroll
    for (int i = 0; i < GIN_PER; i++) {
      const int p = min(tid + i * 256, NA - 1);
      const int y = y0 - 2 + p / AW, x = x0 - 2 + p % AW;
      const bool ok = y >= 0 && y < a0.h && x >= 0 && x < a0.w;
      graw[i] = gin_load<GIN>(a0, min(max(y, 0), a0.h - 1), min(max(x, 0), a0.w - 1), n);
      valid |= (uint32_t)ok << i;
    }
  };
  fetch(t);
  for (;;) {
    const int u = t % (tiles_x * tiles_y), n = t / (tiles_x * tiles_y);
    const int x0 = (u % tiles_x) * TW, y0 = (u / tiles_x) * TH;
    // records -> LDS (zero outside the image: the first layer's padding);
    // the tile's own pixels also to rec_out when asked (pack_input_kernel's
    // layout: the training backward reads them)
#pragma unroll
    for (int i = 0; i < GIN_PER; i++) {
      const int p = tid + i * 256;
      bf16x8 v0 = {}, v1 = {};
      if ((valid >> i) & 1u) gin_pack<GIN>(graw[i], v0, v1);
      if (p < NA) {
        *(bf16x8 *)(s_a + p * 8) = v0;
        *(bf16x8 *)(s_a + PLANE_A + p * 8) = v1;
        const int ry = p / AW, rx = p % AW;
        if (a0.rec_out && ((valid >> i) & 1u) && ry >= 2 && ry < 2 + TH && rx >= 2 && rx < 2 + TW) {
          bf16 *dst = a0.rec_out + (((size_t)n * a0.h + (y0 - 2 + ry)) * a0.w + (x0 - 2 + rx)) * 32;
          *(bf16x8 *)dst = v0;
          *(bf16x8 *)(dst + 8) = v1;
          *(bf16x8 *)(dst + 16) = zero8;
          *(bf16x8 *)(dst + 24) = zero8;
        }
      }
    }
    __syncthreads();
    if (t + (int)gridDim.x < T) fetch(t + (int)gridDim.x);  // in flight during the matrix work
    // first layer: this wave's blocks, A_GROUP at a time (not unrolled: one
    // group's operands live at a time)
#pragma unroll 1
    for (int g = 0; g < A_PER; g += A_GROUP) {
      // this lane's first-layer output pixel in each block of the group
      // (clamped: the tail block's extra lanes compute a duplicate that is
      // not stored)
      int abase[A_GROUP];
#pragma unroll
      for (int j = 0; j < A_GROUP; j++) {
        const int q = min((wave + 4 * (g + j)) * 16 + col, NB - 1);
        abase[j] = ((q / IW) * AW + q % IW) * 8 + (kg & 1) * PLANE_A;
      }
      f32x4 acc0[A_GROUP][MB];
#pragma unroll
      for (int j = 0; j < A_GROUP; j++)
#pragma unroll
        for (int mb = 0; mb < MB; mb++) acc0[j][mb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int p = 0; p < 5; p++)
#pragma unroll
        for (int j = 0; j < A_GROUP; j++) {
          if (wave + 4 * (g + j) >= A_BLOCKS) continue;  // uniform
          bf16x8 bx = *(const bf16x8 *)(s_a + abase[j] + toff[p]);
          if (p == 4 && kg >= 2) bx = zero8;  // the tenth tap: zero weights, and no 0 * inf
#pragma unroll
          for (int mb = 0; mb < MB; mb++)
            acc0[j][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w0[p][mb], bx, acc0[j][mb], 0, 0, 0);
        }
      // bias, ReLU, BF16 (conv_store_row's rounding), zero outside the image
      // (the second layer's padding) -> LDS
#pragma unroll
      for (int j = 0; j < A_GROUP; j++) {
        const int blk = wave + 4 * (g + j);
        const int q = blk * 16 + col;
        if (blk >= A_BLOCKS || q >= NB) continue;
        const int qy = q / IW, qx = q % IW;
        const int y = y0 - 1 + qy, x = x0 - 1 + qx;
        const bool in = y >= 0 && y < a0.h && x >= 0 && x < a0.w;
        // the tile's own pixels (not the halo) also to HBM, when asked
        const bool keep = a0.out0 && in && qy >= 1 && qy <= TH && qx >= 1 && qx <= TW;
        typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int mb = 0; mb < MB; mb++) {
          bf16x4 o;
#pragma unroll
          for (int i = 0; i < 4; i++) {
            float v = acc0[j][mb][i] + bias0[mb][i];
            v = v > 0.f ? v : 0.f;
            o[i] = f2bf(in ? v : 0.f);
          }
          *(bf16x4 *)(s_b + (mb * 2 + (kg >> 1)) * PLANE_B + q * 8 + (kg & 1) * 4) = o;
          if (keep) *(bf16x4 *)(a0.out0 + (((size_t)n * a0.h + y) * a0.w + x) * 32 + mb * 16 + kg * 4) = o;
        }
      }
    }
    __syncthreads();
    // second layer on the LDS tile (conv3x3_lds_kernel's operand order)
    f32x4 acc[WR][WP][MB];
#pragma unroll
    for (int r = 0; r < WR; r++)
#pragma unroll
      for (int pb = 0; pb < WP; pb++)
#pragma unroll
        for (int mb = 0; mb < MB; mb++) acc[r][pb][mb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int dx = 0; dx < 3; dx++) {
      bf16x8 bx[WR + 2][WP];
#pragma unroll
      for (int rr = 0; rr < WR + 2; rr++)
#pragma unroll
        for (int pb = 0; pb < WP; pb++)
          bx[rr][pb] = *(const bf16x8 *)(s_b + kg * PLANE_B + ((wave * WR + rr) * IW + pb * 16 + col + dx) * 8);
#pragma unroll
      for (int dy = 0; dy < 3; dy++) {
        const int tap = dy * 3 + dx;
        bf16x8 aw[MB];
#pragma unroll
        for (int mb = 0; mb < MB; mb++) aw[mb] = *(const bf16x8 *)(s_w + kg * PLANE_W + ((tap * MB + mb) * 16 + col) * 8);
#pragma unroll
        for (int r = 0; r < WR; r++)
#pragma unroll
          for (int pb = 0; pb < WP; pb++)
#pragma unroll
            for (int mb = 0; mb < MB; mb++)
              acc[r][pb][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aw[mb], bx[r + dy][pb], acc[r][pb][mb], 0, 0, 0);
      }
    }
    // 2x2 pool of the stored (BF16) outputs (conv3x3_lds_kernel's pool
    // epilogue: rows r, r + 1 of this lane, columns x, x + 1 of lanes col,
    // col + 1, summed in avgpool_kernel's order)
    {
      const int hh = a1.h >> 1, ww = a1.w >> 1;
#pragma unroll
      for (int r = 0; r + 1 < WR; r += 2) {
        const int y = y0 + wave * WR + r;
#pragma unroll
        for (int pb = 0; pb < WP; pb++) {
          const int x = x0 + pb * 16 + col;
#pragma unroll
          for (int mb = 0; mb < MB; mb++) {
            const int co = mb * 16 + kg * 4;
            float tv[4], bv[4], tn[4], bn[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
              float v0 = acc[r][pb][mb][i] + bias1[mb][i], v1 = acc[r + 1][pb][mb][i] + bias1[mb][i];
              v0 = v0 > 0.f ? v0 : 0.f;
              v1 = v1 > 0.f ? v1 : 0.f;
              tv[i] = bf2f(f2bf(v0));
              bv[i] = bf2f(f2bf(v1));
              tn[i] = lane_xor1(tv[i]);
              bn[i] = lane_xor1(bv[i]);
            }
            if (!(col & 1) && (y >> 1) < hh && (x >> 1) < ww) {
              typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
              bf16x4 o;
#pragma unroll
              for (int i = 0; i < 4; i++) o[i] = f2bf((((tv[i] + tn[i]) + bv[i]) + bn[i]) * 0.25f);
              *(bf16x4 *)(a1.pool_out + (((size_t)n * hh + (y >> 1)) * ww + (x >> 1)) * 32 + co) = o;
            }
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < WR; r++) {
      const int y = y0 + wave * WR + r;
      if (y >= a1.h) break;
      conv_store_row<WP, MB>(a1, acc[r], bias1, n, y, x0, col, kg, 0);
    }
    t += gridDim.x;
    if (t >= T) break;
  }
}

// Weight and bias gradient of a 3x3 convolution:
//   dW[tap][ci][co] = sum_p X[p + shift(tap)][ci] * dY[p][co],  db[co] = sum_p dY[p][co]
// One workgroup: one image, one 32-channel group of the input, a 64-pixel
// column band and `rows` rows, taken RB rows per round.  Per round the RB + 2
// input rows and the RB output-gradient rows are staged in LDS transposed
// (channel-major, written as pixel pairs) so that MFMA operands are 8
// consecutive pixels of one channel; K runs over pixels.  Three aligned
// 16-byte reads of an input row give the operands of its three horizontal
// taps (the shifted ones by an element shift).
// Wave w owns input half cb = w & 1 and the output blocks mb = (w >> 1) + 2j,
// all 9 taps: per (row, k chunk) it reads each of its A fragments once for
// the 9 taps and each tap's B fragment once for its MB / 2 blocks.  Each
// workgroup stores its partial sums (no atomics) to its own slot of the
// scratch buffer, part[group][slot][tap][ci % 32][co], and
// wgrad_reduce_kernel adds the slots in a fixed order.  The workgroups of
// input group 0 also sum the staged output-gradient rows per channel (the
// bias gradient) into part_b[slot][co], reduced by bias_reduce_kernel.
template <int MB, int RB> __global__ __launch_bounds__(256) void conv3x3_wgrad_kernel(const WgradArgs a) {
  constexpr int XS = 64 + 16;  // staged x range [x0 - 8, x0 + 72): keeps dx = 0 reads 16-byte aligned
  constexpr int COUT = MB * 16;
  // LDS layouts.  A 128-bit operand read is served 16 lanes at a time, lanes
  // col = 0..7 and col = 8..15 of one kg together: with channel rows 44 banks
  // apart, lanes 0..7 start on 8 banks 4 apart and lanes 8..15 on the other
  // 8, so the 16 lanes cover the 64 banks once.  A staging store writes one
  // 32-bit pixel pair per lane, the lanes differing in their 8-channel group
  // and pixel pair: channel rows 44 banks apart put 8-channel groups 0 or 32
  // banks apart, and a pad of 512 / COUT banks per 16-channel group (16 for
  // the 32 input channels) separates the rest.  A staged input row holds 32
  // channel rows of XS pixels and is 40 banks longer than a multiple of 64,
  // so a store whose pixel pairs run past the end of a row stays conflict
  // free.
  constexpr int CST = 88;                                   // channel-row stride (bf16), XS and 64 pixels used
  constexpr int XROW = 2896;                                // >= 32 * CST + 2 * 32, 1448 banks = 40 mod 64
  constexpr int DPAD = 1024 / COUT, DROW = COUT * CST + (COUT / 16) * DPAD;
  static_assert(XROW >= 32 * CST + 2 * 32 && (XROW / 2) % 64 == 40, "input row layout");
  __shared__ __align__(16) bf16 xs[(RB + 2) * XROW];
  __shared__ __align__(16) bf16 ds[RB * DROW];
  auto xi = [](int rr, int c, int px) { return rr * XROW + c * CST + (c >> 4) * 32 + px; };
  auto di = [](int rr, int c, int px) { return rr * DROW + c * CST + (c >> 4) * DPAD + px; };
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, kg = lane >> 4;
  const int groups = (a.ca + a.cb) >> 5;
  // a workgroup computes COUT of the a.cout output channels: channel slice hf
  const int halves = a.cout / COUT;
  const int hf = blockIdx.z % halves, zc = blockIdx.z / halves;
  const int n = zc / groups, g = zc % groups;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * a.rows;
  const ConvInput in{a.a, a.ca, a.a_up, a.b, a.cb};
  constexpr int MJ = MB / 2;  // output blocks per wave
  const int cb = wave & 1, mb0 = wave >> 1;
  f32x4 acc[9][MJ];
#pragma unroll
  for (int t = 0; t < 9; t++)
#pragma unroll
    for (int j = 0; j < MJ; j++) acc[t][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // bias partial: thread t sums channel t % COUT over its eighth-, quarter-
  // or half-row slice of pixels (BP = 256 / COUT slices)
  constexpr int BP = 256 / COUT, BW = 64 / BP;
  const int bc = threadIdx.x % COUT, bq = threadIdx.x / COUT;
  float bsum = 0.f;
  typedef unsigned short u16;
  auto pair = [](bf16 lo, bf16 hi) {
    return (uint32_t)__builtin_bit_cast(u16, lo) | ((uint32_t)__builtin_bit_cast(u16, hi) << 16);
  };
  // The next round's rows are loaded into registers while the current
  // round's matrix work runs (loads unconditional, from clamped in-image
  // addresses; what lies outside the image or the round is staged as zero).
  constexpr int XI = ((RB + 2) * (XS / 2) * 4 + 255) / 256;  // input items per thread
  constexpr int DI = (RB * 32 * (COUT / 8) + 255) / 256;    // output-gradient items per thread
  bf16x8 px0[XI], px1[XI], pd0[DI], pd1[DI];
  uint32_t okx = 0, okd = 0;  // bit 2i / 2i+1: item i's first / second pixel lies inside
  auto fetch = [&](int yr, int nr) {
    okx = okd = 0;
#pragma unroll
    for (int i = 0; i < XI; i++) {
      const int e = min((int)threadIdx.x + i * 256, (RB + 2) * (XS / 2) * 4 - 1);
      const int q = e & 3, pp = (e >> 2) % (XS / 2), rr = (e >> 2) / (XS / 2);
      const int y = yr + rr - 1, x = x0 - 8 + pp * 2;
      const bool rowok = rr < nr + 2 && y >= 0 && y < a.h;
      px0[i] = load8_clamped(in, n, y, x, a.h, a.w, g * 32 + q * 8);
      px1[i] = load8_clamped(in, n, y, x + 1, a.h, a.w, g * 32 + q * 8);
      okx |= (uint32_t)(rowok && x >= 0 && x < a.w) << (2 * i);
      okx |= (uint32_t)(rowok && x + 1 >= 0 && x + 1 < a.w) << (2 * i + 1);
    }
#pragma unroll
    for (int i = 0; i < DI; i++) {
      const int e = min((int)threadIdx.x + i * 256, RB * 32 * (COUT / 8) - 1);
      const int q = e % (COUT / 8), pp = (e / (COUT / 8)) % 32, rr = e / (COUT / 8) / 32;
      const int y = min(yr + rr, a.h - 1), x = x0 + pp * 2;
      const bf16 *row = a.dy + ((size_t)n * a.h + y) * a.w * a.cout + hf * COUT + q * 8;
      pd0[i] = *(const bf16x8 *)(row + (size_t)min(x, a.w - 1) * a.cout);
      pd1[i] = *(const bf16x8 *)(row + (size_t)min(x + 1, a.w - 1) * a.cout);
      okd |= (uint32_t)(rr < nr && x < a.w) << (2 * i);
      okd |= (uint32_t)(rr < nr && x + 1 < a.w) << (2 * i + 1);
    }
  };
  if (y0 < a.h) fetch(y0, min(RB, min(a.rows, a.h - y0)));
  for (int r0 = 0; r0 < a.rows; r0 += RB) {
    const int yr = y0 + r0;
    if (yr >= a.h) break;  // uniform across the workgroup
    const int nr = min(RB, min(a.rows - r0, a.h - yr));
    // input rows yr - 1 .. yr + nr (32 channels of group g), transposed, pixel pairs
#pragma unroll
    for (int i = 0; i < XI; i++) {
      const int e = (int)threadIdx.x + i * 256;
      if (e < (RB + 2) * (XS / 2) * 4) {
        const int q = e & 3, pp = (e >> 2) % (XS / 2), rr = (e >> 2) / (XS / 2);
        const bf16x8 z = {};
        const bf16x8 v0 = (okx >> (2 * i)) & 1u ? px0[i] : z, v1 = (okx >> (2 * i + 1)) & 1u ? px1[i] : z;
#pragma unroll
        for (int k = 0; k < 8; k++) *(uint32_t *)&xs[xi(rr, q * 8 + k, pp * 2)] = pair(v0[k], v1[k]);
      }
    }
    // output-gradient rows yr .. yr + nr - 1, transposed, pixel pairs
#pragma unroll
    for (int i = 0; i < DI; i++) {
      const int e = (int)threadIdx.x + i * 256;
      if (e < RB * 32 * (COUT / 8)) {
        const int q = e % (COUT / 8), pp = (e / (COUT / 8)) % 32, rr = e / (COUT / 8) / 32;
        const bf16x8 z = {};
        const bf16x8 v0 = (okd >> (2 * i)) & 1u ? pd0[i] : z, v1 = (okd >> (2 * i + 1)) & 1u ? pd1[i] : z;
#pragma unroll
        for (int k = 0; k < 8; k++) *(uint32_t *)&ds[di(rr, q * 8 + k, pp * 2)] = pair(v0[k], v1[k]);
      }
    }
    __syncthreads();
    {  // the next round's rows, in flight during this round's matrix work
      const int yn = yr + RB;
      if (r0 + RB < a.rows && yn < a.h) fetch(yn, min(RB, min(a.rows - r0 - RB, a.h - yn)));
    }
    if (g == 0) {  // uniform across the workgroup
      for (int rr = 0; rr < nr; rr++)
#pragma unroll
        for (int k = 0; k < BW; k += 8) {
          const bf16x8 v = *(const bf16x8 *)&ds[di(rr, bc, bq * BW + k)];
#pragma unroll
          for (int i = 0; i < 8; i++) bsum += bf2f(v[i]);
        }
    }
    for (int rr = 0; rr < nr; rr++) {
#pragma unroll
      for (int kc = 0; kc < 2; kc++) {
        const int p0 = kc * 32 + kg * 8;
        bf16x8 av[MJ];  // A: co x pixels
#pragma unroll
        for (int j = 0; j < MJ; j++) av[j] = *(const bf16x8 *)&ds[di(rr, (mb0 + 2 * j) * 16 + col, p0)];
#pragma unroll
        for (int dy = 0; dy < 3; dy++) {
          const bf16 *base = &xs[xi(rr + dy, cb * 16 + col, p0 + 8)];  // B: pixels x ci, 16-byte aligned
          const bf16x8 lo = *(const bf16x8 *)(base - 8), mid = *(const bf16x8 *)base,
                       hi = *(const bf16x8 *)(base + 8);
          bf16x8 bm, bp;  // dx = -1: elements base[-1 .. 6]; dx = +1: base[1 .. 8]
          bm[0] = lo[7];
#pragma unroll
          for (int i = 1; i < 8; i++) bm[i] = mid[i - 1];
#pragma unroll
          for (int i = 0; i < 7; i++) bp[i] = mid[i + 1];
          bp[7] = hi[0];
#pragma unroll
          for (int j = 0; j < MJ; j++) {
            acc[dy * 3 + 0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[j], bm, acc[dy * 3 + 0][j], 0, 0, 0);
            acc[dy * 3 + 1][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[j], mid, acc[dy * 3 + 1][j], 0, 0, 0);
            acc[dy * 3 + 2][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[j], bp, acc[dy * 3 + 2][j], 0, 0, 0);
          }
        }
      }
    }
    __syncthreads();
  }
  const size_t slots = (size_t)a.n * gridDim.y * gridDim.x;
  const size_t slot = ((size_t)n * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  float *dst = a.part + ((size_t)g * slots + slot) * (9 * 32 * a.cout) + hf * COUT;
#pragma unroll
  for (int tap = 0; tap < 9; tap++)
#pragma unroll
    for (int j = 0; j < MJ; j++)
      *(f32x4 *)&dst[(tap * 32 + cb * 16 + col) * a.cout + (mb0 + 2 * j) * 16 + kg * 4] = acc[tap][j];
  if (g == 0) {
    // the BP slices of each channel, added in a fixed order (xs is free now)
    float *red = (float *)&xs[0];
    red[bq * COUT + bc] = bsum;
    __syncthreads();
    if (threadIdx.x < COUT) {
      float b = 0.f;
#pragma unroll
      for (int k = 0; k < BP; k++) b += red[k * COUT + threadIdx.x];
      a.part_b[slot * a.cout + hf * COUT + threadIdx.x] = b;
    }
  }
}

// The fixed-order reductions of conv3x3_wgrad_kernel's per-workgroup partial
// sums, in one launch.  Blocks [0, nwb): the weight gradient,
//   dw[co][ci][tap] = sum over slots of part[g][slot][tap][ci % 32][co]
// written in the master layout for ci < cin_log (g = ci / 32); a workgroup
// holds 256 / S float4 elements x S slices, slice k sums slots k, k + S, ...
// and the S slice sums are added in a fixed tree.  Block nwb + co: the bias
// gradient db[co] = sum over slots of part_b[slot][co], 256 strided partial
// sums then a fixed tree.  Neither result depends on the order in which
// workgroups ran.
template <int S>
__global__ __launch_bounds__(256) void wgrad_reduce_kernel(const f32x4 *part, int slots, int per4, int groups,
                                                          int cout, int cin_log, int nwb, const float *part_b,
                                                          float *dw, float *db) {
  if ((int)blockIdx.x >= nwb) {  // uniform per workgroup
    __shared__ float redb[256];
    const int co = (int)blockIdx.x - nwb;
    float sb = 0.f;
    for (int k = threadIdx.x; k < slots; k += 256) sb += part_b[(size_t)k * cout + co];
    redb[threadIdx.x] = sb;
    __syncthreads();
#pragma unroll
    for (int st = 128; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) redb[threadIdx.x] += redb[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) db[co] = redb[0];
    return;
  }
  constexpr int E = 256 / S;
  __shared__ f32x4 red[S][E];
  const int el = threadIdx.x % E, sl = threadIdx.x / E;
  const size_t e = (size_t)blockIdx.x * E + el, total = (size_t)groups * per4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  if (e < total) {
    const size_t g = e / per4, f = e % per4;
    const f32x4 *src = part + g * slots * per4 + f;
    for (int k = sl; k < slots; k += S) acc += src[(size_t)k * per4];
  }
  red[sl][el] = acc;
  __syncthreads();
#pragma unroll
  for (int st = S / 2; st > 0; st >>= 1) {
    if (sl < st) red[sl][el] += red[sl + st][el];
    __syncthreads();
  }
  if (sl == 0 && e < total) {
    const int g = (int)(e / per4), f = (int)(e % per4) * 4;
    const int tap = f / (32 * cout), rem = f % (32 * cout), cl = rem / cout, co = rem % cout;
    const int ci = g * 32 + cl;
    if (ci < cin_log) {
      const f32x4 v = red[0][el];
#pragma unroll
      for (int i = 0; i < 4; i++) dw[((size_t)(co + i) * cin_log + ci) * 9 + tap] = v[i];
    }
  }
}


// ------------------------------------------------------------ glue kernels

// Network input: [noisy irradiance rgb, depth, oct normal xy, albedo rgb, 0...]
// from a compact FP32 record of 12 floats per pixel (the last 3 unused).
__global__ __launch_bounds__(256) void pack_input_kernel(const float *x12, size_t npix, bf16 *x32) {
  const size_t p = (size_t)blockIdx.x * 64 + (threadIdx.x >> 2);
  const int q = threadIdx.x & 3;  // 8-channel quarter
  if (p >= npix) return;
  bf16x8 v = {};
  if (q == 0) {
    for (int i = 0; i < 8; i++) v[i] = f2bf(x12[p * 12 + i]);
  } else if (q == 1) {
    v[0] = f2bf(x12[p * 12 + 8]);
  }
  *(bf16x8 *)(x32 + p * 32 + q * 8) = v;
}

__global__ __launch_bounds__(256) void avgpool_kernel(const bf16 *x, int n, int h, int w, int c, bf16 *y) {
  // y[n][h/2][w/2][c] = mean of the 2x2 children
  const int hh = h >> 1, ww = w >> 1;
  const size_t total = (size_t)n * hh * ww * (c / 8);
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
    const int q = (int)(e % (c / 8));
    size_t p = e / (c / 8);
    const int xx = (int)(p % ww);
    p /= ww;
    const int yy = (int)(p % hh);
    const int nn = (int)(p / hh);
    const size_t b00 = (((size_t)nn * h + 2 * yy) * w + 2 * xx) * c + q * 8;
    const bf16x8 v0 = *(const bf16x8 *)(x + b00), v1 = *(const bf16x8 *)(x + b00 + c);
    const bf16x8 v2 = *(const bf16x8 *)(x + b00 + (size_t)w * c), v3 = *(const bf16x8 *)(x + b00 + (size_t)w * c + c);
    bf16x8 o;
#pragma unroll
    for (int i = 0; i < 8; i++) o[i] = f2bf((bf2f(v0[i]) + bf2f(v1[i]) + bf2f(v2[i]) + bf2f(v3[i])) * 0.25f);
    *(bf16x8 *)(y + (((size_t)nn * hh + yy) * ww + xx) * c + q * 8) = o;
  }
}

// dx[p] = mask(act[p] > 0) * (dx_in[p] + dpool[p/2] / 4): the gradient of a
// tensor feeding both a concatenation (dx_in) and an average pool (dpool).
// Rows (image nn, row yy) over blockIdx.y, their pixels x 8-channel groups
// over the threads: 32-bit index arithmetic, no division in the loop (c / 8
// is a power of two).
__global__ __launch_bounds__(256) void pool_bwd_add_kernel(const bf16 *dx_in, const bf16 *dpool, const bf16 *act, int n,
                                                           int h, int w, int c, bf16 *dx) {
  const int cq = c / 8, cs = __builtin_ctz((unsigned)cq);
  for (int row = blockIdx.y; row < n * h; row += gridDim.y) {
  const int yy = row % h, nn = row / h;
  const size_t r0 = (size_t)row * w * c, rp = ((size_t)nn * (h >> 1) + (yy >> 1)) * (w >> 1) * c;
  for (int e = blockIdx.x * 256 + threadIdx.x; e < w * cq; e += gridDim.x * 256) {
    const int q = e & (cq - 1), xx = e >> cs;
    const size_t i0 = r0 + (size_t)xx * c + q * 8, ip = rp + (size_t)(xx >> 1) * c + q * 8;
    const bf16x8 a = *(const bf16x8 *)(dx_in + i0), b = *(const bf16x8 *)(dpool + ip), m = *(const bf16x8 *)(act + i0);
    bf16x8 o;
#pragma unroll
    for (int i = 0; i < 8; i++) o[i] = bf2f(m[i]) > 0.f ? f2bf(bf2f(a[i]) + 0.25f * bf2f(b[i])) : f2bf(0.f);
    *(bf16x8 *)(dx + i0) = o;
  }
  }
}

// Gradient through a nearest x2 upsample: dhalf[q] = mask(act[q] > 0) * sum of its 4 children.
__global__ __launch_bounds__(256) void upsample_bwd_kernel(const bf16 *dfull, const bf16 *act, int n, int hh, int ww, int c,
                                                           bf16 *dhalf) {
  const int h = hh * 2, w = ww * 2;
  const size_t total = (size_t)n * hh * ww * (c / 8);
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
    const int q = (int)(e % (c / 8));
    size_t p = e / (c / 8);
    const int xx = (int)(p % ww);
    p /= ww;
    const int yy = (int)(p % hh);
    const int nn = (int)(p / hh);
    const size_t b00 = (((size_t)nn * h + 2 * yy) * w + 2 * xx) * c + q * 8;
    const bf16x8 v0 = *(const bf16x8 *)(dfull + b00), v1 = *(const bf16x8 *)(dfull + b00 + c);
    const bf16x8 v2 = *(const bf16x8 *)(dfull + b00 + (size_t)w * c);
    const bf16x8 v3 = *(const bf16x8 *)(dfull + b00 + (size_t)w * c + c);
    const size_t oi = (((size_t)nn * hh + yy) * ww + xx) * c + q * 8;
    const bf16x8 m = *(const bf16x8 *)(act + oi);
    bf16x8 o;
#pragma unroll
    for (int i = 0; i < 8; i++)
      o[i] = bf2f(m[i]) > 0.f ? f2bf(bf2f(v0[i]) + bf2f(v1[i]) + bf2f(v2[i]) + bf2f(v3[i])) : f2bf(0.f);
    *(bf16x8 *)(dhalf + oi) = o;
  }
}

// Output head: out = noisy + W_h * d1 + b_h (1x1, 32 -> 3, FP32 output RGBA).
__global__ __launch_bounds__(256) void head_fwd_kernel(const bf16 *d1, const float *x12, const float *wh,
                                                       const float *bh, size_t npix, int w, int out_pitch,
                                                       float4 *out) {
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= npix) return;
  float o[3];
  for (int c = 0; c < 3; c++) o[c] = x12[p * 12 + c] + bh[c];
  for (int q = 0; q < 4; q++) {
    const bf16x8 v = *(const bf16x8 *)(d1 + p * 32 + q * 8);
    for (int i = 0; i < 8; i++) {
      const float f = bf2f(v[i]);
      for (int c = 0; c < 3; c++) o[c] += wh[c * 32 + q * 8 + i] * f;
    }
  }
  out[(p / w) * out_pitch + p % w] = make_float4(o[0], o[1], o[2], 1.f);
}

// The head for inference: the residual comes from the noisy image itself.
__global__ __launch_bounds__(256) void head_fwd_img_kernel(const bf16 *d1, const float4 *noisy, int np_,
                                                           const float *wh, const float *bh, int w, int h,
                                                           int out_pitch, float4 *out) {
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= (size_t)w * h) return;
  const int x = (int)(p % w), y = (int)(p / w);
  const float4 c = noisy[(size_t)y * np_ + x];
  float o[3] = {c.x + bh[0], c.y + bh[1], c.z + bh[2]};
  for (int q = 0; q < 4; q++) {
    const bf16x8 v = *(const bf16x8 *)(d1 + p * 32 + q * 8);
    for (int i = 0; i < 8; i++) {
      const float f = bf2f(v[i]);
      for (int k = 0; k < 3; k++) o[k] += wh[k * 32 + q * 8 + i] * f;
    }
  }
  out[(size_t)y * out_pitch + x] = make_float4(o[0], o[1], o[2], 1.f);
}

// Training: the 1x1 head, the loss and the head's backward in one pass (the
// head's output is not stored).  One pixel per thread: out = x + b_h + W_h
// act, summed in head_fwd_kernel's order; mean L1: g = sign(out - target) /
// count per channel, mean squared error (l2): g = 2 (out - target) / count;
// dd1 = mask(d1 > 0) W_h^T g (per channel, the three products in channel
// order).  A thread keeps 3 x 32 head-weight partial sums; they are reduced
// over the lanes of a wave with shuffles, over the waves in LDS, and added to
// dW_h, db_h and the loss once per workgroup.
__global__ __launch_bounds__(256) void head_train_kernel(const bf16 *__restrict__ d1, const float *__restrict__ x12,
                                                         const float4 *__restrict__ target, const float *__restrict__ wh,
                                                         const float *__restrict__ bh, size_t npix, float inv_count,
                                                         int l2, bf16 *__restrict__ dd1, float *dwh, float *dbh,
                                                         float *loss) {
  __shared__ float red[4][100];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float aw[3][32], ab[3] = {0.f, 0.f, 0.f}, al = 0.f;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int i = 0; i < 32; i++) aw[c][i] = 0.f;
  for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < npix; p += (size_t)gridDim.x * 256) {
    bf16x8 v[4];
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = *(const bf16x8 *)(d1 + p * 32 + q * 8);
    const float4 t = target[p];
    float o[3];
#pragma unroll
    for (int c = 0; c < 3; c++) o[c] = x12[p * 12 + c] + bh[c];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const float f = bf2f(v[q][i]);
#pragma unroll
        for (int c = 0; c < 3; c++) o[c] += wh[c * 32 + q * 8 + i] * f;
      }
    const float df[3] = {o[0] - t.x, o[1] - t.y, o[2] - t.z};
    float g[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      g[c] = (l2 ? 2.f * df[c] : (df[c] > 0.f ? 1.f : (df[c] < 0.f ? -1.f : 0.f))) * inv_count;
      al += l2 ? df[c] * df[c] : fabsf(df[c]);
      ab[c] += g[c];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      bf16x8 o8;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const float f = bf2f(v[q][i]);
        float sacc = 0.f;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          sacc += wh[c * 32 + q * 8 + i] * g[c];
          aw[c][q * 8 + i] += g[c] * f;
        }
        o8[i] = f > 0.f ? f2bf(sacc) : f2bf(0.f);
      }
      *(bf16x8 *)(dd1 + p * 32 + q * 8) = o8;
    }
  }
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
#pragma unroll
      for (int i = 0; i < 32; i++) aw[c][i] += __shfl_xor(aw[c][i], m);
      ab[c] += __shfl_xor(ab[c], m);
    }
    al += __shfl_xor(al, m);
  }
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
#pragma unroll
      for (int i = 0; i < 32; i++) red[wave][c * 32 + i] = aw[c][i];
      red[wave][96 + c] = ab[c];
    }
    red[wave][99] = al;
  }
  __syncthreads();
  if (threadIdx.x < 100) {
    const float v = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
    if (threadIdx.x < 96) atomicAdd(&dwh[threadIdx.x], v);
    else if (threadIdx.x < 99) atomicAdd(&dbh[threadIdx.x - 96], v);
    else if (loss) atomicAdd(loss, v * inv_count);
  }
}

// Adam on the FP32 master parameters (Kingma & Ba 2015, bias-corrected);
// every updated weight of the convolutions is also written to its BF16
// forward and input-gradient layouts (pack_weights_kernel's placement; the
// zero padding channels are never written).
__global__ __launch_bounds__(256) void adam_pack_kernel(float *p, const float *g, float *m, float *v, size_t n, float lr,
                                                        float b1, float b2, float eps, float bc1, float bc2,
                                                        const PackTable pk) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    p[i] = pi;
#pragma unroll 1
    for (int l = 0; l < pk.n; l++) {
      const PackLayer &L = pk.l[l];
      if (i < L.off || i >= L.off + (size_t)L.cout * L.cin_log * 9) continue;
      const size_t r = i - L.off;
      const int tap = (int)(r % 9), ci = (int)((r / 9) % L.cin_log), co = (int)(r / 9 / L.cin_log);
      const bf16 b = f2bf(pi);
      L.wf[(((size_t)tap * (L.cin_pad / 32) + ci / 32) * L.cout + co) * 32 + (ci & 31)] = b;
      if (L.wt) L.wt[(((size_t)(8 - tap) * (L.cout / 32) + co / 32) * L.cin_pad + ci) * 32 + (co & 31)] = b;
      break;
    }
  }
}

__global__ __launch_bounds__(256) void fill_kernel(float *p, size_t n, float v) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = v;
}

// Pack master weights W[co][ci][3][3] (FP32, ci < cin_log) into the forward
// layout [tap][cin_pad/32][cout][32] and, when wt != null, the input-gradient
// layout: the same convolution read as cout -> cin_pad with flipped taps,
// [tap'][cout/32][cin_pad][32] with tap' = 8 - tap.
__global__ __launch_bounds__(256) void pack_weights_kernel(const float *w, int cout, int cin_log, int cin_pad, bf16 *wf,
                                                           bf16 *wt) {
  const size_t total = (size_t)cout * cin_pad * 9;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
    const int tap = (int)(e % 9);
    const int ci = (int)((e / 9) % cin_pad);
    const int co = (int)(e / 9 / cin_pad);
    const float v = ci < cin_log ? w[((size_t)co * cin_log + ci) * 9 + tap] : 0.f;
    wf[(((size_t)tap * (cin_pad / 32) + ci / 32) * cout + co) * 32 + (ci & 31)] = f2bf(v);
    if (wt) wt[(((size_t)(8 - tap) * (cout / 32) + co / 32) * cin_pad + ci) * 32 + (co & 31)] = f2bf(v);
  }
}

// MFMA operand-layout probe: D = A * B for a 16x32 A and 32x16 B given in
// row-major order, computed by one wave with the layout stated above.
__global__ __launch_bounds__(64) void mfma_probe_kernel(const float *A, const float *B, float *D) {
  const int l = threadIdx.x, col = l & 15, kg = l >> 4;
  bf16x8 a, b;
  for (int i = 0; i < 8; i++) {
    a[i] = f2bf(A[col * 32 + kg * 8 + i]);
    b[i] = f2bf(B[(kg * 8 + i) * 16 + col]);
  }
  f32x4 d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
  for (int i = 0; i < 4; i++) D[(kg * 4 + i) * 16 + col] = d[i];
}

// ------------------------------------------------------------ launchers

static unsigned grid_for(size_t work, size_t per_block) {
  size_t g = (work + per_block - 1) / per_block;
  if (g > 65535u * 16u) g = 65535u * 16u;
  return (unsigned)(g ? g : 1);
}

// Register tiles (R rows, PB pixel blocks) offered per output-channel block
// count MB, accumulators R * PB * MB * 4 VGPRs <= 128; entry 0 is the
// default, the others are the autotuner's candidates.
struct ConvTile {
  int r, pb, lds;  // lds: conv3x3_lds_kernel with WR = r, WP = pb
  int mbw;         // lds: output-channel blocks per workgroup (cout / (16 * mbw) splits)
  int per_cu;      // lds: workgroups per CU walking the tiles, 0 = one workgroup per tile
};
static const ConvTile kTiles1[] = {{4, 2, 0, 1, 0}, {2, 4, 0, 1, 0}, {4, 4, 0, 1, 0}, {1, 4, 0, 1, 0}};
static const ConvTile kTiles2[] = {{4, 2, 0, 2, 0}, {2, 4, 0, 2, 0}, {4, 4, 0, 2, 0}, {1, 4, 0, 2, 0},
                                   {2, 4, 1, 2, 0}, {2, 2, 1, 2, 0}, {1, 4, 1, 2, 0}, {2, 4, 1, 2, 2},
                                   {2, 2, 1, 2, 3}, {1, 4, 1, 2, 3}};
static const ConvTile kTiles4[] = {{2, 2, 0, 4, 0}, {4, 2, 0, 4, 0}, {2, 4, 0, 4, 0}, {1, 4, 0, 4, 0},
                                   {2, 2, 1, 4, 0}, {1, 4, 1, 4, 0}, {1, 2, 1, 4, 0}, {2, 2, 1, 4, 2},
                                   {1, 2, 1, 4, 3}, {2, 4, 1, 2, 0}, {2, 2, 1, 2, 3}};
static const ConvTile kTiles6[] = {{2, 2, 0, 6, 0}, {1, 4, 0, 6, 0}, {1, 2, 0, 6, 0}, {2, 4, 1, 2, 0},
                                   {2, 2, 1, 2, 3}};
static const ConvTile kTiles8[] = {{1, 2, 0, 8, 0}, {2, 2, 0, 8, 0}, {1, 4, 0, 8, 0}, {2, 2, 1, 4, 0},
                                   {1, 2, 1, 4, 3}, {2, 2, 1, 4, 2}};
static const ConvTile kTiles12[] = {{1, 2, 0, 12, 0}, {2, 1, 0, 12, 0}, {2, 2, 1, 4, 0}, {1, 2, 1, 4, 3}};

static const ConvTile *conv_tiles(int mb, int *count) {
#define LGX_TILES(T) \
  *count = (int)(sizeof(T) / sizeof(T[0])); \
  return T
  switch (mb) {
    case 1: LGX_TILES(kTiles1);
    case 2: LGX_TILES(kTiles2);
    case 4: LGX_TILES(kTiles4);
    case 6: LGX_TILES(kTiles6);
    case 8: LGX_TILES(kTiles8);
    case 12: LGX_TILES(kTiles12);
    default: *count = 0; return nullptr;
  }
#undef LGX_TILES
}

int conv3x3_tile_lds(int cout, int tile) {
  int n = 0;
  if (cout % 16) return 0;
  const ConvTile *t = conv_tiles(cout / 16, &n);
  return t && tile >= 0 && tile < n && t[tile].lds && t[tile].mbw * 16 == cout;
}

int conv3x3_tile_pool(int cout, int tile) {
  int n = 0;
  if (cout % 16) return 0;
  const ConvTile *t = conv_tiles(cout / 16, &n);
  return t && tile >= 0 && tile < n && t[tile].lds && t[tile].r % 2 == 0;
}

int conv3x3_tiles(int cout) {
  int n = 0;
  if (cout % 16) return 0;
  (void)conv_tiles(cout / 16, &n);
  return n;
}

template <int MB, int PB, int R> static void launch_conv(const ConvArgs &a, hipStream_t s) {
  dim3 grid((a.w + 4 * PB * 16 - 1) / (4 * PB * 16), (a.h + R - 1) / R, a.n), block(256);
  hipLaunchKernelGGL((conv3x3_kernel<MB, PB, R>), grid, block, 0, s, a);
}

static int cu_count() {
  static int cus = 0;
  if (!cus) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) ==
                                                hipSuccess && v > 0)
      cus = v;
    else
      cus = 256;
  }
  return cus;
}

template <int MBW, int WR, int WP> static void launch_conv_lds(const ConvArgs &a, int per_cu, hipStream_t s) {
  const int tx = (a.w + 16 * WP - 1) / (16 * WP), ty = (a.h + 4 * WR - 1) / (4 * WR), splits = a.cout / (16 * MBW);
  const long tiles = (long)tx * ty * a.n * splits;
  long grid = per_cu > 0 ? std::min(tiles, (long)per_cu * cu_count()) : tiles;
  if constexpr (MBW == 2) {
    if (a.head_out) {  // one tile per workgroup: nothing is prefetched across the head's epilogue
      hipLaunchKernelGGL((conv3x3_lds_kernel<MBW, WR, WP, true>), dim3((unsigned)tiles), dim3(256), 0, s, a, tx, ty,
                         splits);
      return;
    }
    if (a.gin_fmt == 1) {
      hipLaunchKernelGGL((conv3x3_lds_kernel<MBW, WR, WP, false, 1>), dim3((unsigned)grid), dim3(256), 0, s, a, tx, ty,
                         splits);
      return;
    }
    if (a.gin_fmt == 2) {
      hipLaunchKernelGGL((conv3x3_lds_kernel<MBW, WR, WP, false, 2>), dim3((unsigned)grid), dim3(256), 0, s, a, tx, ty,
                         splits);
      return;
    }
  }
  if constexpr (WR % 2 == 0) {
    if (a.up_out) {
      hipLaunchKernelGGL((conv3x3_lds_kernel<MBW, WR, WP, false, 0, true>), dim3((unsigned)grid), dim3(256), 0, s, a,
                         tx, ty, splits);
      return;
    }
  }
  hipLaunchKernelGGL((conv3x3_lds_kernel<MBW, WR, WP, false>), dim3((unsigned)grid), dim3(256), 0, s, a, tx, ty,
                     splits);
}

template <int MB> static hipError_t conv_mb(const ConvArgs &a, const ConvTile &t, hipStream_t s) {
#define LGX_CASE(RR, PP) \
  if (!t.lds && t.r == RR && t.pb == PP) { \
    launch_conv<MB, PP, RR>(a, s); \
    return hipGetLastError(); \
  }
#define LGX_LDS(MBW, RR, PP) \
  if (t.lds && t.mbw == MBW && t.r == RR && t.pb == PP) { \
    launch_conv_lds<MBW, RR, PP>(a, t.per_cu, s); \
    return hipGetLastError(); \
  }
  if constexpr (MB <= 2) {
    LGX_CASE(4, 2) LGX_CASE(2, 4) LGX_CASE(4, 4) LGX_CASE(1, 4)
  } else if constexpr (MB <= 4) {
    LGX_CASE(2, 2) LGX_CASE(4, 2) LGX_CASE(2, 4) LGX_CASE(1, 4)
  } else if constexpr (MB <= 8) {
    LGX_CASE(2, 2) LGX_CASE(1, 4) LGX_CASE(1, 2)  // covers the MB = 6 and MB = 8 tables
  } else {
    LGX_CASE(1, 2) LGX_CASE(2, 1)
  }
  if constexpr (MB % 2 == 0) {
    LGX_LDS(2, 2, 4) LGX_LDS(2, 2, 2) LGX_LDS(2, 1, 4)
  }
  if constexpr (MB % 4 == 0) {
    LGX_LDS(4, 2, 2) LGX_LDS(4, 1, 4) LGX_LDS(4, 1, 2)
  }
#undef LGX_CASE
#undef LGX_LDS
  return hipErrorInvalidValue;
}

hipError_t conv3x3(const ConvArgs &a, hipStream_t s) {
  if (a.cout % 16) return hipErrorInvalidValue;
  if (a.head_out && (a.cout != 32 || a.n != 1 || !conv3x3_tile_lds(a.cout, a.tile))) return hipErrorInvalidValue;
  if (a.gin_fmt && (a.gin_fmt > 2 || a.head_out || a.cout != 32 || a.n != 1 || a.ca != 32 || a.cb ||
                    !conv3x3_tile_lds(a.cout, a.tile)))
    return hipErrorInvalidValue;
  if (a.pool_out && (a.head_out || a.mask0 || a.mask1 || !conv3x3_tile_pool(a.cout, a.tile)))
    return hipErrorInvalidValue;
  if (a.up_out && (a.head_out || a.pool_out || a.mask0 || !a.up_mask || a.split < 16 || a.h % 2 || a.w % 2 ||
                   !conv3x3_tile_pool(a.cout, a.tile)))
    return hipErrorInvalidValue;
  if (a.split < 0 || a.split > a.cout || a.split % 16) return hipErrorInvalidValue;  // whole 16-channel blocks
  const int mb = a.cout / 16;
  int n = 0;
  const ConvTile *tiles = conv_tiles(mb, &n);
  if (!tiles || a.tile < 0 || a.tile >= n) return hipErrorInvalidValue;
  const ConvTile &t = tiles[a.tile];
  switch (mb) {
    case 1: return conv_mb<1>(a, t, s);
    case 2: return conv_mb<2>(a, t, s);
    case 4: return conv_mb<4>(a, t, s);
    case 6: return conv_mb<6>(a, t, s);
    case 8: return conv_mb<8>(a, t, s);
    case 12: return conv_mb<12>(a, t, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t conv_first_pair(const ConvArgs &a0, const ConvArgs &a1, hipStream_t s) {
  if (a0.gin_fmt < 1 || a0.gin_fmt > 3 || (a0.gin_fmt == 3 ? !a0.gin_x12 : a0.n != 1) || a0.n < 1 ||
      a0.cout != 32 || !a0.wt || !a0.bias || a1.n != a0.n || a1.cout != 32 || a1.split != 32 || !a1.out0 ||
      !a1.pool_out || !a1.wt || !a1.bias || a1.h != a0.h || a1.w != a0.w || a0.h < 1 || a0.w < 1)
    return hipErrorInvalidValue;
  const int tx = (a0.w + 31) / 32, ty = (a0.h + 7) / 8;
  const long tiles = (long)tx * ty * a0.n;
  const int grid = (int)std::min(tiles, 2L * cu_count());  // two workgroups per CU fit its LDS and registers
  if (a0.gin_fmt == 1)
    hipLaunchKernelGGL(conv_first_pair_kernel<1>, dim3((unsigned)grid), dim3(256), 0, s, a0, a1, tx, ty);
  else if (a0.gin_fmt == 2)
    hipLaunchKernelGGL(conv_first_pair_kernel<2>, dim3((unsigned)grid), dim3(256), 0, s, a0, a1, tx, ty);
  else
    hipLaunchKernelGGL(conv_first_pair_kernel<3>, dim3((unsigned)grid), dim3(256), 0, s, a0, a1, tx, ty);
  return hipGetLastError();
}

// Weight-gradient workgroups per launch by output-channel count: each stores
// 9 * 32 * cout FP32 partial sums, about 38 MB per launch at these counts.
// Workgroups of conv3x3_wgrad_kernel: about the number that are resident at
// once (2 per CU: the accumulators live in AGPRs, which share the register
// file with the VGPRs, and LDS), so that the launch is one round of
// workgroups rather than a round and a fraction with most CUs idle in the
// fraction.
static int wgrad_target_blocks(int cout) { return (cout == 128 ? 1 : 2) * cu_count(); }  // x2 slices for 128
// Output rows staged per round of conv3x3_wgrad_kernel: a round's loads are
// in flight during the previous round's matrix work, so taller rounds expose
// that latency fewer times, within two workgroups' LDS per CU (64 channels:
// 3 rows).
static constexpr int kWgradRB2 = 4, kWgradRB4 = 3;
static int wgrad_rb(int cout) { return cout <= 32 ? kWgradRB2 : kWgradRB4; }

size_t conv3x3_wgrad_scratch_bytes() {
  size_t m = 0;
  for (int cout : {32, 64, 128}) m = std::max(m, (size_t)2 * wgrad_target_blocks(cout) * (9 * 32 + 1) * cout * 4);
  return m;
}

hipError_t conv3x3_wgrad(WgradArgs a, hipStream_t s) {
  if (a.cout != 32 && a.cout != 64 && a.cout != 128) return hipErrorInvalidValue;
  if ((a.ca + a.cb) % 32 || a.ca % 32 || a.n < 1 || a.h < 1 || a.w < 1 || !a.part || !a.dw || !a.db ||
      a.cin_log < 1 || a.cin_log > a.ca + a.cb)
    return hipErrorInvalidValue;
  const int groups = (a.ca + a.cb) / 32, bands = (a.w + 63) / 64, target = wgrad_target_blocks(a.cout);
  const long segs = (long)bands * a.n * groups;  // workgroups per row of rows
  // rows per workgroup: a multiple of the round height, about `target`
  // workgroups, at most 2 * target
  const int rb = wgrad_rb(a.cout);
  int rows = (int)std::min<long>(a.h, std::max<long>(rb, ((long)a.h * segs + target - 1) / target));
  rows = (rows + rb - 1) / rb * rb;
  long blocks = segs * ((a.h + rows - 1) / rows);
  while (blocks > 2L * target && rows < a.h) {
    rows += rb;
    blocks = segs * ((a.h + rows - 1) / rows);
  }
  const size_t per = (size_t)9 * 32 * a.cout;
  const int slots = (int)(blocks / groups);
  if (((size_t)blocks * per + (size_t)slots * a.cout) * 4 > a.part_bytes) return hipErrorInvalidValue;
  a.rows = rows;
  a.part_b = a.part + (size_t)blocks * per;
  // 128 output channels: the 64-channel kernel on two channel slices (two
  // workgroups per pixel block; the 128-channel accumulators would leave one
  // wave per SIMD)
  const int halves = a.cout == 128 ? 2 : 1;
  dim3 grid(bands, (a.h + rows - 1) / rows, a.n * groups * halves), block(256);
  switch (a.cout / 16) {
    case 2: hipLaunchKernelGGL((conv3x3_wgrad_kernel<2, kWgradRB2>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((conv3x3_wgrad_kernel<4, kWgradRB4>), grid, block, 0, s, a); break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int per4 = (int)(per / 4);
  const size_t total = (size_t)groups * per4;
  const f32x4 *part = (const f32x4 *)a.part;
  auto launch = [&](auto kern, int S) {
    const int nwb = (int)((total + 256 / S - 1) / (256 / S));
    hipLaunchKernelGGL(kern, dim3((unsigned)(nwb + a.cout)), dim3(256), 0, s, part, slots, per4, groups, a.cout,
                       a.cin_log, nwb, (const float *)a.part_b, a.dw, a.db);
  };
  // about 16 loads per thread
  if (slots <= 128) launch(wgrad_reduce_kernel<8>, 8);
  else if (slots <= 256) launch(wgrad_reduce_kernel<16>, 16);
  else if (slots <= 512) launch(wgrad_reduce_kernel<32>, 32);
  else launch(wgrad_reduce_kernel<64>, 64);
  return hipGetLastError();
}

hipError_t pack_input(const float *x12, size_t npix, bf16 *x32, hipStream_t s) {
  hipLaunchKernelGGL(pack_input_kernel, dim3((unsigned)((npix + 63) / 64)), dim3(256), 0, s, x12, npix, x32);
  return hipGetLastError();
}
hipError_t head_fwd_img(const bf16 *d1, const lgx_image &noisy, const float *wh, const float *bh, int out_pitch,
                        float4 *out, hipStream_t s) {
  const size_t npix = (size_t)noisy.w * noisy.h;
  hipLaunchKernelGGL(head_fwd_img_kernel, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s, d1,
                     (const float4 *)noisy.ptr, noisy.pitch, wh, bh, noisy.w, noisy.h, out_pitch, out);
  return hipGetLastError();
}
hipError_t avgpool(const bf16 *x, int n, int h, int w, int c, bf16 *y, hipStream_t s) {
  hipLaunchKernelGGL(avgpool_kernel, dim3(grid_for((size_t)n * (h / 2) * (w / 2) * (c / 8), 256)), dim3(256), 0, s, x,
                     n, h, w, c, y);
  return hipGetLastError();
}
hipError_t pool_bwd_add(const bf16 *dx_in, const bf16 *dpool, const bf16 *act, int n, int h, int w, int c, bf16 *dx,
                        hipStream_t s) {
  if (c % 8 || (c / 8) & (c / 8 - 1) || h < 2 || w < 2) return hipErrorInvalidValue;
  const unsigned bx = (unsigned)std::min((w * (c / 8) + 255) / 256, 64);
  hipLaunchKernelGGL(pool_bwd_add_kernel, dim3(bx, (unsigned)std::min(n * h, 65535)), dim3(256), 0, s, dx_in, dpool,
                     act, n, h, w, c, dx);
  return hipGetLastError();
}
hipError_t upsample_bwd(const bf16 *dfull, const bf16 *act, int n, int hh, int ww, int c, bf16 *dhalf, hipStream_t s) {
  hipLaunchKernelGGL(upsample_bwd_kernel, dim3(grid_for((size_t)n * hh * ww * (c / 8), 256)), dim3(256), 0, s, dfull,
                     act, n, hh, ww, c, dhalf);
  return hipGetLastError();
}
hipError_t head_fwd(const bf16 *d1, const float *x12, const float *wh, const float *bh, size_t npix, int w,
                    int out_pitch, float4 *out, hipStream_t s) {
  hipLaunchKernelGGL(head_fwd_kernel, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s, d1, x12, wh, bh, npix, w,
                     out_pitch, out);
  return hipGetLastError();
}
hipError_t head_train(const bf16 *d1, const float *x12, const float4 *target, const float *wh, const float *bh,
                      size_t npix, float inv_count, int l2, bf16 *dd1, float *dwh, float *dbh, float *loss,
                      hipStream_t s) {
  // persistent: about two workgroups per CU, so as many atomic sums per head parameter
  const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(2 * cu_count(), (npix + 255) / 256));
  hipLaunchKernelGGL(head_train_kernel, dim3(blocks), dim3(256), 0, s, d1, x12, target, wh, bh, npix, inv_count, l2,
                     dd1, dwh, dbh, loss);
  return hipGetLastError();
}
hipError_t adam_pack(float *p, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
                     float bc1, float bc2, const PackTable &pk, hipStream_t s) {
  hipLaunchKernelGGL(adam_pack_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps, bc1,
                     bc2, pk);
  return hipGetLastError();
}

hipError_t fill(float *p, size_t n, float v, hipStream_t s) {
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}
hipError_t pack_weights(const float *w, int cout, int cin_log, int cin_pad, bf16 *wf, bf16 *wt, hipStream_t s) {
  hipLaunchKernelGGL(pack_weights_kernel, dim3(grid_for((size_t)cout * cin_pad * 9, 256)), dim3(256), 0, s, w, cout,
                     cin_log, cin_pad, wf, wt);
  return hipGetLastError();
}
hipError_t mfma_probe(const float *A, const float *B, float *D, hipStream_t s) {
  hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, s, A, B, D);
  return hipGetLastError();
}

}  // namespace unet
}  // namespace lgx
