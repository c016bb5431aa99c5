// U-Net denoiser: layer graph, training step and C ABI (lgx_unet_*).
//
// Three levels, 3x3 convolutions with 32 / 64 / 128 channels, ReLU, 2x2
// average pooling down, nearest x2 upsampling up with concatenated skips,
// and a 1x1 head that predicts a correction added to the noisy input:
//
//   x (9 ch, padded to 32) -> e1a -> e1 (32) -> pool -> e2a -> e2 (64) -> pool
//   -> b1 -> b2 (128) -> [up(b2), e2] (192) -> d2a -> d2 (64)
//   -> [up(d2), e1] (96) -> d1a -> d1 (32) -> head (1x1, 3) + noisy irradiance
//
// Activations and weights BF16, accumulation FP32 (MFMA), FP32 master
// weights updated by Adam.  113,280 multiply-adds per pixel (the padded first
// layer adds 6,624).  Input record per pixel: noisy irradiance rgb, radial
// depth, octahedral normal xy, albedo rgb.
#include <math.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "lgx_internal.h"
#include "unet.h"

namespace lgx {
namespace unet {

struct LayerSpec {
  int cin, cin_pad, cout;
};
static const LayerSpec kLayers[10] = {
    {9, 32, 32},    {32, 32, 32},  // e1a, e1
    {32, 32, 64},   {64, 64, 64},  // e2a, e2
    {64, 64, 128},  {128, 128, 128},  // b1, b2
    {192, 192, 64}, {64, 64, 64},  // d2a, d2
    {96, 96, 32},   {32, 32, 32},  // d1a, d1
};
static const int kHeadIn = 32, kHeadOut = 3;

static size_t layer_params(int l) { return (size_t)kLayers[l].cout * kLayers[l].cin * 9 + kLayers[l].cout; }
static size_t total_params() {
  size_t n = 0;
  for (int l = 0; l < 10; l++) n += layer_params(l);
  return n + kHeadIn * kHeadOut + kHeadOut;
}

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

class Net {
 public:
  int max_n, max_h, max_w;
  std::vector<std::unique_ptr<DevBuf>> owned;
  float *params, *grads, *adam_m, *adam_v;
  size_t np;
  size_t w_off[10], b_off[10], head_w_off, head_b_off;
  bf16 *wf[10], *wt[10];
  // activations
  bf16 *x32, *e1a, *e1, *p1, *e2a, *e2, *p2, *b1, *b2, *d2a, *d2, *d1a, *d1;
  float *out, *x12_scratch, *loss;
  float *wpart;  // weight-gradient partial sums
  size_t wpart_bytes;
  int loss_kind = LGX_UNET_LOSS_L1;
  // gradients of activations
  bf16 *g_d1, *g_d1a, *g_upd2, *g_e1c, *g_d2, *g_d2a, *g_upb2, *g_e2c, *g_b2, *g_b1, *g_p2, *g_e2, *g_e2a, *g_p1, *g_e1,
      *g_e1a;
  int adam_t = 0;
  // register tile of each layer's forward and input-gradient convolution
  // (an index below conv3x3_tiles(cout)); set by autotune()
  int fwd_tile[10] = {}, dx_tile[10] = {};

  void *alloc(size_t bytes) {
    std::unique_ptr<DevBuf> b(new DevBuf);
    b->bytes = (bytes + 255) / 256 * 256;
    if (hipMalloc(&b->p, b->bytes) != hipSuccess) {
      (void)hipGetLastError();
      throw Error(LGX_EINVAL, "lgx_unet: hipMalloc failed");
    }
    void *p = b->p;
    owned.push_back(std::move(b));
    return p;
  }
  struct Error {
    int code;
    const char *what;
    Error(int c, const char *w) : code(c), what(w) {}
  };

  Net(int n, int h, int w) : max_n(n), max_h(h), max_w(w) {
    np = total_params();
    size_t off = 0;
    for (int l = 0; l < 10; l++) {
      w_off[l] = off;
      off += (size_t)kLayers[l].cout * kLayers[l].cin * 9;
      b_off[l] = off;
      off += kLayers[l].cout;
    }
    head_w_off = off;
    off += kHeadIn * kHeadOut;
    head_b_off = off;
    params = (float *)alloc(np * 4);
    grads = (float *)alloc((np + 1) * 4);  // and the loss after the head's gradient
    adam_m = (float *)alloc(np * 4);
    adam_v = (float *)alloc(np * 4);
    for (int l = 0; l < 10; l++) {
      const LayerSpec &L = kLayers[l];
      wf[l] = (bf16 *)alloc((size_t)9 * L.cin_pad * L.cout * 2);
      wt[l] = l == 0 ? nullptr : (bf16 *)alloc((size_t)9 * L.cin_pad * L.cout * 2);
    }
    const size_t P = (size_t)n * h * w, P2 = P / 4, P4 = P / 16;
    auto A = [&](size_t pix, int c) { return (bf16 *)alloc(pix * c * 2); };
    x32 = A(P, 32);
    e1a = A(P, 32);
    e1 = A(P, 32);
    p1 = A(P2, 32);
    e2a = A(P2, 64);
    e2 = A(P2, 64);
    p2 = A(P4, 64);
    b1 = A(P4, 128);
    b2 = A(P4, 128);
    d2a = A(P2, 64);
    d2 = A(P2, 64);
    d1a = A(P, 32);
    d1 = A(P, 32);
    out = (float *)alloc(P * 16);
    x12_scratch = (float *)alloc((size_t)h * w * 48);
    loss = grads + np;
    wpart_bytes = conv3x3_wgrad_scratch_bytes();
    wpart = (float *)alloc(wpart_bytes);
    g_d1 = A(P, 32);
    g_d1a = A(P, 32);
    g_upd2 = A(P, 64);
    g_e1c = A(P, 32);
    g_d2 = A(P2, 64);
    g_d2a = A(P2, 64);
    g_upb2 = A(P2, 128);
    g_e2c = A(P2, 64);
    g_b2 = A(P4, 128);
    g_b1 = A(P4, 128);
    g_p2 = A(P4, 64);
    g_e2 = A(P2, 64);
    g_e2a = A(P2, 64);
    g_p1 = A(P2, 32);
    g_e1 = A(P, 32);
    g_e1a = A(P, 32);
  }

  void check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error((int)e, what);
    }
  }

  void pack(hipStream_t s) {
    for (int l = 0; l < 10; l++)
      check(pack_weights(params + w_off[l], kLayers[l].cout, kLayers[l].cin, kLayers[l].cin_pad, wf[l], wt[l], s),
            "pack_weights");
  }

  void conv(int l, const bf16 *a, int ca, int a_up, const bf16 *b, int cb, int n, int h, int w, bf16 *out_,
            hipStream_t s, bf16 *pool_out = nullptr) {
    ConvArgs c{};
    c.a = a;
    c.ca = ca;
    c.a_up = a_up;
    c.b = b;
    c.cb = cb;
    c.n = n;
    c.h = h;
    c.w = w;
    c.wt = wf[l];
    c.bias = params + b_off[l];
    c.cout = kLayers[l].cout;
    c.relu = 1;
    c.out0 = out_;
    c.split = c.cout;
    c.tile = fwd_tile[l];
    c.pool_out = pool_out;
    check(conv3x3(c, s), "conv3x3");
  }

  // convolution l followed by the 2x2 average pool of its output into pool:
  // in the convolution's epilogue when its tile offers that, else a pass of its own
  void conv_pool(int l, const bf16 *a, int ca, int n, int h, int w, bf16 *out_, bf16 *pool, hipStream_t s) {
    const int cout = kLayers[l].cout;
    if (conv3x3_tile_pool(cout, fwd_tile[l])) {
      conv(l, a, ca, 0, nullptr, 0, n, h, w, out_, s, pool);
    } else {
      conv(l, a, ca, 0, nullptr, 0, n, h, w, out_, s);
      check(avgpool(out_, n, h, w, cout, pool, s), "avgpool");
    }
  }

  // input gradient of layer l: dy (cout channels) -> [split | rest] with masks
  void conv_dx(int l, const bf16 *dy, int n, int h, int w, bf16 *o0, int split, bf16 *o1, const bf16 *m0,
               const bf16 *m1, hipStream_t s, bf16 *up_out = nullptr, const bf16 *up_mask = nullptr) {
    ConvArgs c{};
    c.up_out = up_out;
    c.up_mask = up_mask;
    c.a = dy;
    c.ca = kLayers[l].cout;
    c.n = n;
    c.h = h;
    c.w = w;
    c.wt = wt[l];
    c.cout = kLayers[l].cin_pad;
    c.out0 = o0;
    c.split = split;
    c.out1 = o1;
    c.mask0 = m0;
    c.mask1 = m1;
    c.tile = dx_tile[l];
    check(conv3x3(c, s), "conv3x3 input gradient");
  }

  void wgrad(int l, const bf16 *a, int ca, int a_up, const bf16 *b, int cb, const bf16 *dy, int n, int h, int w,
             hipStream_t s) {
    WgradArgs g{};
    g.a = a;
    g.ca = ca;
    g.a_up = a_up;
    g.b = b;
    g.cb = cb;
    g.dy = dy;
    g.cout = kLayers[l].cout;
    g.n = n;
    g.h = h;
    g.w = w;
    g.cin_log = kLayers[l].cin;
    g.dw = grads + w_off[l];
    g.db = grads + b_off[l];
    g.part = wpart;
    g.part_bytes = wpart_bytes;
    check(conv3x3_wgrad(g, s), "conv3x3_wgrad");
  }

  // forward convolution l on the activations of an n x h x w input
  void fwd_layer(int l, int n, int h, int w, hipStream_t s) {
    const int h2 = h / 2, w2 = w / 2, h4 = h / 4, w4 = w / 4;
    switch (l) {
      case 0: conv(0, x32, 32, 0, nullptr, 0, n, h, w, e1a, s); break;
      case 1: conv_pool(1, e1a, 32, n, h, w, e1, p1, s); break;  // and p1 = pool(e1)
      case 2: conv(2, p1, 32, 0, nullptr, 0, n, h2, w2, e2a, s); break;
      case 3: conv_pool(3, e2a, 64, n, h2, w2, e2, p2, s); break;  // and p2 = pool(e2)
      case 4: conv(4, p2, 64, 0, nullptr, 0, n, h4, w4, b1, s); break;
      case 5: conv(5, b1, 128, 0, nullptr, 0, n, h4, w4, b2, s); break;
      case 6: conv(6, b2, 128, 1, e2, 64, n, h2, w2, d2a, s); break;  // [up(b2), e2]
      case 7: conv(7, d2a, 64, 0, nullptr, 0, n, h2, w2, d2, s); break;
      case 8: conv(8, d2, 64, 1, e1, 32, n, h, w, d1a, s); break;  // [up(d2), e1]
      default: conv(9, d1a, 32, 0, nullptr, 0, n, h, w, d1, s); break;
    }
  }

  // input gradient of convolution l (1..9) in the backward pass
  void dx_layer(int l, int n, int h, int w, hipStream_t s) {
    const int h2 = h / 2, w2 = w / 2, h4 = h / 4, w4 = w / 4;
    switch (l) {
      case 9: conv_dx(9, g_d1, n, h, w, g_d1a, 32, nullptr, d1a, nullptr, s); break;
      case 8:  // [up(d2), e1]: the gradient of d2 through the upsample, in the epilogue when the tile allows
        if (conv3x3_tile_pool(kLayers[8].cin_pad, dx_tile[8])) {
          conv_dx(8, g_d1a, n, h, w, nullptr, 64, g_e1c, nullptr, nullptr, s, g_d2, d2);
        } else {
          conv_dx(8, g_d1a, n, h, w, g_upd2, 64, g_e1c, nullptr, nullptr, s);
          check(upsample_bwd(g_upd2, d2, n, h2, w2, 64, g_d2, s), "upsample_bwd");
        }
        break;
      case 7: conv_dx(7, g_d2, n, h2, w2, g_d2a, 64, nullptr, d2a, nullptr, s); break;
      case 6:  // [up(b2), e2]
        if (conv3x3_tile_pool(kLayers[6].cin_pad, dx_tile[6])) {
          conv_dx(6, g_d2a, n, h2, w2, nullptr, 128, g_e2c, nullptr, nullptr, s, g_b2, b2);
        } else {
          conv_dx(6, g_d2a, n, h2, w2, g_upb2, 128, g_e2c, nullptr, nullptr, s);
          check(upsample_bwd(g_upb2, b2, n, h4, w4, 128, g_b2, s), "upsample_bwd");
        }
        break;
      case 5: conv_dx(5, g_b2, n, h4, w4, g_b1, 128, nullptr, b1, nullptr, s); break;
      case 4: conv_dx(4, g_b1, n, h4, w4, g_p2, 64, nullptr, nullptr, nullptr, s); break;
      case 3: conv_dx(3, g_e2, n, h2, w2, g_e2a, 64, nullptr, e2a, nullptr, s); break;
      case 2: conv_dx(2, g_e2a, n, h2, w2, g_p1, 32, nullptr, nullptr, nullptr, s); break;
      default: conv_dx(1, g_e1, n, h, w, g_e1a, 32, nullptr, e1a, nullptr, s); break;
    }
  }

  // forward convolutions [from, upto) from x32 towards d1 (layers 1 and 3 also pool)
  void body(int n, int h, int w, hipStream_t s, int upto = 10, int from = 0) {
    for (int l = from; l < upto; l++) fwd_layer(l, n, h, w, s);
  }

  // the first two convolutions in one kernel (conv_first_pair): first-layer
  // inputs from the records (x12) or, for denoise, the G-buffer; its output
  // to e1a when kept (the training forward: the backward pass reads it)
  void first_pair(ConvArgs &c0, int n, int h, int w, bf16 *keep, hipStream_t s) {
    c0.ca = 32;
    c0.n = n;
    c0.h = h;
    c0.w = w;
    c0.wt = wf[0];
    c0.bias = params + b_off[0];
    c0.cout = 32;
    c0.relu = 1;
    c0.split = 32;
    c0.out0 = keep;
    ConvArgs c1{};
    c1.ca = 32;
    c1.n = n;
    c1.h = h;
    c1.w = w;
    c1.wt = wf[1];
    c1.bias = params + b_off[1];
    c1.cout = 32;
    c1.relu = 1;
    c1.out0 = e1;
    c1.split = 32;
    c1.pool_out = p1;
    check(conv_first_pair(c0, c1, s), "conv_first_pair");
  }

  // head: also the 1x1 head into out_ (the training step forms the head in
  // its backward pass instead)
  void forward(const float *x12, int n, int h, int w, float *out_, hipStream_t s, int out_pitch = 0,
               bool head = true) {
    const size_t P = (size_t)n * h * w;
    ConvArgs c0{};
    c0.gin_fmt = 3;
    c0.gin_x12 = x12;
    c0.rec_out = x32;  // the packed records: the first layer's weight gradient reads them
    first_pair(c0, n, h, w, e1a, s);
    body(n, h, w, s, 10, 2);
    if (!head) return;
    check(head_fwd(d1, x12, params + head_w_off, params + head_b_off, P, w, out_pitch > 0 ? out_pitch : w,
                   (float4 *)out_, s),
          "head_fwd");
  }

  // An LDS tile of a 32-channel convolution: the autotuned one for layer l, else the first offered.
  int lds_tile32(int l) const {
    if (conv3x3_tile_lds(32, fwd_tile[l])) return fwd_tile[l];
    for (int t = 0; t < conv3x3_tiles(32); t++)
      if (conv3x3_tile_lds(32, t)) return t;
    return 0;
  }

  // Inference straight from a G-buffer and the noisy image (the values of
  // forward on the records they give): the first two convolutions in one
  // kernel that packs the records from the G-buffer while staging them,
  // then layers 2-8, then the last convolution with the head in its epilogue.
  void denoise(const lgx_image &noisy, const lgx_gbuffer &g, const lgx_image &out_, hipStream_t s) {
    const int h = noisy.h, w = noisy.w;
    {
      ConvArgs c{};
      c.gin_fmt = g.normal.format == LGX_RG16F ? 1 : 2;
      c.gin_noisy = (const float4 *)noisy.ptr;
      c.gin_noisy_pitch = noisy.pitch;
      c.gin_depth = (const float *)g.depth.ptr;
      c.gin_depth_pitch = g.depth.pitch;
      c.gin_normal = g.normal.ptr;
      c.gin_normal_pitch = g.normal.pitch;
      c.gin_albedo = g.albedo.ptr;
      c.gin_albedo_pitch = g.albedo.pitch;
      first_pair(c, 1, h, w, nullptr, s);
    }
    body(1, h, w, s, 9, 2);
    // the last convolution with the 1x1 head fused into its epilogue (an LDS
    // tile over all 32 channels: the autotuned one, else the first offered)
    ConvArgs c{};
    c.a = d1a;
    c.ca = 32;
    c.n = 1;
    c.h = h;
    c.w = w;
    c.wt = wf[9];
    c.bias = params + b_off[9];
    c.cout = 32;
    c.relu = 1;
    c.split = 32;
    c.tile = lds_tile32(9);
    c.head_w = params + head_w_off;
    c.head_b = params + head_b_off;
    c.head_noisy = (const float4 *)noisy.ptr;
    c.head_noisy_pitch = noisy.pitch;
    c.head_out = (float4 *)out_.ptr;
    c.head_out_pitch = out_.pitch;
    check(conv3x3(c, s), "conv3x3 with head");
  }

  // Times every register tile of every forward (and, with backward, every
  // input-gradient) convolution at this shape on the network's own buffers
  // and keeps the fastest; returns the table as JSON.
  std::string autotune(int n, int h, int w, bool with_backward, hipStream_t s) {
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    std::string js = "{\"forward\": [";
    auto run = [&](int *tiles, int l, bool dx) {
      const int cout = dx ? kLayers[l].cin_pad : kLayers[l].cout;
      const int nt = conv3x3_tiles(cout);
      float best = 1e30f;
      int bt = 0;
      std::string row = "[";
      for (int t = 0; t < nt; t++) {
        tiles[l] = t;
        auto launch = [&]() {
          if (dx)
            dx_layer(l, n, h, w, s);
          else
            fwd_layer(l, n, h, w, s);
        };
        launch();
        check(hipEventRecord(e0, s), "hipEventRecord");
        for (int r = 0; r < 3; r++) launch();
        check(hipEventRecord(e1, s), "hipEventRecord");
        check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0;
        check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        ms /= 3;
        char b[48];
        snprintf(b, sizeof(b), "%s%.5f", t ? ", " : "", ms);
        row += b;
        if (ms < best) {
          best = ms;
          bt = t;
        }
      }
      tiles[l] = bt;
      char b[64];
      snprintf(b, sizeof(b), "], \"best\": %d}", bt);
      return "{\"layer\": " + std::to_string(l) + ", \"ms\": " + row + b;
    };
    for (int l = 0; l < 10; l++) js += (l ? ", " : "") + run(fwd_tile, l, false);
    js += "], \"input_gradient\": [";
    if (with_backward)
      for (int l = 1; l < 10; l++) js += (l > 1 ? ", " : "") + run(dx_tile, l, true);
    js += "]}";
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return js;
  }

  void backward(const float *x12, const float *target, int n, int h, int w, hipStream_t s) {
    const size_t P = (size_t)n * h * w;
    const int h2 = h / 2, w2 = w / 2, h4 = h / 4, w4 = w / 4;
    forward(x12, n, h, w, out, s, 0, false);
    // the head's gradient and the loss are accumulated atomically; every
    // convolution gradient is overwritten by its reduction
    check(fill(grads + head_w_off, (size_t)kHeadIn * kHeadOut + kHeadOut + 1, 0.f, s), "fill");  // and the loss
    check(head_train(d1, x12, (const float4 *)target, params + head_w_off, params + head_b_off, P, 1.0f / (3.0f * P),
                     loss_kind == LGX_UNET_LOSS_L2, g_d1, grads + head_w_off, grads + head_b_off, loss, s),
          "head_train");
    // d1 = relu(conv9(d1a))
    wgrad(9, d1a, 32, 0, nullptr, 0, g_d1, n, h, w, s);
    dx_layer(9, n, h, w, s);
    // d1a = relu(conv8([up(d2), e1]))
    wgrad(8, d2, 64, 1, e1, 32, g_d1a, n, h, w, s);
    dx_layer(8, n, h, w, s);  // and the gradient of d2 through the upsample
    // d2 = relu(conv7(d2a))
    wgrad(7, d2a, 64, 0, nullptr, 0, g_d2, n, h2, w2, s);
    dx_layer(7, n, h, w, s);
    // d2a = relu(conv6([up(b2), e2]))
    wgrad(6, b2, 128, 1, e2, 64, g_d2a, n, h2, w2, s);
    dx_layer(6, n, h, w, s);  // and the gradient of b2 through the upsample
    // b2 = relu(conv5(b1)), b1 = relu(conv4(p2))
    wgrad(5, b1, 128, 0, nullptr, 0, g_b2, n, h4, w4, s);
    dx_layer(5, n, h, w, s);
    wgrad(4, p2, 64, 0, nullptr, 0, g_b1, n, h4, w4, s);
    dx_layer(4, n, h, w, s);
    // e2 feeds the pool and the skip
    check(pool_bwd_add(g_e2c, g_p2, e2, n, h2, w2, 64, g_e2, s), "pool_bwd_add");
    wgrad(3, e2a, 64, 0, nullptr, 0, g_e2, n, h2, w2, s);
    dx_layer(3, n, h, w, s);
    wgrad(2, p1, 32, 0, nullptr, 0, g_e2a, n, h2, w2, s);
    dx_layer(2, n, h, w, s);
    check(pool_bwd_add(g_e1c, g_p1, e1, n, h, w, 32, g_e1, s), "pool_bwd_add");
    wgrad(1, e1a, 32, 0, nullptr, 0, g_e1, n, h, w, s);
    dx_layer(1, n, h, w, s);
    wgrad(0, x32, 32, 0, nullptr, 0, g_e1a, n, h, w, s);
  }

  void adam_step(float lr, hipStream_t s) {
    adam_t++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.f - powf(b1, (float)adam_t), bc2 = 1.f - powf(b2, (float)adam_t);
    PackTable pk{};
    pk.n = 10;
    for (int l = 0; l < 10; l++)
      pk.l[l] = PackLayer{w_off[l], kLayers[l].cout, kLayers[l].cin, kLayers[l].cin_pad, wf[l], wt[l]};
    check(adam_pack(params, grads, adam_m, adam_v, np, lr, b1, b2, eps, bc1, bc2, pk, s), "adam_pack");
  }
};

// He-normal initialisation from SplitMix64 + Box-Muller; the head starts near
// zero so the untrained network passes the noisy input through.
static void init_params(std::vector<float> &p, uint64_t seed, const size_t *w_off, const size_t *b_off,
                        size_t head_w_off) {
  uint64_t s = seed;
  auto next = [&]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  auto uni = [&]() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
  auto normal = [&]() { return sqrt(-2.0 * log(uni())) * cos(6.283185307179586 * uni()); };
  std::fill(p.begin(), p.end(), 0.f);
  for (int l = 0; l < 10; l++) {
    const double std_ = sqrt(2.0 / (kLayers[l].cin * 9));
    const size_t nw = (size_t)kLayers[l].cout * kLayers[l].cin * 9;
    for (size_t i = 0; i < nw; i++) p[w_off[l] + i] = (float)(normal() * std_);
  }
  for (int i = 0; i < kHeadIn * kHeadOut; i++) p[head_w_off + i] = (float)(normal() * 0.01);
  (void)b_off;
}

}  // namespace unet
}  // namespace lgx

using lgx::unet::Net;

struct lgx_unet {
  std::unique_ptr<Net> net;
};

template <class F> static int unet_guard(F fn) {
  try {
    fn();
    return LGX_OK;
  } catch (const Net::Error &e) {
    return e.code;
  } catch (...) {
    return LGX_EINVAL;
  }
}

extern "C" int64_t lgx_unet_param_count(void) { return (int64_t)lgx::unet::total_params(); }

extern "C" int lgx_unet_create(int32_t max_n, int32_t max_h, int32_t max_w, uint64_t seed, lgx_unet **out) {
  if (!out || max_n < 1 || max_h < 4 || max_w < 4 || max_h % 4 || max_w % 4) return LGX_EINVAL;
  *out = nullptr;
  lgx_unet *u = new lgx_unet;
  int rc = unet_guard([&]() {
    u->net.reset(new Net(max_n, max_h, max_w));
    Net &N = *u->net;
    std::vector<float> p(N.np);
    lgx::unet::init_params(p, seed, N.w_off, N.b_off, N.head_w_off);
    N.check(hipMemcpy(N.params, p.data(), N.np * 4, hipMemcpyHostToDevice), "hipMemcpy");
    N.check(lgx::unet::fill(N.adam_m, N.np, 0.f, nullptr), "fill");
    N.check(lgx::unet::fill(N.adam_v, N.np, 0.f, nullptr), "fill");
    N.pack(nullptr);
    N.check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  });
  if (rc != LGX_OK) {
    delete u;
    return rc;
  }
  *out = u;
  return LGX_OK;
}

static bool shape_ok(const lgx_unet *u, int32_t n, int32_t h, int32_t w) {
  return u && u->net && n >= 1 && h >= 4 && w >= 4 && h % 4 == 0 && w % 4 == 0 &&
         (int64_t)n * h * w <= (int64_t)u->net->max_n * u->net->max_h * u->net->max_w && h * w <= u->net->max_h * u->net->max_w;
}

extern "C" int lgx_unet_forward(lgx_unet *u, const float *x12, int32_t n, int32_t h, int32_t w, float *out,
                                lgx_stream_t stream) {
  if (!shape_ok(u, n, h, w) || !x12 || !out) return LGX_EINVAL;
  return unet_guard([&]() { u->net->forward(x12, n, h, w, out, (hipStream_t)stream); });
}

extern "C" int lgx_unet_backward(lgx_unet *u, const float *x12, const float *target, int32_t n, int32_t h, int32_t w,
                                 float *loss_out, lgx_stream_t stream) {
  if (!shape_ok(u, n, h, w) || !x12 || !target) return LGX_EINVAL;
  return unet_guard([&]() {
    Net &N = *u->net;
    N.backward(x12, target, n, h, w, (hipStream_t)stream);
    if (loss_out)
      N.check(hipMemcpyAsync(loss_out, N.loss, 4, hipMemcpyDeviceToDevice, (hipStream_t)stream), "hipMemcpyAsync");
  });
}

extern "C" int lgx_unet_set_loss(lgx_unet *u, int32_t loss) {
  if (!u || !u->net || (loss != LGX_UNET_LOSS_L1 && loss != LGX_UNET_LOSS_L2)) return LGX_EINVAL;
  u->net->loss_kind = loss;
  return LGX_OK;
}

extern "C" int lgx_unet_adam_step(lgx_unet *u, float lr, lgx_stream_t stream) {
  if (!u || !u->net) return LGX_EINVAL;
  return unet_guard([&]() { u->net->adam_step(lr, (hipStream_t)stream); });
}

extern "C" int lgx_unet_train_step(lgx_unet *u, const float *x12, const float *target, int32_t n, int32_t h, int32_t w,
                                   float lr, float *loss_out, lgx_stream_t stream) {
  int rc = lgx_unet_backward(u, x12, target, n, h, w, loss_out, stream);
  if (rc != LGX_OK) return rc;
  return lgx_unet_adam_step(u, lr, stream);
}

extern "C" int lgx_unet_output(lgx_unet *u, float **out) {
  if (!u || !u->net || !out) return LGX_EINVAL;
  *out = u->net->out;
  return LGX_OK;
}

// which: 0 params, 1 grads, 2 adam m, 3 adam v (host copies, synchronous)
extern "C" int lgx_unet_get_state(lgx_unet *u, int32_t which, float *host, int64_t n, int32_t *adam_t) {
  if (!u || !u->net || !host || which < 0 || which > 3 || n != (int64_t)u->net->np) return LGX_EINVAL;
  return unet_guard([&]() {
    Net &N = *u->net;
    float *src = which == 0 ? N.params : which == 1 ? N.grads : which == 2 ? N.adam_m : N.adam_v;
    N.check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    N.check(hipMemcpy(host, src, N.np * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    if (adam_t) *adam_t = N.adam_t;
  });
}

extern "C" int lgx_unet_set_state(lgx_unet *u, int32_t which, const float *host, int64_t n, int32_t adam_t) {
  if (!u || !u->net || !host || which < 0 || which > 3 || which == 1 || n != (int64_t)u->net->np) return LGX_EINVAL;
  return unet_guard([&]() {
    Net &N = *u->net;
    float *dst = which == 0 ? N.params : which == 2 ? N.adam_m : N.adam_v;
    N.check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    N.check(hipMemcpy(dst, host, N.np * 4, hipMemcpyHostToDevice), "hipMemcpy");
    if (adam_t >= 0) N.adam_t = adam_t;
    if (which == 0) N.pack(nullptr);
    N.check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  });
}

extern "C" int lgx_unet_denoise(lgx_unet *u, const lgx_gbuffer *g, const lgx_image *noisy, lgx_image *out,
                                lgx_stream_t stream) {
  if (!u || !u->net || !g || !noisy || !out) return LGX_EINVAL;
  // every image the packing and the fused head touch, checked before any launch
  using lgx::image_ok;
  if (!image_ok(noisy) || !image_ok(out) || !image_ok(&g->depth) || !image_ok(&g->normal) || !image_ok(&g->albedo))
    return LGX_EINVAL;
  if (noisy->format != LGX_RGBA32F || out->format != LGX_RGBA32F || g->depth.format != LGX_R32F) return LGX_EFORMAT;
  if (!((g->normal.format == LGX_RG16F && g->albedo.format == LGX_RGBA8) ||
        (g->normal.format == LGX_RG32F && g->albedo.format == LGX_RGBA32F)))
    return LGX_EFORMAT;
  const lgx_image *all[4] = {out, &g->depth, &g->normal, &g->albedo};
  for (const lgx_image *im : all)
    if (im->w != noisy->w || im->h != noisy->h) return LGX_ESIZE;
  if (!shape_ok(u, 1, noisy->h, noisy->w)) return LGX_ESIZE;
  return unet_guard([&]() {
    Net &N = *u->net;
    hipStream_t s = (hipStream_t)stream;
    N.denoise(*noisy, *g, *out, s);
  });
}

extern "C" int lgx_unet_conv_tiles(int32_t cout) { return lgx::unet::conv3x3_tiles(cout); }

extern "C" int lgx_unet_set_tiles(lgx_unet *u, const int32_t *fwd, const int32_t *dx) {
  if (!u || !u->net || !fwd || !dx) return LGX_EINVAL;
  for (int l = 0; l < 10; l++) {
    if (fwd[l] < 0 || fwd[l] >= lgx::unet::conv3x3_tiles(lgx::unet::kLayers[l].cout)) return LGX_EINVAL;
    if (dx[l] < 0 || dx[l] >= lgx::unet::conv3x3_tiles(lgx::unet::kLayers[l].cin_pad)) return LGX_EINVAL;
  }
  for (int l = 0; l < 10; l++) {
    u->net->fwd_tile[l] = fwd[l];
    u->net->dx_tile[l] = dx[l];
  }
  return LGX_OK;
}

extern "C" int lgx_unet_get_tiles(const lgx_unet *u, int32_t *fwd, int32_t *dx) {
  if (!u || !u->net || !fwd || !dx) return LGX_EINVAL;
  for (int l = 0; l < 10; l++) {
    fwd[l] = u->net->fwd_tile[l];
    dx[l] = u->net->dx_tile[l];
  }
  return LGX_OK;
}

extern "C" int lgx_unet_autotune(lgx_unet *u, int32_t n, int32_t h, int32_t w, int32_t backward, lgx_stream_t stream,
                                 char *buf, size_t cap) {
  if (!shape_ok(u, n, h, w)) return LGX_EINVAL;
  return unet_guard([&]() {
    std::string js = u->net->autotune(n, h, w, backward != 0, (hipStream_t)stream);
    if (buf && cap) {
      if (js.size() + 1 > cap) throw Net::Error(LGX_EINVAL, "lgx_unet_autotune: buffer too small");
      memcpy(buf, js.c_str(), js.size() + 1);
    }
  });
}

extern "C" int lgx_unet_mfma_probe(const float *A, const float *B, float *D, lgx_stream_t stream) {
  if (!A || !B || !D) return LGX_EINVAL;
  return (int)lgx::unet::mfma_probe(A, B, D, (hipStream_t)stream);
}

extern "C" void lgx_unet_destroy(lgx_unet *u) { delete u; }
