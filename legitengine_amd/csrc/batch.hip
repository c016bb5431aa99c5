// lgx_batch: the denoised GI path for many views of uneven resolution,
// issued on one stream and capturable into one hipGraph.
//
// The views, in ascending pixel count, form a few chunks.  Per chunk the
// stream builds the views' mips / Hi-Z / pre-blur (small launches), gathers
// all of the chunk's views in one persistent launch (lgx_gather_batch) and
// then runs the U-Net of each of the chunk's views, smallest first.  Small
// views therefore complete after a fraction of the batch.  (U-Nets on a
// second stream, overlapping the next chunk's gather, waited for the
// persistent gather's workgroups to drain from the CUs and finished near the
// end of the batch.)  A one-thread kernel stamps the device wall clock at the
// start and after each view, which gives per-view latencies from inside a
// graph replay without host round trips.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "lgx.h"
#include "lgx/render_graph.h"
#include "lgx_internal.h"

namespace lgx {

__global__ void stamp_kernel(unsigned long long *ts, int i) {
  if (threadIdx.x == 0) ts[i] = wall_clock64();
}

static void hipc(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error((int)e, std::string(what) + ": " + hipGetErrorString(e));
  }
}

struct BatchView {
  int w = 0, h = 0;
  lgx_gbuffer g{};
  float view[16], proj[16];
  std::vector<lgx_image> dep, dep_b, rad, rad_b;  // level 0 = the G-buffer's own image
  lgx_image noisy{}, denoised{};
  lgx_pyramid pyr{};
};

}  // namespace lgx

using namespace lgx;

struct lgx_batch {
  lgx_batch_config c{};
  std::vector<BatchView> v;
  std::vector<int> order;                 // ascending pixel count
  std::vector<std::vector<int>> chunks;   // consecutive runs of `order`
  std::vector<void *> blocks;
  std::map<std::pair<int, int>, void *> ws;  // gather tables per extent
  std::vector<lgx_gather_batch *> gbs;    // one persistent gather per chunk
  lgx_unet *net = nullptr;
  unsigned long long *ts = nullptr;  // device [n + 1]
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double wall_khz = 0;
  std::string err;

  ~lgx_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (lgx_gather_batch *gb : gbs) lgx_gather_batch_destroy(gb);
    for (void *p : blocks) (void)hipFree(p);
  }
  void *alloc(size_t bytes) {
    void *p = nullptr;
    hipc(hipMalloc(&p, bytes ? bytes : 1), "lgx_batch hipMalloc");
    blocks.push_back(p);
    return p;
  }
  lgx_image image(int w, int h, int fmt) {
    const int eb = fmt == LGX_R32F ? 4 : fmt == LGX_RGBA16F ? 8 : fmt == LGX_RGBA32F ? 16 : 0;
    if (!eb) throw Error(LGX_EFORMAT, "lgx_batch: unsupported image format");
    lgx_image im{alloc((size_t)w * h * eb), w, h, w, fmt, 0};
    return im;
  }
  lgx_gather_params params(const BatchView &bv) const {
    lgx_gather_params p;
    memset(&p, 0, sizeof(p));
    memcpy(p.view, bv.view, sizeof(p.view));
    memcpy(p.proj, bv.proj, sizeof(p.proj));
    p.width = bv.w;
    p.height = bv.h;
    p.dirs = c.dirs;
    p.max_steps = c.max_steps;
    p.grid = c.grid;
    p.grid_y = c.grid_y;
    p.sampler = c.sampler;
    p.precision = LGX_PREC_F32;
    p.out_layout = LGX_OUT_INTERLEAVED;
    p.tile = c.tile;
    p.footprint_dirs = c.footprint_dirs;
    return p;
  }

  void stamp(int i, hipStream_t s) {
    hipLaunchKernelGGL(stamp_kernel, dim3(1), dim3(64), 0, s, ts, i);
    hipc(hipGetLastError(), "stamp_kernel");
  }

  void pyramids(BatchView &bv, hipStream_t s) {
    const lgx_stream_t ls = (lgx_stream_t)s;
    const int L = c.levels;
    if (L < 2) return;
    std::vector<int32_t> radius((size_t)(L - 1), 2);
    Check(lgx_mip_build(bv.rad.data(), L, LGX_MIP_AVG, ls), "lgx_mip_build radiance");
    Check(lgx_mip_build(bv.dep.data(), L, c.sampler == LGX_SAMPLER_HIZ_POINT ? LGX_MIP_MAX : LGX_MIP_AVG, ls),
          "lgx_mip_build depth");
    Check(lgx_box_blur(bv.rad.data() + 1, bv.rad_b.data() + 1, radius.data(), L - 1, ls), "lgx_box_blur radiance");
    if (c.sampler == LGX_SAMPLER_TRILINEAR)
      Check(lgx_box_blur(bv.dep.data() + 1, bv.dep_b.data() + 1, radius.data(), L - 1, ls), "lgx_box_blur depth");
  }

  void enqueue(hipStream_t s) {
    stamp(0, s);
    for (size_t k = 0; k < chunks.size(); k++) {
      for (int i : chunks[k]) pyramids(v[i], s);
      Check(lgx_gather_batch_launch(gbs[k], c.workgroups, (lgx_stream_t)s), "lgx_gather_batch_launch");
      for (int i : chunks[k]) {
        if (net)
          Check(lgx_unet_denoise(net, &v[i].g, &v[i].noisy, &v[i].denoised, (lgx_stream_t)s), "lgx_unet_denoise");
        stamp(1 + i, s);
      }
    }
  }

  // Persistent gathers of every chunk, back to back (the autotuner's unit).
  void launch_gathers(hipStream_t s) {
    for (lgx_gather_batch *gb : gbs)
      Check(lgx_gather_batch_launch(gb, c.workgroups, (lgx_stream_t)s), "lgx_gather_batch_launch");
  }
};

template <class F> static int bguard(lgx_batch *b, F &&f) {
  try {
    f();
    b->err.clear();
    return LGX_OK;
  } catch (const Error &e) {
    b->err = e.what();
    return e.code();
  } catch (const std::exception &e) {
    b->err = e.what();
    return LGX_EINVAL;
  }
}

extern "C" int lgx_batch_create(const lgx_batch_config *c, const lgx_batch_view *views, int32_t n, lgx_unet *net,
                                lgx_batch **out) {
  if (!c || !views || n < 1 || !out) return LGX_EINVAL;
  *out = nullptr;
  if (c->levels < 1 || c->levels > LGX_MAX_LEVELS || c->grid < 1 || c->workgroups < 0) return LGX_EINVAL;
  lgx_batch *b = new (std::nothrow) lgx_batch;
  if (!b) return LGX_EINVAL;
  b->c = *c;
  b->net = net;
  int rc = bguard(b, [&]() {
    b->v.resize((size_t)n);
    for (int i = 0; i < n; i++) {
      const lgx_batch_view &in = views[i];
      BatchView &bv = b->v[i];
      bv.w = in.width;
      bv.h = in.height;
      bv.g = in.gbuffer;
      memcpy(bv.view, in.view, sizeof(bv.view));
      memcpy(bv.proj, in.proj, sizeof(bv.proj));
      if (bv.w < 1 || bv.h < 1 || bv.g.depth.w != bv.w || bv.g.depth.h != bv.h || bv.g.radiance.w != bv.w ||
          bv.g.radiance.h != bv.h)
        throw Error(LGX_ESIZE, "lgx_batch_create: G-buffer extent differs from the view's");
      const int rf = bv.g.radiance.format;
      bv.dep.push_back(bv.g.depth);
      bv.rad.push_back(bv.g.radiance);
      int lw = bv.w, lh = bv.h;
      for (int l = 1; l < c->levels; l++) {
        lw /= 2;
        lh /= 2;
        if (lw < 1 || lh < 1) throw Error(LGX_ESIZE, "lgx_batch_create: too many levels for the extent");
        bv.dep.push_back(b->image(lw, lh, LGX_R32F));
        bv.rad.push_back(b->image(lw, lh, rf));
      }
      bv.rad_b = bv.rad;  // level 0: the pre-blur at r = 0 is the image itself
      bv.dep_b = bv.dep;
      for (int l = 1; l < c->levels; l++) {
        bv.rad_b[l] = b->image(bv.rad[l].w, bv.rad[l].h, rf);
        if (c->sampler == LGX_SAMPLER_TRILINEAR) bv.dep_b[l] = b->image(bv.dep[l].w, bv.dep[l].h, LGX_R32F);
      }
      bv.pyr.n_levels = c->levels;
      for (int l = 0; l < c->levels; l++) {
        bv.pyr.depth[l] = bv.dep_b[l];
        bv.pyr.radiance[l] = bv.rad_b[l];
      }
      bv.noisy = b->image(bv.w, bv.h, LGX_RGBA32F);
      bv.denoised = net ? b->image(bv.w, bv.h, LGX_RGBA32F) : bv.noisy;
      auto key = std::make_pair(bv.w, bv.h);
      if (!b->ws.count(key)) {
        lgx_gather_params p = b->params(bv);
        p.tile = gather_unswizzled_tile(p.tile);  // tables only: the gather batch holds the swizzled copies
        const size_t bytes = lgx_gather_workspace_bytes(&p);
        if (!bytes) throw Error(LGX_EINVAL, "lgx_batch_create: gather parameters");
        void *w = b->alloc(bytes);
        Check(lgx_gather_prepare(&p, w), "lgx_gather_prepare");
        b->ws[key] = w;
      }
    }
    b->order.resize((size_t)n);
    for (int i = 0; i < n; i++) b->order[i] = i;
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
      return (long)b->v[x].w * b->v[x].h < (long)b->v[y].w * b->v[y].h;
    });
    // chunks: consecutive runs of the ascending order, equal view counts
    const int K = std::min(n, c->chunks > 0 ? (int)c->chunks : 4);
    for (int k = 0; k < K; k++) {
      std::vector<int> ch(b->order.begin() + (size_t)k * n / K, b->order.begin() + (size_t)(k + 1) * n / K);
      std::vector<lgx_gather_params> ps;
      std::vector<const void *> wss;
      std::vector<lgx_gbuffer> gs;
      std::vector<lgx_pyramid> pyrs;
      std::vector<lgx_image> outs;
      for (int i : ch) {
        BatchView &bv = b->v[i];
        ps.push_back(b->params(bv));
        wss.push_back(b->ws[std::make_pair(bv.w, bv.h)]);
        gs.push_back(bv.g);
        pyrs.push_back(bv.pyr);
        outs.push_back(bv.noisy);
      }
      lgx_gather_batch *gb = nullptr;
      Check(lgx_gather_batch_create(ps.data(), wss.data(), gs.data(), pyrs.data(), outs.data(), (int32_t)ch.size(),
                                    &gb),
            "lgx_gather_batch_create");
      b->gbs.push_back(gb);
      b->chunks.push_back(ch);
    }
    b->ts = (unsigned long long *)b->alloc(sizeof(unsigned long long) * ((size_t)n + 1));
    int dev = 0, khz = 0;
    hipc(hipGetDevice(&dev), "hipGetDevice");
    hipc(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "hipDeviceGetAttribute");
    if (khz <= 0) throw Error(LGX_ENODEVICE, "lgx_batch_create: no device wall clock rate");
    b->wall_khz = khz;
  });
  if (rc != LGX_OK) {
    fprintf(stderr, "lgx_batch_create: %s\n", b->err.c_str());
    delete b;
    return rc;
  }
  *out = b;
  return LGX_OK;
}

extern "C" int lgx_batch_run(lgx_batch *b, lgx_stream_t stream) {
  if (!b) return LGX_EINVAL;
  return bguard(b, [&]() { b->enqueue((hipStream_t)stream); });
}

extern "C" int lgx_batch_capture(lgx_batch *b, lgx_stream_t stream) {
  if (!b) return LGX_EINVAL;
  if (!stream) return LGX_EINVAL;  // the legacy default stream cannot be captured
  return bguard(b, [&]() {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    hipStream_t s = (hipStream_t)stream;
    hipc(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      b->enqueue(s);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    hipc(hipStreamEndCapture(s, &b->graph), "hipStreamEndCapture");
    hipc(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  });
}

extern "C" int lgx_batch_replay(lgx_batch *b, lgx_stream_t stream) {
  if (!b) return LGX_EINVAL;
  if (!b->exec) return LGX_EINVAL;
  return bguard(b, [&]() { hipc(hipGraphLaunch(b->exec, (hipStream_t)stream), "hipGraphLaunch"); });
}

extern "C" int lgx_batch_autotune(lgx_batch *b, lgx_stream_t stream, char *buf, size_t cap) {
  if (!b) return LGX_EINVAL;
  return bguard(b, [&]() {
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, cus = 0;
    hipc(hipGetDevice(&dev), "hipGetDevice");
    hipc(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    // run the pyramids once so the gather reads this batch's real inputs
    b->enqueue(s);
    hipc(hipStreamSynchronize(s), "hipStreamSynchronize");
    hipEvent_t e0, e1;
    hipc(hipEventCreate(&e0), "hipEventCreate");
    hipc(hipEventCreate(&e1), "hipEventCreate");
    std::string js = "[";
    double best = 1e30;
    int best_tile = b->c.tile, best_wg = b->c.workgroups;
    const int wg_mult[] = {0, 4, 8, 16};
    const int saved_wg = b->c.workgroups;
    try {
      for (int t = 0; t < lgx_gather_tiles(); t++) {
        for (lgx_gather_batch *gb : b->gbs) Check(lgx_gather_batch_set_tile(gb, t), "lgx_gather_batch_set_tile");
        for (int m : wg_mult) {
          const int wg = m * cus;
          b->c.workgroups = wg;
          b->launch_gathers(s);  // warm-up
          std::vector<float> ms;
          for (int r = 0; r < 3; r++) {
            hipc(hipEventRecord(e0, s), "hipEventRecord");
            b->launch_gathers(s);
            hipc(hipEventRecord(e1, s), "hipEventRecord");
            hipc(hipEventSynchronize(e1), "hipEventSynchronize");
            float x = 0;
            hipc(hipEventElapsedTime(&x, e0, e1), "hipEventElapsedTime");
            ms.push_back(x);
          }
          std::sort(ms.begin(), ms.end());
          char row[128];
          snprintf(row, sizeof(row), "%s{\"tile\": %d, \"workgroups\": %d, \"ms\": %.5f}", js.size() > 1 ? ", " : "",
                   t, wg, ms[1]);
          js += row;
          if (ms[1] < best) {
            best = ms[1];
            best_tile = t;
            best_wg = wg;
          }
        }
      }
    } catch (...) {
      b->c.workgroups = saved_wg;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    js += "]";
    for (lgx_gather_batch *gb : b->gbs) Check(lgx_gather_batch_set_tile(gb, best_tile), "lgx_gather_batch_set_tile");
    b->c.tile = best_tile;
    b->c.workgroups = best_wg;
    if (buf && cap) {
      char head[160];
      snprintf(head, sizeof(head), "{\"best\": {\"tile\": %d, \"workgroups\": %d, \"ms\": %.5f}, \"cus\": %d, \"table\": ",
               best_tile, best_wg, best, cus);
      std::string all = std::string(head) + js + "}";
      if (all.size() + 1 > cap) throw Error(LGX_EINVAL, "lgx_batch_autotune: buffer too small");
      memcpy(buf, all.c_str(), all.size() + 1);
    }
  });
}

extern "C" int lgx_batch_latencies(lgx_batch *b, double *ms, int32_t n) {
  if (!b || !ms || n != (int32_t)b->v.size()) return LGX_EINVAL;
  return bguard(b, [&]() {
    std::vector<unsigned long long> t(b->v.size() + 1);
    hipc(hipDeviceSynchronize(), "hipDeviceSynchronize");
    hipc(hipMemcpy(t.data(), b->ts, t.size() * sizeof(t[0]), hipMemcpyDeviceToHost), "hipMemcpy");
    for (size_t i = 0; i < b->v.size(); i++) ms[i] = (double)(t[i + 1] - t[0]) / b->wall_khz;
  });
}

extern "C" int lgx_batch_output(lgx_batch *b, int32_t i, int32_t which, lgx_image *out) {
  if (!b || !out || i < 0 || i >= (int32_t)b->v.size() || which < 0 || which > 1) return LGX_EINVAL;
  *out = which == 0 ? b->v[i].noisy : b->v[i].denoised;
  return LGX_OK;
}

extern "C" int lgx_batch_shape(const lgx_batch *b, int32_t *tile, int32_t *workgroups) {
  if (!b) return LGX_EINVAL;
  if (tile) *tile = b->c.tile;
  if (workgroups) *workgroups = b->c.workgroups;
  return LGX_OK;
}

extern "C" const char *lgx_batch_last_error(const lgx_batch *b) { return b ? b->err.c_str() : "null batch"; }

extern "C" void lgx_batch_destroy(lgx_batch *b) { delete b; }
