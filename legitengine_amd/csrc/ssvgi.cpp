// Pass builders, the SSVGI renderer and its C ABI (lgx_renderer_*).
#include "lgx/ssvgi.h"

#include <cstdio>
#include <cstring>
#include <string>

namespace lgx {

// Profiler colours (legit::Colors names at the reference call sites).
static const uint32_t kNephritis = 0x27ae60, kWisteria = 0x8e44ad, kSunFlower = 0xf1c40f, kGreenSea = 0x16a085,
                      kCarrot = 0xe67e22, kBelizeHole = 0x2980b9;

UnmippedProxy::UnmippedProxy(RenderGraph *rg, int32_t format, uvec2 size, uint32_t usage) : baseSize(size) {
  imageProxy = rg->AddImage(format, 1, 1, size, usage);
  imageViewProxy = rg->AddImageView(imageProxy.Id(), 0, 1, 0, 1);
}

MippedProxy::MippedProxy(RenderGraph *rg, int32_t fmt, uvec2 size, uint32_t usage, uint32_t mips, bool ext0)
    : renderGraph(rg), baseSize(size), format(fmt), externalLevel0(ext0) {
  if (ext0) {
    // level 0 is caller memory: a placeholder descriptor until bound
    lgx_image ph{nullptr, (int32_t)size.x, (int32_t)size.y, (int32_t)size.x, fmt, 0};
    mipImageViewProxies.push_back(rg->AddExternalImageView(ph));
    if (mips > 1) {
      imageProxy = rg->AddImage(fmt, mips - 1, 1, uvec2{size.x / 2, size.y / 2}, usage);
      for (uint32_t m = 0; m + 1 < mips; m++) mipImageViewProxies.push_back(rg->AddImageView(imageProxy.Id(), m, 1, 0, 1));
    }
  } else {
    imageProxy = rg->AddImage(fmt, mips, 1, size, usage);
    for (uint32_t m = 0; m < mips; m++) mipImageViewProxies.push_back(rg->AddImageView(imageProxy.Id(), m, 1, 0, 1));
  }
}

void MippedProxy::BindExternalLevel0(const lgx_image &image) {
  if (!externalLevel0) throw Error(LGX_EINVAL, "MippedProxy::BindExternalLevel0: level 0 is owned by the graph");
  renderGraph->UpdateExternalImageView(mipImageViewProxies[0].Id(), image);
}

static std::vector<RenderGraph::ImageViewProxyId> ids(const std::vector<RenderGraph::ImageViewProxyUnique> &v,
                                                      size_t from = 0) {
  std::vector<RenderGraph::ImageViewProxyId> out;
  for (size_t i = from; i < v.size(); i++) out.push_back(v[i].Id());
  return out;
}

void MipBuilder::BuildMips(RenderGraph *rg, const MippedProxy &proxy, FilterTypes filter) {
  const auto levels = ids(proxy.mipImageViewProxies);
  if (levels.size() < 2) return;
  const int32_t mode = filter == FilterTypes::Avg ? LGX_MIP_AVG : LGX_MIP_MAX;
  rg->AddPass(RenderGraph::ComputePassDesc()
                  .SetInputImages({levels[0]})
                  .SetStorageImages(std::vector<RenderGraph::ImageViewProxyId>(levels.begin() + 1, levels.end()))
                  .SetProfilerInfo(kNephritis, "MipBuilderPass")
                  .SetRecordFunc([levels, mode](RenderGraph::PassContext ctx) {
                    std::vector<lgx_image> d;
                    for (auto id : levels) d.push_back(ctx.GetImageView(id));
                    Check(lgx_mip_build(d.data(), (int32_t)d.size(), mode, ctx.GetStream()), "lgx_mip_build");
                  }));
}

void BlurBuilder::ApplyBlur(RenderGraph *rg, RenderGraph::ImageViewProxyId src, RenderGraph::ImageViewProxyId dst,
                            int radius) {
  ApplyBlurs(rg, {src}, {dst}, {radius});
}

void BlurBuilder::ApplyBlurs(RenderGraph *rg, const std::vector<RenderGraph::ImageViewProxyId> &src,
                             const std::vector<RenderGraph::ImageViewProxyId> &dst, const std::vector<int> &radius) {
  if (src.size() != dst.size() || src.size() != radius.size()) throw Error(LGX_EINVAL, "BlurBuilder::ApplyBlurs");
  if (src.empty()) return;
  for (size_t i = 0; i < src.size(); i++)
    if (!(rg->GetMipSize(src[i], 0) == rg->GetMipSize(dst[i], 0)))  // BlurBuilder.h:16-17
      throw Error(LGX_ESIZE, "BlurBuilder::ApplyBlur: source and destination extents differ");
  rg->AddPass(RenderGraph::ComputePassDesc()
                  .SetInputImages(src)
                  .SetStorageImages(dst)
                  .SetProfilerInfo(kWisteria, "BlurPass")
                  .SetRecordFunc([src, dst, radius](RenderGraph::PassContext ctx) {
                    std::vector<lgx_image> s, d;
                    std::vector<int32_t> r(radius.begin(), radius.end());
                    for (auto id : src) s.push_back(ctx.GetImageView(id));
                    for (auto id : dst) d.push_back(ctx.GetImageView(id));
                    Check(lgx_box_blur(s.data(), d.data(), r.data(), (int32_t)s.size(), ctx.GetStream()),
                          "lgx_box_blur");
                  }));
}

void InterleaveBuilder::Deinterleave(RenderGraph *rg, RenderGraph::ImageViewProxyId interleaved,
                                     RenderGraph::ImageViewProxyId deinterleaved, uvec2 grid) {
  rg->AddPass(RenderGraph::ComputePassDesc()
                  .SetInputImages({interleaved})
                  .SetStorageImages({deinterleaved})
                  .SetProfilerInfo(kBelizeHole, "DeinterleavePass")
                  .SetRecordFunc([interleaved, deinterleaved, grid](RenderGraph::PassContext ctx) {
                    lgx_image a = ctx.GetImageView(interleaved), b = ctx.GetImageView(deinterleaved);
                    Check(lgx_deinterleave(&a, &b, (int32_t)grid.x, (int32_t)grid.y, ctx.GetStream()),
                          "lgx_deinterleave");
                  }));
}

void InterleaveBuilder::Interleave(RenderGraph *rg, RenderGraph::ImageViewProxyId deinterleaved,
                                   RenderGraph::ImageViewProxyId interleaved, uvec2 grid) {
  // the interleave program (interleave.frag), which InterleaveBuilder.h:60-62 meant to bind
  rg->AddPass(RenderGraph::ComputePassDesc()
                  .SetInputImages({deinterleaved})
                  .SetStorageImages({interleaved})
                  .SetProfilerInfo(kBelizeHole, "InterleavePass")
                  .SetRecordFunc([interleaved, deinterleaved, grid](RenderGraph::PassContext ctx) {
                    lgx_image a = ctx.GetImageView(deinterleaved), b = ctx.GetImageView(interleaved);
                    Check(lgx_interleave(&a, &b, (int32_t)grid.x, (int32_t)grid.y, ctx.GetStream()),
                          "lgx_interleave");
                  }));
}

// ---------------------------------------------------------------- renderer

struct SsvgiRenderer::ViewportResources {
  ViewportResources(RenderGraph *rg, const SsvgiConfig &c, uvec2 size)
      : depth(rg, LGX_R32F, size, kUsageStorage, (uint32_t)c.levels, true),
        radiance(rg, c.radiance_format, size, kUsageStorage, (uint32_t)c.levels, true),
        blurredRadiance(rg, c.radiance_format, size, kUsageStorage, (uint32_t)c.levels, true),
        indirectSlices(rg, LGX_RGBA32F, size, kUsageStorage),
        indirect(rg, LGX_RGBA32F, size, kUsageStorage),
        denoised{UnmippedProxy(rg, LGX_RGBA32F, size, kUsageStorage), UnmippedProxy(rg, LGX_RGBA32F, size, kUsageStorage)},
        final_(rg, LGX_RGBA32F, size, kUsageStorage) {
    if (c.sampler == LGX_SAMPLER_TRILINEAR)
      blurredDepth.reset(new MippedProxy(rg, LGX_R32F, size, kUsageStorage, (uint32_t)c.levels, true));
    lgx_image ph{nullptr, (int32_t)size.x, (int32_t)size.y, (int32_t)size.x, 0, 0};
    ph.format = c.normal_format;
    normal = rg->AddExternalImageView(ph);
    ph.format = c.albedo_format;
    albedo = rg->AddExternalImageView(ph);
    lgx_gather_params p = GatherParams(c, nullptr, nullptr);
    size_t ws = lgx_gather_workspace_bytes(&p);
    if (ws == 0) throw Error(LGX_EUNSUPPORTED, "SsvgiRenderer: gather parameters");
    workspace = rg->AddRawBuffer(ws);
    // the workspace depends on the resolution and parameters only: fill it once
    Check(lgx_gather_prepare(&p, rg->BufferPointer(workspace.Id())), "lgx_gather_prepare");
  }
  // The pattern divides the frame: the gather writes slices and the
  // InterleavePass restores the image (SSVGIRenderer.h's structure); else
  // the gather writes the image directly.
  static bool Sliced(const SsvgiConfig &c) {
    const int gy = c.grid_y > 0 ? c.grid_y : c.grid;
    return c.width % c.grid == 0 && c.height % gy == 0;
  }
  static lgx_gather_params GatherParams(const SsvgiConfig &c, const float *view, const float *proj) {
    lgx_gather_params p;
    memset(&p, 0, sizeof(p));
    if (view) memcpy(p.view, view, 64);
    if (proj) memcpy(p.proj, proj, 64);
    p.width = c.width;
    p.height = c.height;
    p.dirs = c.dirs;
    p.max_steps = c.max_steps;
    p.grid = c.grid;
    p.grid_y = c.grid_y;
    p.sampler = c.sampler;
    p.precision = LGX_PREC_F32;
    p.out_layout = Sliced(c) ? LGX_OUT_SLICES : LGX_OUT_INTERLEAVED;
    p.tile = c.tile;
    p.footprint_dirs = c.footprint_dirs;
    return p;
  }
  MippedProxy depth, radiance, blurredRadiance;
  std::unique_ptr<MippedProxy> blurredDepth;
  RenderGraph::ImageViewProxyUnique normal, albedo;
  RenderGraph::BufferProxyUnique workspace;
  UnmippedProxy indirectSlices, indirect;
  UnmippedProxy denoised[2];
  UnmippedProxy final_;
  RenderGraph::ImageViewProxyId output;
};

SsvgiRenderer::SsvgiRenderer(RenderGraph *rg, const SsvgiConfig &cfg) : rg_(rg), cfg_(cfg) {
  RecreateSwapchainResources(uvec2{(uint32_t)cfg.width, (uint32_t)cfg.height});
}

SsvgiRenderer::~SsvgiRenderer() = default;

void SsvgiRenderer::RecreateSwapchainResources(uvec2 size) {
  cfg_.width = (int32_t)size.x;
  cfg_.height = (int32_t)size.y;
  const int gy = cfg_.grid_y > 0 ? cfg_.grid_y : cfg_.grid;
  if (cfg_.levels < 1 || cfg_.levels > LGX_MAX_LEVELS || cfg_.grid < 1 || cfg_.grid > 16 || gy > 16 ||
      cfg_.grid * gy > 64)
    throw Error(LGX_EUNSUPPORTED, "SsvgiRenderer: levels / grid");
  vr_.reset();
  vr_.reset(new ViewportResources(rg_, cfg_, size));
}

void SsvgiRenderer::RenderFrame(const SsvgiFrame &f) {
  ViewportResources &v = *vr_;
  const SsvgiConfig c = cfg_;
  const lgx_gbuffer &g = f.gbuffer;
  v.depth.BindExternalLevel0(g.depth);
  v.radiance.BindExternalLevel0(g.radiance);
  v.blurredRadiance.BindExternalLevel0(g.radiance);  // r = 0 at level 0 is an exact copy (blurLayerBuilder.frag:30-31)
  if (v.blurredDepth) v.blurredDepth->BindExternalLevel0(g.depth);
  rg_->UpdateExternalImageView(v.normal.Id(), g.normal);
  rg_->UpdateExternalImageView(v.albedo.Id(), g.albedo);

  // SSVGIRenderer.h:207-208 mips; Hi-Z max for depth in the Hi-Z point-fetch mode
  mipBuilder_.BuildMips(rg_, v.depth,
                        c.sampler == LGX_SAMPLER_HIZ_POINT ? MipBuilder::FilterTypes::Depth : MipBuilder::FilterTypes::Avg);
  mipBuilder_.BuildMips(rg_, v.radiance, MipBuilder::FilterTypes::Avg);
  // :209-221 pre-blur r = 2 above level 0
  {
    std::vector<int> r(c.levels - 1, 2);
    blurBuilder_.ApplyBlurs(rg_, ids(v.radiance.mipImageViewProxies, 1), ids(v.blurredRadiance.mipImageViewProxies, 1),
                            r);
    if (v.blurredDepth)
      blurBuilder_.ApplyBlurs(rg_, ids(v.depth.mipImageViewProxies, 1), ids(v.blurredDepth->mipImageViewProxies, 1), r);
  }
  // :223-263 IndirectLightPass
  {
    auto dlev = v.blurredDepth ? ids(v.blurredDepth->mipImageViewProxies) : ids(v.depth.mipImageViewProxies);
    auto rlev = ids(v.blurredRadiance.mipImageViewProxies);
    std::vector<RenderGraph::ImageViewProxyId> inputs = dlev;
    inputs.insert(inputs.end(), rlev.begin(), rlev.end());
    inputs.push_back(v.normal.Id());
    auto out = ViewportResources::Sliced(c) ? v.indirectSlices.imageViewProxy.Id() : v.indirect.imageViewProxy.Id();
    auto ws = v.workspace.Id();
    auto normal = v.normal.Id();
    lgx_gather_params p = ViewportResources::GatherParams(c, f.view, f.proj);
    rg_->AddPass(RenderGraph::ComputePassDesc()
                     .SetInputImages(inputs)
                     .SetStorageImages({out})
                     .SetStorageBuffers({ws})
                     .SetProfilerInfo(kSunFlower, "IndirectLightPass")
                     .SetRecordFunc([p, dlev, rlev, out, ws, normal](RenderGraph::PassContext ctx) {
                       lgx_pyramid pyr;
                       memset(&pyr, 0, sizeof(pyr));
                       pyr.n_levels = (int32_t)dlev.size();
                       for (size_t l = 0; l < dlev.size(); l++) {
                         pyr.depth[l] = ctx.GetImageView(dlev[l]);
                         pyr.radiance[l] = ctx.GetImageView(rlev[l]);
                       }
                       lgx_gbuffer gb;
                       memset(&gb, 0, sizeof(gb));
                       gb.depth = pyr.depth[0];
                       gb.normal = ctx.GetImageView(normal);
                       lgx_image o = ctx.GetImageView(out);
                       lgx_gather_params pp = p;
                       Check(lgx_gather(&pp, ctx.GetBuffer(ws), &gb, &pyr, &o, ctx.GetStream()), "lgx_gather");
                     }));
  }
  if (ViewportResources::Sliced(c))
    interleaveBuilder_.Interleave(rg_, v.indirectSlices.imageViewProxy.Id(), v.indirect.imageViewProxy.Id(),
                                  uvec2{(uint32_t)c.grid, (uint32_t)(c.grid_y > 0 ? c.grid_y : c.grid)});
  // :265-302 DenoiserPass
  RenderGraph::ImageViewProxyId cur = v.indirect.imageViewProxy.Id();
  if (c.denoiser == LGX_DENOISER_UNET) {
    if (!c.unet) throw Error(LGX_EINVAL, "SsvgiRenderer: LGX_DENOISER_UNET without an attached network");
    auto depth0 = v.depth.mipImageViewProxies[0].Id();
    auto normal = v.normal.Id(), albedo = v.albedo.Id();
    auto dst = v.denoised[0].imageViewProxy.Id();
    auto src = cur;
    lgx_unet *net = c.unet;
    rg_->AddPass(RenderGraph::ComputePassDesc()
                     .SetInputImages({src, depth0, normal, albedo})
                     .SetStorageImages({dst})
                     .SetProfilerInfo(kGreenSea, "DenoiserPass")
                     .SetRecordFunc([net, src, dst, depth0, normal, albedo](RenderGraph::PassContext ctx) {
                       lgx_gbuffer gb;
                       memset(&gb, 0, sizeof(gb));
                       gb.depth = ctx.GetImageView(depth0);
                       gb.normal = ctx.GetImageView(normal);
                       gb.albedo = ctx.GetImageView(albedo);
                       lgx_image a = ctx.GetImageView(src), o = ctx.GetImageView(dst);
                       Check(lgx_unet_denoise(net, &gb, &a, &o, ctx.GetStream()), "lgx_unet_denoise");
                     }));
    cur = dst;
  } else if (c.denoiser != LGX_DENOISER_NONE) {
    auto depth0 = v.depth.mipImageViewProxies[0].Id();
    auto normal = v.normal.Id();
    const int passes = c.denoiser == LGX_DENOISER_REGRESSION ? 1 : c.denoise_passes;
    for (int i = 0; i < passes; i++) {
      auto dst = v.denoised[i & 1].imageViewProxy.Id();
      auto src = cur;
      lgx_bilateral_params bp{c.denoise_radius, c.depth_sharpness, c.normal_sharpness, 0};
      const int kind = c.denoiser;
      rg_->AddPass(RenderGraph::ComputePassDesc()
                       .SetInputImages({src, depth0, normal})
                       .SetStorageImages({dst})
                       .SetProfilerInfo(kGreenSea, "DenoiserPass")
                       .SetRecordFunc([bp, kind, src, dst, depth0, normal](RenderGraph::PassContext ctx) {
                         lgx_image a = ctx.GetImageView(src), o = ctx.GetImageView(dst);
                         lgx_gbuffer gb;
                         memset(&gb, 0, sizeof(gb));
                         gb.depth = ctx.GetImageView(depth0);
                         gb.normal = ctx.GetImageView(normal);
                         lgx_bilateral_params b = bp;
                         if (kind == LGX_DENOISER_REGRESSION)
                           Check(lgx_regression_denoise(&a, &gb, 2, &o, ctx.GetStream()), "lgx_regression_denoise");
                         else
                           Check(lgx_bilateral(&b, &a, &gb, &o, ctx.GetStream()), "lgx_bilateral");
                       }));
      cur = dst;
    }
  }
  // :304-342 final gathering
  if (c.composite) {
    auto radiance0 = v.radiance.mipImageViewProxies[0].Id();
    auto albedo = v.albedo.Id();
    auto dst = v.final_.imageViewProxy.Id();
    auto src = cur;
    rg_->AddPass(RenderGraph::ComputePassDesc()
                     .SetInputImages({src, radiance0, albedo})
                     .SetStorageImages({dst})
                     .SetProfilerInfo(kCarrot, "GatheringPass")
                     .SetRecordFunc([src, dst, radiance0, albedo](RenderGraph::PassContext ctx) {
                       lgx_gbuffer gb;
                       memset(&gb, 0, sizeof(gb));
                       gb.radiance = ctx.GetImageView(radiance0);
                       gb.albedo = ctx.GetImageView(albedo);
                       lgx_image in = ctx.GetImageView(src), o = ctx.GetImageView(dst);
                       Check(lgx_composite(&gb, &in, &o, ctx.GetStream()), "lgx_composite");
                     }));
    cur = dst;
  }
  v.output = cur;
}

RenderGraph::ImageViewProxyId SsvgiRenderer::Output() const { return vr_->output; }

lgx_image SsvgiRenderer::Intermediate(const char *name, uint32_t mip) const {
  const ViewportResources &v = *vr_;
  auto view_of = [&](RenderGraph::ImageViewProxyId id) { return rg_->DescribeImageView(id, 0); };
  auto level = [&](const MippedProxy &m) {
    if (mip >= m.mipImageViewProxies.size()) throw Error(LGX_EINVAL, "Intermediate: mip");
    return view_of(m.mipImageViewProxies[mip].Id());
  };
  std::string n(name ? name : "");
  if (n == "depth_pyramid") return level(v.depth);
  if (n == "radiance_pyramid") return level(v.radiance);
  if (n == "blurred_radiance") return level(v.blurredRadiance);
  if (n == "blurred_depth" && v.blurredDepth) return level(*v.blurredDepth);
  if (n == "indirect_slices") return view_of(v.indirectSlices.imageViewProxy.Id());
  if (n == "indirect") return view_of(v.indirect.imageViewProxy.Id());
  if (n == "denoised" || n == "final") return view_of(v.output);
  throw Error(LGX_EINVAL, "SsvgiRenderer::Intermediate: unknown name");
}

}  // namespace lgx

// ---------------------------------------------------------------- C ABI

struct lgx_renderer {
  lgx::RenderGraph rg;
  std::unique_ptr<lgx::SsvgiRenderer> r;
  std::string last_error;
};

static int guard(lgx_renderer *r, const std::function<void()> &fn) {
  try {
    fn();
    return LGX_OK;
  } catch (const lgx::Error &e) {
    if (r) r->last_error = e.what();
    return e.code();
  } catch (const std::exception &e) {
    if (r) r->last_error = e.what();
    return LGX_EINVAL;
  }
}

extern "C" int lgx_renderer_create(const lgx_renderer_config *c, lgx_renderer **out) {
  if (!c || !out) return LGX_EINVAL;
  *out = nullptr;
  lgx_renderer *r = new lgx_renderer;
  int rc = guard(r, [&]() {
    lgx::SsvgiConfig s;
    s.width = c->width;
    s.height = c->height;
    s.dirs = c->dirs;
    s.max_steps = c->max_steps;
    s.footprint_dirs = c->footprint_dirs;
    s.levels = c->levels;
    s.grid = c->grid;
    s.grid_y = c->grid_y;
    s.sampler = c->sampler;
    s.denoiser = c->denoiser;
    s.denoise_passes = c->denoise_passes;
    s.denoise_radius = c->denoise_radius;
    s.depth_sharpness = c->depth_sharpness;
    s.normal_sharpness = c->normal_sharpness;
    s.normal_format = c->normal_format;
    s.albedo_format = c->albedo_format;
    s.radiance_format = c->radiance_format;
    s.tile = c->tile;
    s.composite = c->composite != 0;
    r->rg.SetProfiling(c->profile != 0);
    r->rg.SetValidation(c->validate != 0);
    r->r.reset(new lgx::SsvgiRenderer(&r->rg, s));
  });
  if (rc != LGX_OK) {
    delete r;
    return rc;
  }
  *out = r;
  return LGX_OK;
}

static void declare(lgx_renderer *r, const lgx_gbuffer *g, const float *view, const float *proj) {
  if (!g || !view || !proj) throw lgx::Error(LGX_EINVAL, "lgx_renderer: null frame input");
  lgx::SsvgiFrame f;
  memcpy(f.view, view, 64);
  memcpy(f.proj, proj, 64);
  f.gbuffer = *g;
  try {
    r->r->RenderFrame(f);
  } catch (...) {
    r->rg.DiscardPasses();
    throw;
  }
}

extern "C" int lgx_renderer_render(lgx_renderer *r, const lgx_gbuffer *g, const float view[16], const float proj[16],
                                   lgx_stream_t stream) {
  if (!r) return LGX_EINVAL;
  return guard(r, [&]() {
    declare(r, g, view, proj);
    r->rg.Execute(stream);
  });
}

extern "C" int lgx_renderer_capture(lgx_renderer *r, const lgx_gbuffer *g, const float view[16],
                                    const float proj[16], lgx_stream_t stream) {
  if (!r) return LGX_EINVAL;
  return guard(r, [&]() {
    declare(r, g, view, proj);
    r->rg.Capture(stream);
  });
}

extern "C" int lgx_renderer_replay(lgx_renderer *r, lgx_stream_t stream) {
  if (!r) return LGX_EINVAL;
  return guard(r, [&]() { r->rg.Replay(stream); });
}

extern "C" int lgx_renderer_attach_unet(lgx_renderer *r, lgx_unet *u) {
  if (!r || !u) return LGX_EINVAL;
  if (r->r->Config().denoiser != LGX_DENOISER_UNET) return LGX_EINVAL;
  r->r->AttachUnet(u);
  return LGX_OK;
}

extern "C" int lgx_renderer_intermediate(lgx_renderer *r, const char *name, int32_t mip, lgx_image *out) {
  if (!r || !out || mip < 0) return LGX_EINVAL;
  return guard(r, [&]() { *out = r->r->Intermediate(name, (uint32_t)mip); });
}

extern "C" int lgx_renderer_pass_times(lgx_renderer *r, int32_t mode, char *buf, size_t cap) {
  if (!r || !buf || cap == 0 || mode < 0 || mode > 2) return LGX_EINVAL;
  return guard(r, [&]() {
    std::string s = "[";
    if (mode == 2) {
      r->rg.ResetGpuProfilerData();
      buf[0] = 0;
      return;
    }
    const auto t = mode == 0 ? r->rg.GetLastFrameGpuProfilerData() : r->rg.GetAccumulatedGpuProfilerData();
    for (size_t i = 0; i < t.size(); i++) {
      char item[160];
      snprintf(item, sizeof(item), "%s{\"name\":\"%s\",\"ms\":%.6f}", i ? "," : "", t[i].name.c_str(), t[i].ms);
      s += item;
    }
    s += "]";
    if (s.size() + 1 > cap) throw lgx::Error(LGX_ESIZE, "lgx_renderer_pass_times: buffer too small");
    memcpy(buf, s.c_str(), s.size() + 1);
  });
}

extern "C" const char *lgx_renderer_last_error(const lgx_renderer *r) {
  return r ? r->last_error.c_str() : "null renderer";
}

extern "C" void lgx_renderer_destroy(lgx_renderer *r) { delete r; }
