// Pass builders and the SSVGI renderer on lgx::RenderGraph.
//
//   UnmippedProxy / MippedProxy   src/Render/Common/MipBuilder.h:3-33
//   MipBuilder::BuildMips         MipBuilder.h:142-181 (FilterTypes :136-140)
//   BlurBuilder::ApplyBlur        src/Render/Common/BlurBuilder.h:14-46
//   InterleaveBuilder             src/Render/Common/InterleaveBuilder.h:14-80
//   SsvgiRenderer                 src/Render/Renderers/SSVGIRenderer.h (RecreateSwapchainResources :24-29,
//                                 RenderFrame :30-350, ViewportResources :391-419)
//
// Differences kept on purpose (each is a pass-count or launch-shape change,
// not a change of results): BuildMips declares one pass that builds up to
// five levels per launch instead of one pass per level; the per-level
// pre-blur of a pyramid is one batched pass; the r = 0 copy of level 0 is an
// alias of level 0 (an exact copy in the reference); the gather writes its
// output per 4x4 interleave slice and an InterleavePass restores the image.
#pragma once

#include <memory>
#include <vector>

#include "lgx/render_graph.h"

namespace lgx {

struct UnmippedProxy {
  UnmippedProxy(RenderGraph *rg, int32_t format, uvec2 size, uint32_t usage);
  RenderGraph::ImageProxyUnique imageProxy;
  RenderGraph::ImageViewProxyUnique imageViewProxy;
  uvec2 baseSize;
};

// A mip chain: one view per level.  Level 0 may be bound to caller memory
// (BindExternalLevel0), e.g. the G-buffer image the pyramid starts from.
struct MippedProxy {
  MippedProxy(RenderGraph *rg, int32_t format, uvec2 size, uint32_t usage, uint32_t mips,
              bool external_level0 = false);
  void BindExternalLevel0(const lgx_image &image);
  RenderGraph *renderGraph;
  RenderGraph::ImageProxyUnique imageProxy;  // levels 0..mips-1 (or 1..mips-1 when level 0 is external)
  std::vector<RenderGraph::ImageViewProxyUnique> mipImageViewProxies;
  uvec2 baseSize;
  int32_t format;
  bool externalLevel0;
};

// The builders take the reference's constructor (legit::Core *) and
// argument lists, memory pool included (MipBuilder.h:142,
// BlurBuilder.h:14, InterleaveBuilder.h:14,48), beside shorter forms
// without the pool.
class MipBuilder {
 public:
  MipBuilder() = default;
  explicit MipBuilder(Core *) {}
  enum struct FilterTypes { Avg, Depth };  // Depth = max-Z reduction
  void BuildMips(RenderGraph *rg, ShaderMemoryPool *memoryPool, const MippedProxy &proxy,
                 FilterTypes filter = FilterTypes::Avg) {
    BuildMips(rg, proxy, filter);
  }
  void BuildMips(RenderGraph *rg, const MippedProxy &proxy, FilterTypes filter = FilterTypes::Avg);
};

class BlurBuilder {
 public:
  BlurBuilder() = default;
  explicit BlurBuilder(Core *) {}
  void ApplyBlur(RenderGraph *rg, ShaderMemoryPool *memoryPool, RenderGraph::ImageViewProxyId src,
                 RenderGraph::ImageViewProxyId dst, int radius) {
    ApplyBlur(rg, src, dst, radius);
  }
  void ApplyBlur(RenderGraph *rg, RenderGraph::ImageViewProxyId src, RenderGraph::ImageViewProxyId dst, int radius);
  // one pass blurring every (src[i], dst[i], radius[i])
  void ApplyBlurs(RenderGraph *rg, const std::vector<RenderGraph::ImageViewProxyId> &src,
                  const std::vector<RenderGraph::ImageViewProxyId> &dst, const std::vector<int> &radius);
};

class InterleaveBuilder {
 public:
  InterleaveBuilder() = default;
  explicit InterleaveBuilder(Core *) {}
  void Deinterleave(RenderGraph *rg, ShaderMemoryPool *memoryPool, RenderGraph::ImageViewProxyId interleaved,
                    RenderGraph::ImageViewProxyId deinterleaved, uvec2 gridSize) {
    Deinterleave(rg, interleaved, deinterleaved, gridSize);
  }
  void Interleave(RenderGraph *rg, ShaderMemoryPool *memoryPool, RenderGraph::ImageViewProxyId deinterleaved,
                  RenderGraph::ImageViewProxyId interleaved, uvec2 gridSize) {
    Interleave(rg, deinterleaved, interleaved, gridSize);
  }
  void Deinterleave(RenderGraph *rg, RenderGraph::ImageViewProxyId interleaved,
                    RenderGraph::ImageViewProxyId deinterleaved, uvec2 gridSize);
  void Interleave(RenderGraph *rg, RenderGraph::ImageViewProxyId deinterleaved,
                  RenderGraph::ImageViewProxyId interleaved, uvec2 gridSize);
};



struct SsvgiConfig {
  int32_t width = 1920, height = 1080;
  int32_t dirs = 32, max_steps = 64, levels = 6, grid = 4;
  int32_t footprint_dirs = 0;  // gather step growth / lod direction count, 0 = dirs
  int32_t grid_y = 0;          // interleave pattern height, 0 = grid (square)
  int32_t sampler = LGX_SAMPLER_HIZ_POINT;
  int32_t denoiser = LGX_DENOISER_BILATERAL;
  int32_t denoise_passes = 2, denoise_radius = 4;
  float depth_sharpness = 10.5f, normal_sharpness = 20.0f;
  int32_t normal_format = LGX_RG16F, albedo_format = LGX_RGBA8, radiance_format = LGX_RGBA16F;
  int32_t tile = 0;
  bool composite = false;
  struct lgx_unet *unet = nullptr;  // LGX_DENOISER_UNET
};

struct SsvgiFrame {
  float view[16], proj[16];
  lgx_gbuffer gbuffer;  // caller's device images for this frame
};

class SsvgiRenderer {
 public:
  SsvgiRenderer(RenderGraph *rg, const SsvgiConfig &cfg);
  ~SsvgiRenderer();
  // Allocates the viewport resources for a new extent (SSVGIRenderer.h:24-29).
  void RecreateSwapchainResources(uvec2 size);
  // Declares the frame's passes (SSVGIRenderer.h:207-302); the caller runs
  // them with RenderGraph::Execute.
  void RenderFrame(const SsvgiFrame &frame);
  // Final image of the frame (denoised irradiance, or the composite).
  RenderGraph::ImageViewProxyId Output() const;
  // Named intermediates for inspection: "depth_pyramid", "radiance_pyramid",
  // "blurred_radiance", "blurred_depth", "indirect_slices", "indirect",
  // "denoised", "final".
  lgx_image Intermediate(const char *name, uint32_t mip) const;
  const SsvgiConfig &Config() const { return cfg_; }
  void AttachUnet(struct lgx_unet *u) { cfg_.unet = u; }

 private:
  struct ViewportResources;
  RenderGraph *rg_;
  SsvgiConfig cfg_;
  MipBuilder mipBuilder_;
  BlurBuilder blurBuilder_;
  InterleaveBuilder interleaveBuilder_;
  std::unique_ptr<ViewportResources> vr_;
};

}  // namespace lgx
