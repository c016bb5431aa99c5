// BaseRenderer-shaped SSVGI renderer, Camera and the in-flight frame queue.
#include "lgx/base_renderer.h"

#include <cmath>
#include <cstring>

namespace lgx {

static const uint32_t kPeterRiver = 0x3498db;  // legit::Colors name for the G-buffer pass

static void mul4(const float *a, const float *b, float *out) {  // column-major a * b
  float r[16];
  for (int c = 0; c < 4; c++)
    for (int row = 0; row < 4; row++) {
      float s = 0.f;
      for (int k = 0; k < 4; k++) s += a[k * 4 + row] * b[c * 4 + k];
      r[c * 4 + row] = s;
    }
  memcpy(out, r, sizeof(r));
}

void Camera::GetTransformMatrix(float out[16]) const {
  const float ch = std::cos(horAngle), sh = std::sin(horAngle), cv = std::cos(vertAngle), sv = std::sin(vertAngle);
  const float t[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, pos[0], pos[1], pos[2], 1};
  const float ry[16] = {ch, 0, -sh, 0, 0, 1, 0, 0, sh, 0, ch, 0, 0, 0, 0, 1};
  const float rx[16] = {1, 0, 0, 0, 0, cv, sv, 0, 0, -sv, cv, 0, 0, 0, 0, 1};
  float a[16];
  mul4(t, ry, a);
  mul4(a, rx, out);
}

InFlightQueue::InFlightQueue(Core *core, lgx_stream_t stream, size_t inFlightFramesCount)
    : core_(core), stream_(stream), inFlightFramesCount_(inFlightFramesCount) {
  if (!core || inFlightFramesCount < 1) throw Error(LGX_EINVAL, "InFlightQueue: core and frame count");
}

FrameInfo InFlightQueue::BeginFrame() {
  if (inFrame_) throw Error(LGX_EINVAL, "InFlightQueue::BeginFrame: the previous frame was not ended");
  inFrame_ = true;
  FrameInfo f;
  f.memoryPool = &memoryPool_;
  f.frameIndex = frameIndex_ % inFlightFramesCount_;
  f.stream = stream_;
  return f;
}

void InFlightQueue::EndFrame() {
  if (!inFrame_) throw Error(LGX_EINVAL, "InFlightQueue::EndFrame: no frame begun");
  inFrame_ = false;
  frameIndex_++;
  try {
    core_->GetRenderGraph()->Execute(stream_);
  } catch (...) {
    core_->GetRenderGraph()->DiscardPasses();
    throw;
  }
}

const std::vector<RenderGraph::PassTiming> &InFlightQueue::GetLastFrameGpuProfilerData() {
  return core_->GetRenderGraph()->GetLastFrameGpuProfilerData();
}

SSVGIRenderer::SSVGIRenderer(Core *core, const SsvgiConfig &cfg) : core_(core), cfg_(cfg) {
  if (!core) throw Error(LGX_EINVAL, "SSVGIRenderer: null core");
}

SSVGIRenderer::~SSVGIRenderer() = default;

void SSVGIRenderer::RecreateSwapchainResources(Extent2D extent, size_t inFlightFramesCount) {
  if (extent.width == 0 || extent.height == 0) throw Error(LGX_ESIZE, "SSVGIRenderer: empty viewport");
  RenderGraph *rg = core_->GetRenderGraph();
  gi_.reset();
  depth_.reset();
  normal_.reset();
  albedo_.reset();
  radiance_.reset();
  const uvec2 size{extent.width, extent.height};
  depth_.reset(new UnmippedProxy(rg, LGX_R32F, size, kUsageStorage));
  normal_.reset(new UnmippedProxy(rg, cfg_.normal_format, size, kUsageStorage));
  albedo_.reset(new UnmippedProxy(rg, cfg_.albedo_format, size, kUsageStorage));
  radiance_.reset(new UnmippedProxy(rg, cfg_.radiance_format, size, kUsageStorage));
  SsvgiConfig c = cfg_;
  c.width = (int32_t)extent.width;
  c.height = (int32_t)extent.height;
  gi_.reset(new SsvgiRenderer(rg, c));
}

lgx_gbuffer SSVGIRenderer::GetGBuffer() const {
  if (!gi_) throw Error(LGX_EINVAL, "SSVGIRenderer: RecreateSwapchainResources was not called");
  RenderGraph *rg = core_->GetRenderGraph();
  lgx_gbuffer g;
  g.depth = rg->DescribeImageView(depth_->imageViewProxy.Id(), 0);
  g.normal = rg->DescribeImageView(normal_->imageViewProxy.Id(), 0);
  g.albedo = rg->DescribeImageView(albedo_->imageViewProxy.Id(), 0);
  g.radiance = rg->DescribeImageView(radiance_->imageViewProxy.Id(), 0);
  return g;
}

void SSVGIRenderer::RenderFrame(const FrameInfo &frameInfo, const Camera &camera, const Camera &light, Scene *scene,
                                GLFWwindow *window) {
  if (!gi_) throw Error(LGX_EINVAL, "SSVGIRenderer: RecreateSwapchainResources was not called");
  if (!scene) throw Error(LGX_EINVAL, "SSVGIRenderer::RenderFrame: null scene");
  RenderGraph *rg = core_->GetRenderGraph();
  const SsvgiConfig &c = gi_->Config();
  // the scene drawn from this frame's camera
  lgx_scene s = *scene;
  memcpy(s.cam_pos, camera.pos, sizeof(s.cam_pos));
  s.cam_hor = camera.horAngle;
  s.cam_vert = camera.vertAngle;
  Check(lgx_scene_matrices(&s, c.width, c.height, view_, proj_), "lgx_scene_matrices");
  const lgx_gbuffer g = GetGBuffer();
  // G-buffer, direct light and hard shadows of the analytic scene
  // (in place of ShadowPass / GBufferPass / LightPass, SSVGIRenderer.h:61-205)
  rg->AddPass(RenderGraph::ComputePassDesc()
                  .SetStorageImages({depth_->imageViewProxy.Id(), normal_->imageViewProxy.Id(),
                                     albedo_->imageViewProxy.Id(), radiance_->imageViewProxy.Id()})
                  .SetProfilerInfo(kPeterRiver, "GBufferPass")
                  .SetRecordFunc([s, g](RenderGraph::PassContext ctx) {
                    Check(lgx_scene_render_gpu(&s, &g, ctx.GetStream()), "lgx_scene_render_gpu");
                  }));
  SsvgiFrame f;
  memcpy(f.view, view_, sizeof(view_));
  memcpy(f.proj, proj_, sizeof(proj_));
  f.gbuffer = g;
  gi_->RenderFrame(f);
}

lgx_image SSVGIRenderer::GetOutput() const {
  if (!gi_) throw Error(LGX_EINVAL, "SSVGIRenderer: RecreateSwapchainResources was not called");
  return core_->GetRenderGraph()->DescribeImageView(gi_->Output(), 0);
}

}  // namespace lgx
