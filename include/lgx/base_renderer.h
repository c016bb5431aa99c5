// The renderer-facing layer of LegitEngine in its own shape, so that a frame
// loop written against the reference's renderers drives the lgx GI path:
//
//   BaseRenderer         src/Render/Renderers/BaseRenderer.h:6-15
//   SSVGIRenderer        src/Render/Renderers/SSVGIRenderer.h:5-30 (constructor from the core,
//                        RecreateSceneResources / RecreateSwapchainResources), RenderFrame :30-350
//   Camera               src/Scene/Scene.h:20-34
//   InFlightQueue        the frame loop of src/main.cpp:269-312 (BeginFrame -> RenderFrame -> EndFrame,
//                        GetLastFrameGpuProfilerData one frame late)
//
// What differs, and why:
//  - Scene is the seeded analytic scene (lgx_scene): its analytic G-buffer
//    and directional light with hard shadows replace the reference's
//    ShadowPass, GBufferPass and LightPass over OBJ meshes (:61-205).  The
//    `light` camera is accepted and unused: the scene carries its light.
//  - The projection uses the scene's vertical field of view with the
//    reference's perspectiveRH_NO(fov, aspect, 0.01, 1000) * scale(1,-1,-1)
//    (:54-58); the reference hard-codes 1 radian.
//  - FrameInfo carries the HIP stream the frame runs on; the window handle
//    is accepted and unused.
#pragma once

#include <memory>

#include "lgx/ssvgi.h"

struct GLFWwindow;

namespace lgx {

struct Extent2D {  // vk::Extent2D
  uint32_t width = 0, height = 0;
};

// Camera::GetTransformMatrix (Scene.h:30-33): translate(pos) * rotate(horAngle, y) * rotate(vertAngle, x).
struct Camera {
  float pos[3] = {0.f, 0.f, 0.f};
  float vertAngle = 0.f, horAngle = 0.f;
  void GetTransformMatrix(float out[16]) const;  // column-major
};

using Scene = lgx_scene;

// InFlightQueue::FrameInfo (main.cpp:269).
struct FrameInfo {
  ShaderMemoryPool *memoryPool = nullptr;
  size_t frameIndex = 0;
  lgx_stream_t stream = nullptr;
};

// The frame loop's queue: BeginFrame hands out the frame; EndFrame runs the
// passes declared since then on the queue's stream, in declaration order.
class InFlightQueue {
 public:
  InFlightQueue(Core *core, lgx_stream_t stream, size_t inFlightFramesCount = 2);
  FrameInfo BeginFrame();
  void EndFrame();
  // Per-pass GPU times of the last executed frame (needs profiling on the graph).
  const std::vector<RenderGraph::PassTiming> &GetLastFrameGpuProfilerData();
  size_t GetInFlightFramesCount() const { return inFlightFramesCount_; }

 private:
  Core *core_;
  lgx_stream_t stream_;
  size_t inFlightFramesCount_, frameIndex_ = 0;
  bool inFrame_ = false;
  ShaderMemoryPool memoryPool_;
};

class BaseRenderer {
 public:
  BaseRenderer() {}
  virtual ~BaseRenderer() {}
  virtual void RecreateSceneResources(Scene *scene) {}
  virtual void RecreateSwapchainResources(Extent2D viewportExtent, size_t inFlightFramesCount) {}
  virtual void RenderFrame(const FrameInfo &frameInfo, const Camera &camera, const Camera &light, Scene *scene,
                           GLFWwindow *window) {}
  virtual void ReloadShaders() {}
  virtual void ChangeView() {}
};

// The SSVGI renderer behind BaseRenderer: a GBufferPass that renders the
// scene's analytic G-buffer into graph images, then the GI passes of
// SsvgiRenderer (mips / Hi-Z, pre-blur, IndirectLightPass, InterleavePass,
// DenoiserPass, optional GatheringPass).
class SSVGIRenderer : public BaseRenderer {
 public:
  explicit SSVGIRenderer(Core *core, const SsvgiConfig &cfg = SsvgiConfig());
  ~SSVGIRenderer() override;
  void RecreateSceneResources(Scene *scene) override {}
  void RecreateSwapchainResources(Extent2D viewportExtent, size_t inFlightFramesCount) override;
  void RenderFrame(const FrameInfo &frameInfo, const Camera &camera, const Camera &light, Scene *scene,
                   GLFWwindow *window) override;
  // The GI passes (intermediates, attached network) and the frame's output.
  SsvgiRenderer &GetPasses() { return *gi_; }
  lgx_image GetOutput() const;
  // The G-buffer the last frame rendered (graph images).
  lgx_gbuffer GetGBuffer() const;
  // View and projection of the last frame (column-major).
  const float *GetViewMatrix() const { return view_; }
  const float *GetProjMatrix() const { return proj_; }

 private:
  Core *core_;
  SsvgiConfig cfg_;
  std::unique_ptr<UnmippedProxy> depth_, normal_, albedo_, radiance_;
  std::unique_ptr<SsvgiRenderer> gi_;
  float view_[16] = {}, proj_[16] = {};
};

}  // namespace lgx
