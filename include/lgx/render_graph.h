// lgx::RenderGraph — the compute-pass subset of LegitEngine's render graph
// API, executing on one HIP stream of an MI355X.
//
// Shape follows the reference call sites (the LegitVulkan submodule that
// implements RenderGraph is not vendored; its interface is inferred from its
// callers):
//   AddImage / AddImageView / AddBuffer<T> / AddExternal*   MipBuilder.h:8-26, LSGIRenderer.h:646,
//                                                           MipBuilder.h:65, PointRenderer.h:988
//   GetMipSize(view, mip)                                   BlurBuilder.h:16
//   AddPass(ComputePassDesc()                               MipBuilder.h:196-229, LSGIRenderer.h:268-308
//     .SetInputImages / .SetStorageImages / .SetStorageBuffers
//     .SetProfilerInfo(color, name) / .SetRecordFunc(lambda(PassContext)))
//   PassContext::GetImageView / GetBuffer                   MipBuilder.h:213-216
//   execution at the end of the frame, in declaration order (InFlightQueue::EndFrame, main.cpp:312)
//   per-pass GPU timings one frame late (GetLastFrameGpuProfilerData, main.cpp:273)
//
// MI355X mapping: one in-order HIP stream replaces barriers (passes run in
// declaration order, so every declared read-after-write is ordered); the
// command buffer becomes the stream handed to record funcs; images live in
// HBM with one pitched allocation per mip level; per-pass hipEvents give the
// profiler data; a captured frame can be replayed as a hipGraph.  Errors from
// the C ABI are thrown as lgx::Error, as the reference reports failures by
// exception (main.cpp:314-318).
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "lgx.h"

namespace lgx {

struct uvec2 {
  uint32_t x, y;
  bool operator==(const uvec2 &o) const { return x == o.x && y == o.y; }
};

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string &what);
  int code() const { return code_; }

 private:
  int code_;
};

// Throws lgx::Error when an lgx_* call fails.
void Check(int code, const char *what);

// Usage bits, recorded for validation (the reference's vk::ImageUsageFlags).
enum ImageUsage : uint32_t { kUsageSampled = 1, kUsageStorage = 2 };

class RenderGraph {
 public:
  struct ImageProxyId {
    uint32_t v = UINT32_MAX;
    bool operator==(const ImageProxyId &o) const { return v == o.v; }
  };
  struct ImageViewProxyId {
    uint32_t v = UINT32_MAX;
    bool operator==(const ImageViewProxyId &o) const { return v == o.v; }
  };
  struct BufferProxyId {
    uint32_t v = UINT32_MAX;
    bool operator==(const BufferProxyId &o) const { return v == o.v; }
  };

  // RAII owners (ImageProxyUnique & co.): release the proxy on destruction.
  template <class IdT> class Unique {
   public:
    Unique() = default;
    Unique(RenderGraph *g, IdT id) : g_(g), id_(id) {}
    Unique(Unique &&o) noexcept : g_(o.g_), id_(o.id_) { o.g_ = nullptr; }
    Unique &operator=(Unique &&o) noexcept {
      reset();
      g_ = o.g_;
      id_ = o.id_;
      o.g_ = nullptr;
      return *this;
    }
    Unique(const Unique &) = delete;
    Unique &operator=(const Unique &) = delete;
    ~Unique() { reset(); }
    IdT Id() const { return id_; }
    const Unique *operator->() const { return this; }  // proxy->Id() as in the reference
    void reset();

   private:
    RenderGraph *g_ = nullptr;
    IdT id_{};
  };
  using ImageProxyUnique = Unique<ImageProxyId>;
  using ImageViewProxyUnique = Unique<ImageViewProxyId>;
  using BufferProxyUnique = Unique<BufferProxyId>;

  class PassContext {
   public:
    // The stream record funcs launch on (the reference's GetCommandBuffer()).
    lgx_stream_t GetStream() const { return stream_; }
    lgx_stream_t GetCommandBuffer() const { return stream_; }
    // Descriptor of the view's base mip.
    lgx_image GetImageView(ImageViewProxyId id) const;
    // Descriptor of mip `i` (relative to the view's base mip).
    lgx_image GetImageViewMip(ImageViewProxyId id, uint32_t i) const;
    uint32_t GetImageViewMipCount(ImageViewProxyId id) const;
    void *GetBuffer(BufferProxyId id) const;

   private:
    friend class RenderGraph;
    const RenderGraph *g_ = nullptr;
    size_t pass_ = 0;
    lgx_stream_t stream_ = nullptr;
  };

  class ComputePassDesc {
   public:
    ComputePassDesc &SetInputImages(std::vector<ImageViewProxyId> v) {
      inputs = std::move(v);
      return *this;
    }
    ComputePassDesc &SetStorageImages(std::vector<ImageViewProxyId> v) {
      storage = std::move(v);
      return *this;
    }
    ComputePassDesc &SetStorageBuffers(std::vector<BufferProxyId> v) {
      buffers = std::move(v);
      return *this;
    }
    ComputePassDesc &SetProfilerInfo(uint32_t c, std::string n) {
      color = c;
      name = std::move(n);
      return *this;
    }
    ComputePassDesc &SetRecordFunc(std::function<void(PassContext)> f) {
      record = std::move(f);
      return *this;
    }
    std::vector<ImageViewProxyId> inputs, storage;
    std::vector<BufferProxyId> buffers;
    uint32_t color = 0;
    std::string name = "ComputePass";
    std::function<void(PassContext)> record;
  };

  struct PassTiming {
    std::string name;
    uint32_t color;
    float ms;
  };

  RenderGraph();
  ~RenderGraph();
  RenderGraph(const RenderGraph &) = delete;
  RenderGraph &operator=(const RenderGraph &) = delete;

  // Image of `mips` levels (floor halving) allocated in HBM now; formats are LGX_*.
  ImageProxyUnique AddImage(int32_t format, uint32_t mips, uint32_t layers, uvec2 size, uint32_t usage);
  ImageViewProxyUnique AddImageView(ImageProxyId image, uint32_t base_mip, uint32_t mip_count, uint32_t base_layer,
                                    uint32_t layer_count);
  // Caller-owned memory bound as a one-mip view; UpdateExternalImageView
  // rebinds it (e.g. a new G-buffer each frame) without redeclaring passes.
  ImageViewProxyUnique AddExternalImageView(const lgx_image &image, uint32_t usage = kUsageSampled);
  void UpdateExternalImageView(ImageViewProxyId id, const lgx_image &image);
  template <class T> BufferProxyUnique AddBuffer(size_t count) { return AddRawBuffer(count * sizeof(T)); }
  BufferProxyUnique AddRawBuffer(size_t bytes);
  BufferProxyUnique AddExternalBuffer(void *ptr, size_t bytes);
  uvec2 GetMipSize(ImageViewProxyId view, uint32_t mip) const;
  // Host-side inspection outside passes (tests, resource set-up).
  lgx_image DescribeImageView(ImageViewProxyId view, uint32_t mip) const;
  void *BufferPointer(BufferProxyId id) const;

  void AddPass(ComputePassDesc &desc);
  void AddPass(ComputePassDesc &&desc) { AddPass(desc); }

  // Runs the declared passes in order on `stream` and clears the pass list
  // (passes are declared per frame, as RenderFrame does in the reference).
  void Execute(lgx_stream_t stream);
  // Runs the declared passes once under hipGraph capture and keeps the graph;
  // Replay() launches it again (inputs are read at replay time).  The pass
  // list is cleared after capture.
  void Capture(lgx_stream_t stream);
  void Replay(lgx_stream_t stream);
  bool HasCapture() const;
  // Drops passes declared but not yet executed (a frame whose declaration failed).
  void DiscardPasses() { passes_.clear(); }

  void SetProfiling(bool on) { profiling_ = on; }
  // Validation: record funcs may only touch views/buffers their pass declared.
  void SetValidation(bool on) { validate_ = on; }
  // Timings of the last executed frame (waits for it to finish).
  const std::vector<PassTiming> &GetLastFrameGpuProfilerData();
  // Mean per-pass time (by pass position) over the frames timed since the
  // last reset (waits for them to finish).
  std::vector<PassTiming> GetAccumulatedGpuProfilerData();
  void ResetGpuProfilerData();
  size_t DeclaredPassCount() const { return passes_.size(); }

 private:
  struct Level {
    void *ptr;
    int32_t w, h, pitch;
  };
  struct Image {
    bool alive = false, external = false;
    int32_t format = 0;
    uint32_t usage = 0;
    std::vector<Level> levels;
    void *block = nullptr;  // owning allocation
  };
  struct View {
    bool alive = false;
    uint32_t image = 0, base = 0, count = 0;
  };
  struct Buffer {
    bool alive = false, external = false;
    void *ptr = nullptr;
    size_t bytes = 0;
  };
  void ReleaseImage(ImageProxyId id);
  void ReleaseView(ImageViewProxyId id);
  void ReleaseBuffer(BufferProxyId id);
  void CheckAccess(size_t pass, ImageViewProxyId id) const;
  void CheckAccess(size_t pass, BufferProxyId id) const;
  void RunPasses(lgx_stream_t stream, bool time);
  void ResolveFrame(int slot);
  const View &ViewAt(ImageViewProxyId id) const;

  std::vector<Image> images_;
  std::vector<View> views_;
  std::vector<Buffer> buffers_;
  std::vector<ComputePassDesc> passes_;
  bool profiling_ = false, validate_ = false;
  struct FrameTiming;
  std::unique_ptr<FrameTiming> timing_;
  std::vector<PassTiming> last_timings_;
  struct CapturedGraph;
  std::unique_ptr<CapturedGraph> captured_;

  template <class IdT> friend class Unique;
};

// legit::ShaderMemoryPool, the per-frame uniform pool the reference's
// builders and record funcs take (frameInfo.memoryPool, SSVGIRenderer.h:207-221,
// 238-246).  The lgx passes carry their parameters as kernel arguments, so
// the pool holds nothing; the type lets call sites written against the
// reference builders compile unchanged.
class ShaderMemoryPool {};

// legit::Core (main.cpp:132): owns the render graph the renderers and
// builders declare passes on (Core::GetRenderGraph).
class Core {
 public:
  RenderGraph *GetRenderGraph() { return &renderGraph_; }

 private:
  RenderGraph renderGraph_;
};

template <class IdT> void RenderGraph::Unique<IdT>::reset() {
  if (!g_) return;
  if constexpr (std::is_same<IdT, ImageProxyId>::value) g_->ReleaseImage(id_);
  if constexpr (std::is_same<IdT, ImageViewProxyId>::value) g_->ReleaseView(id_);
  if constexpr (std::is_same<IdT, BufferProxyId>::value) g_->ReleaseBuffer(id_);
  g_ = nullptr;
}

}  // namespace lgx
