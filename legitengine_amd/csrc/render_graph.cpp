// lgx::RenderGraph implementation: proxies over HBM allocations, passes
// recorded in declaration order on one HIP stream, per-pass hipEvent timing,
// optional access validation and hipGraph capture of a whole frame.
#include "lgx/render_graph.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <string>

namespace lgx {

int format_bytes_host(int f) {
  switch (f) {
    case LGX_R32F: case LGX_RG16F: case LGX_RGBA8: return 4;
    case LGX_RG32F: case LGX_RGBA16F: return 8;
    case LGX_RGBA32F: return 16;
    case LGX_RGBA64F: return 32;
    case LGX_R16F: return 2;
    default: return 0;
  }
}

Error::Error(int code, const std::string &what)
    : std::runtime_error(what + ": " + lgx_strerror(code) + " (" + std::to_string(code) + ")"), code_(code) {}

void Check(int code, const char *what) {
  if (code != LGX_OK) throw Error(code, what);
}

static void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    (void)hipGetLastError();  // do not leave our failure to be reported by the next launch
    throw Error((int)e, what);
  }
}

// Per-pass event pairs for up to kFramesInFlight frames, so reading the
// timings of frame n does not stall the recording of frame n+1 (the
// reference reads GPU timestamps one frame late, main.cpp:273).
struct RenderGraph::FrameTiming {
  static const int kFramesInFlight = 4;
  struct Frame {
    std::vector<hipEvent_t> events;  // 2 per pass
    std::vector<std::pair<std::string, uint32_t>> names;
    bool pending = false;
  };
  Frame frames[kFramesInFlight];
  int next = 0;       // slot of the next frame
  int last = -1;      // slot of the most recent recorded frame
  std::vector<std::pair<double, int>> acc;  // per-pass sum of ms and count, by pass position
  std::vector<std::string> acc_names;
  ~FrameTiming() {
    for (auto &f : frames)
      for (auto e : f.events) (void)hipEventDestroy(e);
  }
};

struct RenderGraph::CapturedGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~CapturedGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

RenderGraph::RenderGraph() : timing_(new FrameTiming) {}

RenderGraph::~RenderGraph() {
  captured_.reset();
  for (auto &im : images_)
    if (im.alive && !im.external && im.block) (void)hipFree(im.block);
  for (auto &b : buffers_)
    if (b.alive && !b.external && b.ptr) (void)hipFree(b.ptr);
}

RenderGraph::ImageProxyUnique RenderGraph::AddImage(int32_t format, uint32_t mips, uint32_t layers, uvec2 size,
                                                    uint32_t usage) {
  const int eb = format_bytes_host(format);
  if (eb <= 0) throw Error(LGX_EFORMAT, "RenderGraph::AddImage");
  if (layers != 1 || mips < 1 || mips > LGX_MAX_LEVELS || size.x == 0 || size.y == 0)
    throw Error(LGX_EUNSUPPORTED, "RenderGraph::AddImage");
  Image im;
  im.alive = true;
  im.format = format;
  im.usage = usage;
  // one block, levels 256-byte aligned; rows pitched to 64 texels so row
  // starts stay aligned for vector loads
  size_t total = 0;
  uint32_t w = size.x, h = size.y;
  std::vector<size_t> offs;
  for (uint32_t l = 0; l < mips; l++) {
    if (w == 0 || h == 0) throw Error(LGX_ESIZE, "RenderGraph::AddImage: too many mips for the extent");
    int32_t pitch = (int32_t)((w + 63) / 64 * 64);
    offs.push_back(total);
    total += ((size_t)pitch * h * eb + 255) / 256 * 256;
    im.levels.push_back(Level{nullptr, (int32_t)w, (int32_t)h, pitch});
    w /= 2;  // MipBuilder.h:148-149
    h /= 2;
  }
  hip_check(hipMalloc(&im.block, total), "RenderGraph::AddImage hipMalloc");
  for (uint32_t l = 0; l < mips; l++) im.levels[l].ptr = (char *)im.block + offs[l];
  images_.push_back(std::move(im));
  return ImageProxyUnique(this, ImageProxyId{(uint32_t)images_.size() - 1});
}

RenderGraph::ImageViewProxyUnique RenderGraph::AddImageView(ImageProxyId image, uint32_t base_mip, uint32_t mip_count,
                                                            uint32_t base_layer, uint32_t layer_count) {
  if (image.v >= images_.size() || !images_[image.v].alive) throw Error(LGX_EINVAL, "RenderGraph::AddImageView");
  const Image &im = images_[image.v];
  if (base_layer != 0 || layer_count != 1 || mip_count == 0 || base_mip + mip_count > im.levels.size())
    throw Error(LGX_EUNSUPPORTED, "RenderGraph::AddImageView");
  views_.push_back(View{true, image.v, base_mip, mip_count});
  return ImageViewProxyUnique(this, ImageViewProxyId{(uint32_t)views_.size() - 1});
}

RenderGraph::ImageViewProxyUnique RenderGraph::AddExternalImageView(const lgx_image &image, uint32_t usage) {
  Image im;
  im.alive = true;
  im.external = true;
  im.format = image.format;
  im.usage = usage;
  im.levels.push_back(Level{image.ptr, image.w, image.h, image.pitch});
  images_.push_back(std::move(im));
  views_.push_back(View{true, (uint32_t)images_.size() - 1, 0, 1});
  return ImageViewProxyUnique(this, ImageViewProxyId{(uint32_t)views_.size() - 1});
}

void RenderGraph::UpdateExternalImageView(ImageViewProxyId id, const lgx_image &image) {
  const View &v = ViewAt(id);
  Image &im = images_[v.image];
  if (!im.external) throw Error(LGX_EINVAL, "RenderGraph::UpdateExternalImageView: not an external view");
  if (image.format != im.format || image.w != im.levels[0].w || image.h != im.levels[0].h)
    throw Error(LGX_ESIZE, "RenderGraph::UpdateExternalImageView: extent or format changed");
  im.levels[0] = Level{image.ptr, image.w, image.h, image.pitch};
}

RenderGraph::BufferProxyUnique RenderGraph::AddRawBuffer(size_t bytes) {
  Buffer b;
  b.alive = true;
  b.bytes = bytes;
  hip_check(hipMalloc(&b.ptr, bytes ? bytes : 1), "RenderGraph::AddBuffer hipMalloc");
  buffers_.push_back(b);
  return BufferProxyUnique(this, BufferProxyId{(uint32_t)buffers_.size() - 1});
}

RenderGraph::BufferProxyUnique RenderGraph::AddExternalBuffer(void *ptr, size_t bytes) {
  Buffer b;
  b.alive = true;
  b.external = true;
  b.ptr = ptr;
  b.bytes = bytes;
  buffers_.push_back(b);
  return BufferProxyUnique(this, BufferProxyId{(uint32_t)buffers_.size() - 1});
}

const RenderGraph::View &RenderGraph::ViewAt(ImageViewProxyId id) const {
  if (id.v >= views_.size() || !views_[id.v].alive) throw Error(LGX_EINVAL, "RenderGraph: unknown image view proxy");
  return views_[id.v];
}

uvec2 RenderGraph::GetMipSize(ImageViewProxyId view, uint32_t mip) const {
  const View &v = ViewAt(view);
  if (mip >= v.count) throw Error(LGX_EINVAL, "RenderGraph::GetMipSize");
  const Level &L = images_[v.image].levels[v.base + mip];
  return uvec2{(uint32_t)L.w, (uint32_t)L.h};
}

lgx_image RenderGraph::DescribeImageView(ImageViewProxyId id, uint32_t mip) const {
  const View &v = ViewAt(id);
  if (mip >= v.count) throw Error(LGX_EINVAL, "RenderGraph::DescribeImageView");
  const Image &im = images_[v.image];
  const Level &L = im.levels[v.base + mip];
  return lgx_image{L.ptr, L.w, L.h, L.pitch, im.format, 0};
}

void *RenderGraph::BufferPointer(BufferProxyId id) const {
  if (id.v >= buffers_.size() || !buffers_[id.v].alive) throw Error(LGX_EINVAL, "RenderGraph::BufferPointer");
  return buffers_[id.v].ptr;
}

void RenderGraph::ReleaseImage(ImageProxyId id) {
  if (id.v >= images_.size() || !images_[id.v].alive) return;
  Image &im = images_[id.v];
  if (!im.external && im.block) (void)hipFree(im.block);
  im = Image();
}
void RenderGraph::ReleaseView(ImageViewProxyId id) {
  if (id.v >= views_.size() || !views_[id.v].alive) return;
  View &v = views_[id.v];
  if (images_[v.image].external) images_[v.image] = Image();  // an external view owns its slot
  v = View();
}
void RenderGraph::ReleaseBuffer(BufferProxyId id) {
  if (id.v >= buffers_.size() || !buffers_[id.v].alive) return;
  Buffer &b = buffers_[id.v];
  if (!b.external && b.ptr) (void)hipFree(b.ptr);
  b = Buffer();
}

void RenderGraph::AddPass(ComputePassDesc &desc) {
  for (auto id : desc.inputs) (void)ViewAt(id);
  for (auto id : desc.storage) (void)ViewAt(id);
  for (auto id : desc.buffers)
    if (id.v >= buffers_.size() || !buffers_[id.v].alive) throw Error(LGX_EINVAL, "RenderGraph::AddPass: buffer");
  passes_.push_back(desc);
}

void RenderGraph::CheckAccess(size_t pass, ImageViewProxyId id) const {
  if (!validate_) return;
  const ComputePassDesc &p = passes_[pass];
  // a view is declared when it, or a view of the same image covering the
  // accessed mip range, is in the pass's input or storage list
  const View &v = ViewAt(id);
  auto covers = [&](ImageViewProxyId d) {
    const View &dv = views_[d.v];
    return d == id || (dv.image == v.image && dv.base <= v.base && dv.base + dv.count >= v.base + v.count);
  };
  for (auto d : p.inputs)
    if (covers(d)) return;
  for (auto d : p.storage)
    if (covers(d)) return;
  throw Error(LGX_EINVAL, "RenderGraph validation: pass '" + p.name + "' touched an undeclared image view");
}
void RenderGraph::CheckAccess(size_t pass, BufferProxyId id) const {
  if (!validate_) return;
  const ComputePassDesc &p = passes_[pass];
  if (std::find(p.buffers.begin(), p.buffers.end(), id) != p.buffers.end()) return;
  throw Error(LGX_EINVAL, "RenderGraph validation: pass '" + p.name + "' touched an undeclared buffer");
}

lgx_image RenderGraph::PassContext::GetImageView(ImageViewProxyId id) const { return GetImageViewMip(id, 0); }

lgx_image RenderGraph::PassContext::GetImageViewMip(ImageViewProxyId id, uint32_t i) const {
  g_->CheckAccess(pass_, id);
  const View &v = g_->ViewAt(id);
  if (i >= v.count) throw Error(LGX_EINVAL, "PassContext::GetImageViewMip");
  const Image &im = g_->images_[v.image];
  const Level &L = im.levels[v.base + i];
  return lgx_image{L.ptr, L.w, L.h, L.pitch, im.format, 0};
}

uint32_t RenderGraph::PassContext::GetImageViewMipCount(ImageViewProxyId id) const { return g_->ViewAt(id).count; }

void *RenderGraph::PassContext::GetBuffer(BufferProxyId id) const {
  g_->CheckAccess(pass_, id);
  if (id.v >= g_->buffers_.size() || !g_->buffers_[id.v].alive) throw Error(LGX_EINVAL, "PassContext::GetBuffer");
  return g_->buffers_[id.v].ptr;
}

void RenderGraph::ResolveFrame(int slot) {
  FrameTiming &T = *timing_;
  FrameTiming::Frame &F = T.frames[slot];
  if (!F.pending) return;
  std::vector<PassTiming> out;
  for (size_t i = 0; i < F.names.size(); i++) {
    hip_check(hipEventSynchronize(F.events[2 * i + 1]), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, F.events[2 * i], F.events[2 * i + 1]), "hipEventElapsedTime");
    out.push_back(PassTiming{F.names[i].first, F.names[i].second, ms});
    if (T.acc.size() <= i) {
      T.acc.resize(i + 1, {0.0, 0});
      T.acc_names.resize(i + 1);
    }
    T.acc[i].first += ms;
    T.acc[i].second += 1;
    T.acc_names[i] = F.names[i].first;
  }
  F.pending = false;
  if (slot == T.last) last_timings_ = out;
}

void RenderGraph::RunPasses(lgx_stream_t stream, bool time) {
  FrameTiming &T = *timing_;
  FrameTiming::Frame *F = nullptr;
  if (time) {
    F = &T.frames[T.next];
    if (F->pending) ResolveFrame(T.next);  // kFramesInFlight frames old: normally complete
    size_t need = passes_.size() * 2;
    while (F->events.size() < need) {
      hipEvent_t e;
      hip_check(hipEventCreate(&e), "hipEventCreate");
      F->events.push_back(e);
    }
    F->names.clear();
  }
  for (size_t i = 0; i < passes_.size(); i++) {
    PassContext ctx;
    ctx.g_ = this;
    ctx.pass_ = i;
    ctx.stream_ = stream;
    if (time) hip_check(hipEventRecord(F->events[2 * i], (hipStream_t)stream), "hipEventRecord");
    if (passes_[i].record) passes_[i].record(ctx);
    if (time) {
      hip_check(hipEventRecord(F->events[2 * i + 1], (hipStream_t)stream), "hipEventRecord");
      F->names.emplace_back(passes_[i].name, passes_[i].color);
    }
  }
  if (time) {
    F->pending = true;
    T.last = T.next;
    T.next = (T.next + 1) % FrameTiming::kFramesInFlight;
  }
}

void RenderGraph::Execute(lgx_stream_t stream) {
  try {
    RunPasses(stream, profiling_);
  } catch (...) {
    passes_.clear();
    throw;
  }
  passes_.clear();
}

void RenderGraph::Capture(lgx_stream_t stream) {
  if (!stream) {
    passes_.clear();
    throw Error(LGX_EINVAL, "RenderGraph::Capture: the legacy default stream cannot be captured");
  }
  captured_.reset(new CapturedGraph);
  hipStream_t s = (hipStream_t)stream;
  hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  try {
    RunPasses(stream, false);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    passes_.clear();
    captured_.reset();
    throw;
  }
  passes_.clear();
  hip_check(hipStreamEndCapture(s, &captured_->graph), "hipStreamEndCapture");
  hip_check(hipGraphInstantiate(&captured_->exec, captured_->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
}

void RenderGraph::Replay(lgx_stream_t stream) {
  if (!captured_ || !captured_->exec) throw Error(LGX_EINVAL, "RenderGraph::Replay: nothing captured");
  hip_check(hipGraphLaunch(captured_->exec, (hipStream_t)stream), "hipGraphLaunch");
}

bool RenderGraph::HasCapture() const { return captured_ && captured_->exec; }

const std::vector<RenderGraph::PassTiming> &RenderGraph::GetLastFrameGpuProfilerData() {
  FrameTiming &T = *timing_;
  // resolve oldest first so the accumulation stays in frame order
  for (int k = 0; k < FrameTiming::kFramesInFlight; k++) ResolveFrame((T.next + k) % FrameTiming::kFramesInFlight);
  return last_timings_;
}

std::vector<RenderGraph::PassTiming> RenderGraph::GetAccumulatedGpuProfilerData() {
  (void)GetLastFrameGpuProfilerData();
  std::vector<PassTiming> out;
  const FrameTiming &T = *timing_;
  for (size_t i = 0; i < T.acc.size(); i++)
    out.push_back(PassTiming{T.acc_names[i], 0u, T.acc[i].second ? (float)(T.acc[i].first / T.acc[i].second) : 0.f});
  return out;
}

void RenderGraph::ResetGpuProfilerData() {
  (void)GetLastFrameGpuProfilerData();
  timing_->acc.clear();
  timing_->acc_names.clear();
}

}  // namespace lgx
