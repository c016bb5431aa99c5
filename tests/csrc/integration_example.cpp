// A C++ consumer of the lgx headers written the way LegitEngine drives its
// renderers: the frame loop of src/main.cpp:132-160 and :269-312
// (BeginFrame -> RenderFrame -> EndFrame) over a BaseRenderer, and the
// builder calls of SSVGIRenderer.h:207-221 with their reference argument
// lists.  The C++ in INTEGRATION.md section 1 is the text between the
// snippet markers below (a CPU test checks that it stays so), and build()
// compiles this file, so the documented usage compiles.
//
// Run on a GPU, it renders one frame through lgx::SSVGIRenderer and the same
// frame through the C ABI (lgx_renderer_*), builds a Hi-Z pyramid, a
// pre-blurred radiance pyramid and an interleave round trip through the
// reference-signature builders, and exits 0 only if every comparison is
// bit-identical.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "lgx/base_renderer.h"

static int fail(const char *what) {
  std::fprintf(stderr, "integration: FAILED: %s\n", what);
  return 1;
}

static size_t texel_bytes(int f) {
  switch (f) {
    case LGX_R32F: case LGX_RG16F: case LGX_RGBA8: return 4;
    case LGX_RG32F: case LGX_RGBA16F: return 8;
    case LGX_RGBA32F: return 16;
    default: return 0;
  }
}

static std::vector<unsigned char> download(const lgx_image &im) {
  const size_t row = (size_t)im.w * texel_bytes(im.format);
  std::vector<unsigned char> out(row * im.h);
  if (hipMemcpy2D(out.data(), row, im.ptr, (size_t)im.pitch * texel_bytes(im.format), row, im.h,
                  hipMemcpyDeviceToHost) != hipSuccess)
    out.clear();
  return out;
}

static lgx_image alloc_image(int w, int h, int format) {
  lgx_image im{nullptr, w, h, w, format, 0};
  if (hipMalloc(&im.ptr, (size_t)w * h * texel_bytes(format)) != hipSuccess) im.ptr = nullptr;
  return im;
}

int main() {
  const uint32_t width = 320, height = 192;
  lgx_scene scene;
  if (lgx_scene_make(7, 8, 2, &scene) != LGX_OK) return fail("lgx_scene_make");
  hipStream_t stream;
  if (hipStreamCreate(&stream) != hipSuccess) return fail("hipStreamCreate");
  lgx::SsvgiConfig cfg;
  cfg.dirs = 16;
  cfg.max_steps = 32;
  cfg.levels = 5;
  cfg.composite = true;  // indirect light times albedo plus direct light (finalGatherer.frag:42-58)
  try {
    // [integration-snippet-begin]
    lgx::Core core;                                      // legit::Core (main.cpp:132)
    lgx::InFlightQueue inFlightQueue(&core, stream);     // frames run on `stream`, in declaration order
    std::unique_ptr<lgx::BaseRenderer> renderer(new lgx::SSVGIRenderer(&core, cfg));  // CreateRenderer (main.cpp:63-75)
    renderer->RecreateSceneResources(&scene);
    renderer->RecreateSwapchainResources({width, height}, inFlightQueue.GetInFlightFramesCount());

    lgx::Camera camera, light;                           // Camera (Scene.h:20-34)
    std::memcpy(camera.pos, scene.cam_pos, sizeof(camera.pos));
    camera.horAngle = scene.cam_hor;
    camera.vertAngle = scene.cam_vert;

    auto frameInfo = inFlightQueue.BeginFrame();                       // main.cpp:269
    renderer->RenderFrame(frameInfo, camera, light, &scene, nullptr);  // main.cpp:278
    inFlightQueue.EndFrame();                                          // main.cpp:312: the passes run here

    // the reference's builders, with its argument lists (SSVGIRenderer.h:207-221)
    lgx::RenderGraph *rg = core.GetRenderGraph();
    auto &ssvgi = static_cast<lgx::SSVGIRenderer &>(*renderer);
    const lgx_gbuffer gbuffer = ssvgi.GetGBuffer();
    lgx::MippedProxy depthMips(rg, LGX_R32F, {width, height}, lgx::kUsageStorage, cfg.levels, true);
    lgx::MippedProxy lightMips(rg, LGX_RGBA16F, {width, height}, lgx::kUsageStorage, cfg.levels, true);
    lgx::MippedProxy blurredLight(rg, LGX_RGBA16F, {width, height}, lgx::kUsageStorage, cfg.levels, true);
    depthMips.BindExternalLevel0(gbuffer.depth);
    lightMips.BindExternalLevel0(gbuffer.radiance);
    blurredLight.BindExternalLevel0(gbuffer.radiance);
    lgx::MipBuilder mipBuilder(&core);
    lgx::BlurBuilder blurBuilder(&core);
    lgx::InterleaveBuilder interleaveBuilder(&core);
    lgx::UnmippedProxy slices(rg, LGX_RGBA32F, {width, height}, lgx::kUsageStorage);
    lgx::UnmippedProxy restored(rg, LGX_RGBA32F, {width, height}, lgx::kUsageStorage);
    auto output = rg->AddExternalImageView(ssvgi.GetOutput());

    frameInfo = inFlightQueue.BeginFrame();
    mipBuilder.BuildMips(rg, frameInfo.memoryPool, depthMips, lgx::MipBuilder::FilterTypes::Depth);  // max-Z Hi-Z
    mipBuilder.BuildMips(rg, frameInfo.memoryPool, lightMips);
    for (uint32_t mip = 1; mip < lightMips.mipImageViewProxies.size(); mip++)
      blurBuilder.ApplyBlur(rg, frameInfo.memoryPool, lightMips.mipImageViewProxies[mip]->Id(),
                            blurredLight.mipImageViewProxies[mip]->Id(), 2);
    interleaveBuilder.Deinterleave(rg, frameInfo.memoryPool, output->Id(), slices.imageViewProxy->Id(), {4, 4});
    interleaveBuilder.Interleave(rg, frameInfo.memoryPool, slices.imageViewProxy->Id(),
                                 restored.imageViewProxy->Id(), {4, 4});
    inFlightQueue.EndFrame();
    // [integration-snippet-end]
    if (hipStreamSynchronize(stream) != hipSuccess) return fail("hipStreamSynchronize");

    // the same frame through the C ABI renderer
    lgx_renderer_config rc;
    std::memset(&rc, 0, sizeof(rc));
    rc.width = (int32_t)width;
    rc.height = (int32_t)height;
    rc.dirs = cfg.dirs;
    rc.max_steps = cfg.max_steps;
    rc.levels = cfg.levels;
    rc.grid = cfg.grid;
    rc.sampler = cfg.sampler;
    rc.denoiser = cfg.denoiser;
    rc.denoise_passes = cfg.denoise_passes;
    rc.denoise_radius = cfg.denoise_radius;
    rc.depth_sharpness = cfg.depth_sharpness;
    rc.normal_sharpness = cfg.normal_sharpness;
    rc.normal_format = cfg.normal_format;
    rc.albedo_format = cfg.albedo_format;
    rc.radiance_format = cfg.radiance_format;
    rc.composite = 1;
    lgx_renderer *cr = nullptr;
    if (lgx_renderer_create(&rc, &cr) != LGX_OK) return fail("lgx_renderer_create");
    lgx_gbuffer g2;
    g2.depth = alloc_image(width, height, LGX_R32F);
    g2.normal = alloc_image(width, height, cfg.normal_format);
    g2.albedo = alloc_image(width, height, cfg.albedo_format);
    g2.radiance = alloc_image(width, height, cfg.radiance_format);
    if (!g2.depth.ptr || !g2.normal.ptr || !g2.albedo.ptr || !g2.radiance.ptr) return fail("hipMalloc");
    float view[16], proj[16];
    if (lgx_scene_matrices(&scene, (int32_t)width, (int32_t)height, view, proj) != LGX_OK)
      return fail("lgx_scene_matrices");
    if (lgx_scene_render_gpu(&scene, &g2, stream) != LGX_OK) return fail("lgx_scene_render_gpu");
    if (lgx_renderer_render(cr, &g2, view, proj, stream) != LGX_OK) return fail(lgx_renderer_last_error(cr));
    if (hipStreamSynchronize(stream) != hipSuccess) return fail("hipStreamSynchronize");
    lgx_image cfinal;
    if (lgx_renderer_intermediate(cr, "final", 0, &cfinal) != LGX_OK) return fail("lgx_renderer_intermediate");

    const auto a = download(ssvgi.GetOutput()), b = download(cfinal);
    if (a.empty() || a != b) return fail("frame differs from lgx_renderer_render");
    if (download(gbuffer.radiance) != download(g2.radiance)) return fail("G-buffer differs");
    if (std::memcmp(view, ssvgi.GetViewMatrix(), 64) || std::memcmp(proj, ssvgi.GetProjMatrix(), 64))
      return fail("camera matrices differ");
    int levels_checked = 0;
    for (uint32_t mip = 1; mip < (uint32_t)cfg.levels; mip++) {
      lgx_image ref_d, ref_r;
      if (lgx_renderer_intermediate(cr, "depth_pyramid", (int32_t)mip, &ref_d) != LGX_OK ||
          lgx_renderer_intermediate(cr, "blurred_radiance", (int32_t)mip, &ref_r) != LGX_OK)
        return fail("lgx_renderer_intermediate (pyramids)");
      if (download(rg->DescribeImageView(depthMips.mipImageViewProxies[mip]->Id(), 0)) != download(ref_d))
        return fail("Hi-Z level differs");
      if (download(rg->DescribeImageView(blurredLight.mipImageViewProxies[mip]->Id(), 0)) != download(ref_r))
        return fail("pre-blurred radiance level differs");
      levels_checked++;
    }
    if (download(rg->DescribeImageView(restored.imageViewProxy->Id(), 0)) != a)
      return fail("Interleave(Deinterleave(frame)) differs from the frame");
    std::printf("integration: %ux%u frame through SSVGIRenderer bit-identical to lgx_renderer_render (%zu bytes); "
                "BuildMips/ApplyBlur with the memory pool match the renderer's %d pyramid levels; "
                "interleave round trip identical; passes last frame: %zu\n",
                width, height, a.size(), levels_checked, inFlightQueue.GetLastFrameGpuProfilerData().size());
    lgx_renderer_destroy(cr);
    (void)hipFree(g2.depth.ptr);
    (void)hipFree(g2.normal.ptr);
    (void)hipFree(g2.albedo.ptr);
    (void)hipFree(g2.radiance.ptr);
  } catch (const lgx::Error &e) {
    std::fprintf(stderr, "integration: lgx::Error %d: %s\n", e.code(), e.what());
    return 1;
  }
  return 0;
}
