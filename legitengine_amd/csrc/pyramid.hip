// Mip pyramid builder and per-mip box pre-blur.
//
// MipBuilder::BuildMips (src/Render/Common/MipBuilder.h:142-181) records one
// full-screen pass per level; here one launch builds up to five levels:
// each 256-thread workgroup reduces a 32x32 texel tile of the source level
// through LDS down to a single texel (levels +1 .. +5), so the source is read
// from HBM once.  Extents halve with floor division (MipBuilder.h:148-149);
// a texel inside a level's extent always has its four children inside the
// level below, so out-of-extent LDS entries never feed a stored texel.
// Reductions follow mipLevelBuilder.frag:17-43 (children (0,0),(1,0),(0,1),
// (1,1) in that order); every stored level is re-rounded to its storage
// format before it feeds the next level, as the reference reads each level
// back from the image.
//
// BlurBuilder::ApplyBlur (src/Render/Common/BlurBuilder.h:14-46) is one
// pass per mip; lgx_box_blur batches any number of (src, dst, radius)
// triples into one launch.
#include "lgx_internal.h"

namespace lgx {

struct MipArgs {
  const void *src;
  int src_pitch;
  void *dst[5];
  int dst_pitch[5];
  int lw[6], lh[6];  // extents of the source (0) and the built levels
  int n;             // levels built in this launch (1..5)
};

template <int MODE> LGX_D float4 combine(float4 a, float4 b, float4 c, float4 d) {
  if (MODE == LGX_MIP_AVG) {
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    s.x += a.x; s.y += a.y; s.z += a.z; s.w += a.w;
    s.x += b.x; s.y += b.y; s.z += b.z; s.w += b.w;
    s.x += c.x; s.y += c.y; s.z += c.z; s.w += c.w;
    s.x += d.x; s.y += d.y; s.z += d.z; s.w += d.w;
    return make_float4(s.x / 4.0f, s.y / 4.0f, s.z / 4.0f, s.w / 4.0f);
  } else {
    float m = -1e5f;  // mipLevelBuilder.frag:32
    m = m < a.x ? a.x : m;
    m = m < b.x ? b.x : m;
    m = m < c.x ? c.x : m;
    m = m < d.x ? d.x : m;
    return make_float4(m, 0.f, 0.f, 1.f);
  }
}

template <int FMT, int MODE>
__global__ __launch_bounds__(256) void mip_kernel(const MipArgs a) {
  __shared__ float4 buf[2][16 * 16];
  const int tid = threadIdx.x;
  // level +1: 16x16 texels per tile, children read from the source
  {
    const int lx = tid & 15, ly = tid >> 4;
    const int i = blockIdx.x * 16 + lx, j = blockIdx.y * 16 + ly;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < a.lw[1] && j < a.lh[1]) {
      const size_t r0 = (size_t)(2 * j) * a.src_pitch, r1 = (size_t)(2 * j + 1) * a.src_pitch;
      float4 c00 = Texel<FMT>::load(a.src, r0 + 2 * i), c10 = Texel<FMT>::load(a.src, r0 + 2 * i + 1);
      float4 c01 = Texel<FMT>::load(a.src, r1 + 2 * i), c11 = Texel<FMT>::load(a.src, r1 + 2 * i + 1);
      v = Store<FMT>::round(combine<MODE>(c00, c10, c01, c11));
      Store<FMT>::store(a.dst[0], (size_t)j * a.dst_pitch[0] + i, v);
    }
    buf[0][tid] = v;
  }
  __syncthreads();
  int size = 16;
  for (int m = 1; m < a.n; m++) {
    const int ns = size >> 1;
    const float4 *prev = buf[(m - 1) & 1];
    float4 *cur = buf[m & 1];
    if (tid < ns * ns) {
      const int lx = tid % ns, ly = tid / ns;
      const int i = blockIdx.x * ns + lx, j = blockIdx.y * ns + ly;
      float4 v = Store<FMT>::round(combine<MODE>(prev[(2 * ly) * size + 2 * lx], prev[(2 * ly) * size + 2 * lx + 1],
                                                  prev[(2 * ly + 1) * size + 2 * lx],
                                                  prev[(2 * ly + 1) * size + 2 * lx + 1]));
      if (i < a.lw[m + 1] && j < a.lh[m + 1]) Store<FMT>::store(a.dst[m], (size_t)j * a.dst_pitch[m] + i, v);
      cur[tid] = v;
    }
    __syncthreads();
    size = ns;
  }
}

struct BlurJob {
  const void *src;
  void *dst;
  int w, h, sp, dp, r, block0, tiles_x;
};
struct BlurArgs {
  BlurJob job[LGX_MAX_LEVELS * 2];
  int n;
};

template <int FMT> __global__ __launch_bounds__(256) void blur_kernel(const BlurArgs a) {
  int jb = 0;
  while (jb + 1 < a.n && (int)blockIdx.x >= a.job[jb + 1].block0) jb++;
  const BlurJob &J = a.job[jb];
  const int t = (int)blockIdx.x - J.block0;
  const int px = (t % J.tiles_x) * 16 + (threadIdx.x & 15), py = (t / J.tiles_x) * 16 + (threadIdx.x >> 4);
  if (px >= J.w || py >= J.h) return;
  float4 out;
  if (J.r == 0) {
    out = Texel<FMT>::load(J.src, (size_t)py * J.sp + px);  // blurLayerBuilder.frag:30-31
  } else {
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    float tw = 0.0f;
    for (int x = -J.r; x < J.r; x++)  // :22-29, x outer, y inner
      for (int y = -J.r; y < J.r; y++) {
        int sx = min(max(px + x, 0), J.w - 1), sy = min(max(py + y, 0), J.h - 1);
        float4 v = Texel<FMT>::load(J.src, (size_t)sy * J.sp + sx);
        s.x += v.x;
        s.y += v.y;
        s.z += v.z;
        s.w += v.w;
        tw += 1.0f;
      }
    out = make_float4(s.x / tw, s.y / tw, s.z / tw, s.w / tw);
  }
  Store<FMT>::store(J.dst, (size_t)py * J.dp + px, Store<FMT>::round(out));
}

template <int FMT, int MODE> static hipError_t launch_mips(const MipArgs &a, hipStream_t s) {
  dim3 grid((a.lw[1] + 15) / 16, (a.lh[1] + 15) / 16), block(256);
  hipLaunchKernelGGL((mip_kernel<FMT, MODE>), grid, block, 0, s, a);
  return hipGetLastError();
}

}  // namespace lgx

using namespace lgx;

extern "C" int lgx_mip_build(const lgx_image *levels, int32_t n_levels, int32_t mode, lgx_stream_t stream) {
  if (!levels || n_levels < 1 || n_levels > LGX_MAX_LEVELS) return LGX_EINVAL;
  if (mode != LGX_MIP_AVG && mode != LGX_MIP_MAX) return LGX_EINVAL;
  const int fmt = levels[0].format;
  for (int l = 0; l < n_levels; l++) {
    if (!image_ok(&levels[l])) return LGX_EINVAL;
    if (levels[l].format != fmt) return LGX_EFORMAT;
    if (l > 0 && (levels[l].w != levels[l - 1].w / 2 || levels[l].h != levels[l - 1].h / 2)) return LGX_ESIZE;
  }
  if (mode == LGX_MIP_MAX && fmt != LGX_R32F) return LGX_EFORMAT;
  if (mode == LGX_MIP_AVG && fmt != LGX_R32F && fmt != LGX_RGBA16F && fmt != LGX_RGBA32F && fmt != LGX_RG32F)
    return LGX_EFORMAT;
  hipStream_t s = (hipStream_t)stream;
  for (int base = 0; base + 1 < n_levels; base += 5) {
    MipArgs a;
    memset(&a, 0, sizeof(a));
    a.n = n_levels - 1 - base < 5 ? n_levels - 1 - base : 5;
    a.src = levels[base].ptr;
    a.src_pitch = levels[base].pitch;
    a.lw[0] = levels[base].w;
    a.lh[0] = levels[base].h;
    for (int m = 0; m < a.n; m++) {
      a.dst[m] = levels[base + 1 + m].ptr;
      a.dst_pitch[m] = levels[base + 1 + m].pitch;
      a.lw[m + 1] = levels[base + 1 + m].w;
      a.lh[m + 1] = levels[base + 1 + m].h;
    }
    hipError_t e;
    if (mode == LGX_MIP_MAX)
      e = launch_mips<LGX_R32F, LGX_MIP_MAX>(a, s);
    else if (fmt == LGX_R32F)
      e = launch_mips<LGX_R32F, LGX_MIP_AVG>(a, s);
    else if (fmt == LGX_RG32F)
      e = launch_mips<LGX_RG32F, LGX_MIP_AVG>(a, s);
    else if (fmt == LGX_RGBA16F)
      e = launch_mips<LGX_RGBA16F, LGX_MIP_AVG>(a, s);
    else
      e = launch_mips<LGX_RGBA32F, LGX_MIP_AVG>(a, s);
    if (e != hipSuccess) return hip_status(e);
  }
  return LGX_OK;
}

extern "C" int lgx_box_blur(const lgx_image *src, const lgx_image *dst, const int32_t *radius, int32_t n,
                            lgx_stream_t stream) {
  if (!src || !dst || !radius || n < 1 || n > LGX_MAX_LEVELS * 2) return LGX_EINVAL;
  const int fmt = src[0].format;
  BlurArgs a;
  memset(&a, 0, sizeof(a));
  a.n = n;
  int blocks = 0;
  for (int i = 0; i < n; i++) {
    if (!image_ok(&src[i]) || !image_ok(&dst[i]) || radius[i] < 0 || radius[i] > 64) return LGX_EINVAL;
    if (src[i].format != fmt || dst[i].format != fmt) return LGX_EFORMAT;
    if (src[i].w != dst[i].w || src[i].h != dst[i].h) return LGX_ESIZE;  // BlurBuilder.h:17
    BlurJob &J = a.job[i];
    J.src = src[i].ptr;
    J.dst = dst[i].ptr;
    J.w = src[i].w;
    J.h = src[i].h;
    J.sp = src[i].pitch;
    J.dp = dst[i].pitch;
    J.r = radius[i];
    J.tiles_x = (J.w + 15) / 16;
    J.block0 = blocks;
    blocks += J.tiles_x * ((J.h + 15) / 16);
  }
  hipStream_t s = (hipStream_t)stream;
  dim3 grid(blocks), block(256);
  switch (fmt) {
    case LGX_R32F: hipLaunchKernelGGL(blur_kernel<LGX_R32F>, grid, block, 0, s, a); break;
    case LGX_RG32F: hipLaunchKernelGGL(blur_kernel<LGX_RG32F>, grid, block, 0, s, a); break;
    case LGX_RGBA16F: hipLaunchKernelGGL(blur_kernel<LGX_RGBA16F>, grid, block, 0, s, a); break;
    case LGX_RGBA32F: hipLaunchKernelGGL(blur_kernel<LGX_RGBA32F>, grid, block, 0, s, a); break;
    default: return LGX_EFORMAT;
  }
  return hip_status(hipGetLastError());
}
