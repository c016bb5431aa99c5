// Denoise and composite passes.
//
//  * lgx_bilateral: the geometry-aware filter whose weights are written (and
//    commented out) in denoiser.frag:87-110: w = exp(-|d - dc| * 10.5) *
//    exp(-|n - nc|^2 * 20) over the taps [-r, r)^2 (y outer, x inner), at
//    pixel centres with clamp to edge, colour / (sumW + 1e-7).  One 16x16
//    output tile per workgroup; the tile plus its halo of irradiance, radial
//    depth and decoded normals is staged in LDS once, so each texel is read
//    from memory once per pass instead of (2r)^2 times.
//  * lgx_regression_denoise: the live DenoiserPass (denoiser.frag:72-185).
//  * lgx_composite: finalGatherer.frag:42-58.
#include <math.h>

#include "lgx_internal.h"

namespace lgx {

struct BilatArgs {
  const float4 *in;
  int in_pitch;
  const float *depth;
  int depth_pitch;
  const void *normal;
  int normal_pitch;
  float4 *out;
  int out_pitch;
  int w, h, r;
  float ds, ns;
};

constexpr int kBilatMaxR = 8;
constexpr int kBilatSpan = 16 + 2 * kBilatMaxR;

template <int NFMT> __global__ __launch_bounds__(256) void bilateral_kernel(const BilatArgs a) {
  __shared__ float4 s_col[kBilatSpan * kBilatSpan];
  __shared__ float4 s_geo[kBilatSpan * kBilatSpan];  // depth, normal xyz
  const int r = a.r;
  const int span = 16 + 2 * r - 1;  // taps reach -r .. r-1
  const int x0 = blockIdx.x * 16 - r, y0 = blockIdx.y * 16 - r;
  for (int i = threadIdx.x; i < span * span; i += 256) {
    const int lx = i % span, ly = i / span;
    const int gx = min(max(x0 + lx, 0), a.w - 1), gy = min(max(y0 + ly, 0), a.h - 1);
    s_col[ly * kBilatSpan + lx] = a.in[(size_t)gy * a.in_pitch + gx];
    const float4 nt = Texel<NFMT>::load(a.normal, (size_t)gy * a.normal_pitch + gx);
    const Vec3<float> n = oct_decode<float>(nt.x, nt.y);
    s_geo[ly * kBilatSpan + lx] = make_float4(a.depth[(size_t)gy * a.depth_pitch + gx], n.x, n.y, n.z);
  }
  __syncthreads();
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int px = blockIdx.x * 16 + tx, py = blockIdx.y * 16 + ty;
  if (px >= a.w || py >= a.h) return;
  const float4 gc = s_geo[(ty + r) * kBilatSpan + tx + r];
  // exp(-|dd| ds) exp(-|dn|^2 ns) = exp2(-(|dd| ds + |dn|^2 ns) log2(e)): one
  // hardware exp2 (1 ulp) per tap instead of two library expf
  const float kd = a.ds * 1.44269504f, kn = a.ns * 1.44269504f;
  float c0 = 0.f, c1 = 0.f, c2 = 0.f, sw = 0.f;
  for (int y = -r; y < r; y++)
    for (int x = -r; x < r; x++) {
      const int li = (ty + r + y) * kBilatSpan + tx + r + x;
      const float4 gs = s_geo[li];
      const float dx = gs.y - gc.y, dy = gs.z - gc.z, dz = gs.w - gc.w;
      const float wgt = __builtin_amdgcn_exp2f(-(fabsf(gs.x - gc.x) * kd + (dx * dx + dy * dy + dz * dz) * kn));
      const float4 c = s_col[li];
      c0 += c.x * wgt;
      c1 += c.y * wgt;
      c2 += c.z * wgt;
      sw += wgt;
    }
  a.out[(size_t)py * a.out_pitch + px] = make_float4(c0 / (sw + 1e-7f), c1 / (sw + 1e-7f), c2 / (sw + 1e-7f), 1.0f);
}

__global__ __launch_bounds__(256) void copy4_kernel(const float4 *in, int ip, float4 *out, int op, int w, int h) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x < w && y < h) out[(size_t)y * op + x] = in[(size_t)y * ip + x];
}

struct RegArgs {
  const float4 *in;
  int in_pitch;
  const float *depth;
  int depth_pitch;
  float4 *out;
  int out_pitch;
  int w, h;
};

__global__ __launch_bounds__(256) void regression_kernel(const RegArgs a) {
  const int px = blockIdx.x * 16 + (threadIdx.x & 15), py = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (px >= a.w || py >= a.h) return;
  float dep[16];
  float3 col[16];
  int i = 0;
  for (int y = -2; y < 2; y++)  // denoiser.frag:116-123
    for (int x = -2; x < 2; x++) {
      const int sx = min(max(px + x, 0), a.w - 1), sy = min(max(py + y, 0), a.h - 1);
      dep[i] = a.depth[(size_t)sy * a.depth_pitch + sx];
      const float4 c = a.in[(size_t)sy * a.in_pitch + sx];
      col[i] = make_float3(c.x, c.y, c.z);
      i++;
    }
  float g00 = 0.f, g01 = 0.f, g11 = 0.f;  // :135-147
  for (int k = 0; k < 16; k++) g00 += dep[k] * dep[k];
  for (int k = 0; k < 16; k++) g01 += dep[k] * 1.0f;
  for (int k = 0; k < 16; k++) g11 += 1.0f * 1.0f;
  g00 += 1e-7f;
  g11 += 1e-7f;
  const float det = g00 * g11 - g01 * g01;  // 2x2 block of inverse(mat4) (:148)
  const float i00 = g11 / det, i01 = -g01 / det, i11 = g00 / det;
  const float dc = a.depth[(size_t)py * a.depth_pitch + px];
  float res[3];
  for (int ch = 0; ch < 3; ch++) {  // :154-184
    float m0 = 0.f, m1 = 0.f;
    for (int k = 0; k < 16; k++) m0 += dep[k] * (&col[k].x)[ch];
    for (int k = 0; k < 16; k++) m1 += 1.0f * (&col[k].x)[ch];
    const float c0 = i00 * m0 + i01 * m1;
    const float c1 = i01 * m0 + i11 * m1;
    res[ch] = c0 * dc + c1 * 1.0f;
  }
  a.out[(size_t)py * a.out_pitch + px] = make_float4(res[0], res[1], res[2], 1.0f);
}

template <int RFMT, int AFMT>
__global__ __launch_bounds__(256) void composite_kernel(const void *rad, int rp, const void *alb, int ap,
                                                        const float4 *ind, int ip, float4 *out, int op, int w, int h) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= w || y >= h) return;
  const float4 d = Texel<RFMT>::load(rad, (size_t)y * rp + x);
  const float4 al = Texel<AFMT>::load(alb, (size_t)y * ap + x);
  const float4 in = ind[(size_t)y * ip + x];
  out[(size_t)y * op + x] = make_float4(d.x + in.x * al.x, d.y + in.y * al.y, d.z + in.z * al.z, d.w + in.w * al.w);
}

}  // namespace lgx

using namespace lgx;

extern "C" int lgx_bilateral(const lgx_bilateral_params *p, const lgx_image *in, const lgx_gbuffer *g,
                             lgx_image *out, lgx_stream_t stream) {
  if (!p || !g || !image_ok(in) || !image_ok(out) || !image_ok(&g->depth) || !image_ok(&g->normal))
    return LGX_EINVAL;
  if (in->format != LGX_RGBA32F || out->format != LGX_RGBA32F || g->depth.format != LGX_R32F) return LGX_EFORMAT;
  if (g->normal.format != LGX_RG16F && g->normal.format != LGX_RG32F) return LGX_EFORMAT;
  if (in->w != out->w || in->h != out->h || g->depth.w != in->w || g->depth.h != in->h || g->normal.w != in->w ||
      g->normal.h != in->h)
    return LGX_ESIZE;
  if (p->radius < 0 || p->radius > kBilatMaxR) return LGX_EUNSUPPORTED;
  if (in->ptr == out->ptr) return LGX_EINVAL;  // taps read neighbours: no in-place
  hipStream_t s = (hipStream_t)stream;
  if (p->radius == 0) {  // denoiser.frag:82-86
    dim3 grid((in->w + 63) / 64, (in->h + 3) / 4);
    hipLaunchKernelGGL(copy4_kernel, grid, dim3(256), 0, s, (const float4 *)in->ptr, in->pitch, (float4 *)out->ptr,
                       out->pitch, in->w, in->h);
    return hip_status(hipGetLastError());
  }
  BilatArgs a{(const float4 *)in->ptr, in->pitch, (const float *)g->depth.ptr, g->depth.pitch, g->normal.ptr,
              g->normal.pitch, (float4 *)out->ptr, out->pitch, in->w, in->h, p->radius, p->depth_sharpness,
              p->normal_sharpness};
  dim3 grid((in->w + 15) / 16, (in->h + 15) / 16);
  if (g->normal.format == LGX_RG16F)
    hipLaunchKernelGGL(bilateral_kernel<LGX_RG16F>, grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(bilateral_kernel<LGX_RG32F>, grid, dim3(256), 0, s, a);
  return hip_status(hipGetLastError());
}

extern "C" int lgx_regression_denoise(const lgx_image *in, const lgx_gbuffer *g, int32_t radius, lgx_image *out,
                                      lgx_stream_t stream) {
  if (!g || !image_ok(in) || !image_ok(out) || !image_ok(&g->depth)) return LGX_EINVAL;
  if (in->format != LGX_RGBA32F || out->format != LGX_RGBA32F || g->depth.format != LGX_R32F) return LGX_EFORMAT;
  if (in->w != out->w || in->h != out->h || g->depth.w != in->w || g->depth.h != in->h) return LGX_ESIZE;
  if (in->ptr == out->ptr) return LGX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  if (radius == 0) {
    dim3 grid((in->w + 63) / 64, (in->h + 3) / 4);
    hipLaunchKernelGGL(copy4_kernel, grid, dim3(256), 0, s, (const float4 *)in->ptr, in->pitch, (float4 *)out->ptr,
                       out->pitch, in->w, in->h);
    return hip_status(hipGetLastError());
  }
  RegArgs a{(const float4 *)in->ptr, in->pitch, (const float *)g->depth.ptr, g->depth.pitch, (float4 *)out->ptr,
            out->pitch, in->w, in->h};
  dim3 grid((in->w + 15) / 16, (in->h + 15) / 16);
  hipLaunchKernelGGL(regression_kernel, grid, dim3(256), 0, s, a);
  return hip_status(hipGetLastError());
}

extern "C" int lgx_composite(const lgx_gbuffer *g, const lgx_image *indirect, lgx_image *out, lgx_stream_t stream) {
  if (!g || !image_ok(indirect) || !image_ok(out) || !image_ok(&g->radiance) || !image_ok(&g->albedo))
    return LGX_EINVAL;
  if (indirect->format != LGX_RGBA32F || out->format != LGX_RGBA32F) return LGX_EFORMAT;
  const int w = out->w, h = out->h;
  if (indirect->w != w || indirect->h != h || g->radiance.w != w || g->radiance.h != h || g->albedo.w != w ||
      g->albedo.h != h)
    return LGX_ESIZE;
  hipStream_t s = (hipStream_t)stream;
  dim3 grid((w + 63) / 64, (h + 3) / 4), block(256);
  const int rf = g->radiance.format, af = g->albedo.format;
  const void *rp = g->radiance.ptr, *ap = g->albedo.ptr;
  const float4 *ip = (const float4 *)indirect->ptr;
  float4 *op = (float4 *)out->ptr;
  if (rf == LGX_RGBA16F && af == LGX_RGBA8)
    hipLaunchKernelGGL((composite_kernel<LGX_RGBA16F, LGX_RGBA8>), grid, block, 0, s, rp, g->radiance.pitch, ap,
                       g->albedo.pitch, ip, indirect->pitch, op, out->pitch, w, h);
  else if (rf == LGX_RGBA32F && af == LGX_RGBA32F)
    hipLaunchKernelGGL((composite_kernel<LGX_RGBA32F, LGX_RGBA32F>), grid, block, 0, s, rp, g->radiance.pitch, ap,
                       g->albedo.pitch, ip, indirect->pitch, op, out->pitch, w, h);
  else if (rf == LGX_RGBA16F && af == LGX_RGBA32F)
    hipLaunchKernelGGL((composite_kernel<LGX_RGBA16F, LGX_RGBA32F>), grid, block, 0, s, rp, g->radiance.pitch, ap,
                       g->albedo.pitch, ip, indirect->pitch, op, out->pitch, w, h);
  else if (rf == LGX_RGBA32F && af == LGX_RGBA8)
    hipLaunchKernelGGL((composite_kernel<LGX_RGBA32F, LGX_RGBA8>), grid, block, 0, s, rp, g->radiance.pitch, ap,
                       g->albedo.pitch, ip, indirect->pitch, op, out->pitch, w, h);
  else
    return LGX_EFORMAT;
  return hip_status(hipGetLastError());
}
