// 4x4 interleave permutations (InterleaveBuilder, src/Render/Common/InterleaveBuilder.h:14-80).
//
// Pure gathers: every thread writes one destination texel (coalesced
// stores) and reads the source texel the permutation names.  Texels are
// moved as opaque 4/8/16/32-byte elements, so the result is bit-exact for
// every format.
#include "lgx_internal.h"

namespace lgx {

struct PermArgs {
  const void *src;
  void *dst;
  int w, h, sp, dp, gx, gy, cx, cy;
  int inverse;  // 0: deinterleave, 1: interleave
};

template <class E> __global__ __launch_bounds__(256) void perm_kernel(const PermArgs a) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= a.w || y >= a.h) return;
  int sx, sy;
  if (!a.inverse) {
    // deinterleave.frag:16-21 InterleavePixel: patternIndex = d / cell; src = (d % cell) * grid + patternIndex
    sx = (x % a.cx) * a.gx + x / a.cx;
    sy = (y % a.cy) * a.gy + y / a.cy;
  } else {
    // interleave.frag:15-20 DeinterleavePixel: src = (i % grid) * cell + i / grid
    sx = (x % a.gx) * a.cx + x / a.gx;
    sy = (y % a.gy) * a.cy + y / a.gy;
  }
  ((E *)a.dst)[(size_t)y * a.dp + x] = ((const E *)a.src)[(size_t)sy * a.sp + sx];
}

struct Elem32 {
  uint4 a, b;
};

static int permute(const lgx_image *in, lgx_image *out, int gx, int gy, int inverse, hipStream_t s) {
  if (!image_ok(in) || !image_ok(out) || gx < 1 || gy < 1) return LGX_EINVAL;
  if (in->format != out->format) return LGX_EFORMAT;
  if (in->w != out->w || in->h != out->h) return LGX_ESIZE;
  if (in->w % gx || in->h % gy) return LGX_ESIZE;  // bijective only when the grid divides the extent
  PermArgs a{in->ptr, out->ptr, in->w, in->h, in->pitch, out->pitch, gx, gy, in->w / gx, in->h / gy, inverse};
  dim3 grid((in->w + 63) / 64, (in->h + 3) / 4), block(256);
  switch (format_bytes(in->format)) {
    case 2: hipLaunchKernelGGL(perm_kernel<uint16_t>, grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL(perm_kernel<uint32_t>, grid, block, 0, s, a); break;
    case 8: hipLaunchKernelGGL(perm_kernel<uint2>, grid, block, 0, s, a); break;
    case 16: hipLaunchKernelGGL(perm_kernel<uint4>, grid, block, 0, s, a); break;
    case 32: hipLaunchKernelGGL(perm_kernel<Elem32>, grid, block, 0, s, a); break;
    default: return LGX_EFORMAT;
  }
  return hip_status(hipGetLastError());
}

}  // namespace lgx

extern "C" int lgx_deinterleave(const lgx_image *in, lgx_image *out, int32_t gx, int32_t gy, lgx_stream_t stream) {
  return lgx::permute(in, out, gx, gy, 0, (hipStream_t)stream);
}

extern "C" int lgx_interleave(const lgx_image *in, lgx_image *out, int32_t gx, int32_t gy, lgx_stream_t stream) {
  return lgx::permute(in, out, gx, gy, 1, (hipStream_t)stream);
}
