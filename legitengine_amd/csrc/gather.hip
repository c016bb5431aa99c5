// Horizon gather: IndirectLightPass (src/Render/Renderers/SSVGIRenderer.h:223-263)
// running SSVGI/indirectLighting.frag:114-272 as a gfx950 compute kernel.
//
// Design (MI355X):
//  * Thread mapping per 4x4 interleave slice: every workgroup covers a tile
//    of ONE slice (all pixels with the same pattern index), so a wavefront's
//    64 lanes share the Hammersley offsets, i.e. the screen directions and
//    the geometric step offsets.  Those come from per-slice tables in the
//    workspace (scalar loads), and all lanes march parallel rays 4 px apart:
//    coherent pyramid accesses and near-uniform trip counts.
//  * Exact early exit: the side fade (:220-228) is zero for every sample
//    outside the open screen rectangle, the ray leaves that convex region
//    once and never re-enters, so every later step contributes exactly 0.
//  * Per step, the sample's world-space delta is evaluated from per-direction
//    dot products (the ray through q is affine in the step offset before
//    normalisation), the horizon test uses a monotone pseudo-angle, and the
//    atan2 plus double-angle terms run only when the horizon rises.
//  * Sample positions q = c + dir * off and the texel/level selection use the
//    shader's exact operation order (no contraction): the tables come from
//    the host in the shader's float arithmetic, so point fetches pick the same
//    texels as the scalar restatement.
//  * XCD-aware block order: the 16 slices of a tile run back to back on one
//    XCD, neighbouring tiles stay on that XCD's L2.
#include <limits.h>
#include <math.h>

#include <mutex>
#include <unordered_map>
#include <vector>

#include "lgx_internal.h"

namespace lgx {

// Buffer resource word 3 of the gfx9 family (raw, untyped 32-bit access).
static constexpr int kBufferConfig = 0x00020000;

template <class T> struct GatherArgs {
  T inv_vp[16];
  T cam[3];
  T e0[3], ex[3], ey[3];  // unnormalised ray through screen point s: e0 + s.x*ex + s.y*ey
  T W, H, near_step, log_ratio, ambient;
  int w, h, grid, grid_y, cells_x, cells_y, D, S_tab;  // pattern grid x grid_y
  int d_pow2;      // D is a power of two: x / D == x * inv_d exactly
  double inv_d;    // 1 / D
  T rW, rH;        // RN(1/W), RN(1/H) for the exact division by the extent
  const T *dirs;   // [grid * grid_y][D][4] = (cos, sin, 1 / cos, 1 / sin)
  const T *steps;  // [grid * grid_y][S_tab][2] = (offset, lod)
  const int *level_start;  // [grid * grid_y][LGX_MAX_LEVELS + 1]: first step at or above each Hi-Z level
  const float *depth;
  int depth_pitch;
  const void *normal;
  int normal_pitch;
  int n_levels;
  const float *pd[LGX_MAX_LEVELS];
  const void *pr[LGX_MAX_LEVELS];
  int pd_pitch[LGX_MAX_LEVELS], pr_pitch[LGX_MAX_LEVELS];
  int lw[LGX_MAX_LEVELS], lh[LGX_MAX_LEVELS];
  // Column-swizzled levels (SWZ kernels): texel column x of level l lives at
  // column (x & swz_mask[l]) * swz_w[l] + (x >> swz_shift[l]) of pd / pr
  // (mask and shift 0 on levels kept as they are)
  int swz_mask[LGX_MAX_LEVELS], swz_shift[LGX_MAX_LEVELS], swz_w[LGX_MAX_LEVELS];
  void *out;
  int out_pitch, out_slices;
  int *step_counts;
  int *hit_counts;  // COUNT only, may be null
  int tile_w, tile_h, tiles_x, n_blocks;
  int occ;   // host side: which register allocation to launch
  int pipe;  // host side: which step loop to launch
  int swz;   // host side: the fine levels are column-swizzled copies
  int sx0, sy0, lcells_x, lcells_y;  // slice cells covered by the launch (region of interest)
  int rx0, ry0, rx1, ry1;            // pixels gathered: [rx0, rx1) x [ry0, ry1)
};

template <class T> LGX_D T fast_rsqrt(T v);
template <> LGX_D float fast_rsqrt<float>(float v) { return __builtin_amdgcn_rsqf(v); }
template <> LGX_D double fast_rsqrt<double>(double v) { return 1.0 / sqrt(v); }

template <class T> LGX_D T fma_(T a, T b, T c) { return fma(a, b, c); }

template <class T> LGX_D T fast_rcp(T v);
template <> LGX_D float fast_rcp<float>(float v) { return __builtin_amdgcn_rcpf(v); }
template <> LGX_D double fast_rcp<double>(double v) { return 1.0 / v; }

// q / b for the viewport extent b, correctly rounded: y = RN(q * RN(1/b)),
// one fused correction with the exact residual q - y*b (Markstein).  The
// result equals the IEEE quotient for every q that can select a texel; this
// is checked exhaustively for the supported extents (tests/csrc/check_const_div.c),
// so texel selection and the on-screen test match the shader's q / viewport.
template <class T> LGX_D T div_extent(T q, T b, T rb) {
  if (sizeof(T) == 8) return q / b;
  const T y = q * rb;
  const T e = fma(-y, b, q);
  return fma(e, rb, y);
}

// atan2(y, x) for finite arguments (atan2(0, 0) = 0), float: octant
// reduction to t = min(|x|, |y|) / max(|x|, |y|) in [0, 1] and
// atan(t) = t P(t^2) with a degree-7 polynomial fitted to atan(t) / t on
// [0, 1] by iteratively reweighted least squares (max error 4e-8; 1.2e-7
// evaluated in float).  The hit branch of the march calls it on every horizon
// rise, where the library routine's handling of infinities and NaNs (which
// cannot occur there) costs more than the polynomial.
template <class T> LGX_D T atan2_finite(T y, T x);
template <> LGX_D double atan2_finite<double>(double y, double x) { return atan2(y, x); }
template <> LGX_D float atan2_finite<float>(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  const float t = mx > 0.f ? mn * __builtin_amdgcn_rcpf(mx) : 0.f;
  const float s = t * t;
  float p = -0.004054928198456764f;
  p = fmaf(p, s, 0.021864214912056923f);
  p = fmaf(p, s, -0.05591406673192978f);
  p = fmaf(p, s, 0.0964231863617897f);
  p = fmaf(p, s, -0.1390867531299591f);
  p = fmaf(p, s, 0.19946573674678802f);
  p = fmaf(p, s, -0.33329862356185913f);
  p = fmaf(p, s, 0.9999993443489075f);
  float r = t * p;
  r = ay > ax ? 1.5707963705062866f - r : r;
  r = x < 0.f ? 3.1415927410125732f - r : r;
  return copysignf(r, y);
}

// Radiance texel through a buffer resource (32-bit byte offset against a
// scalar base): the raw load, and its widening to float, so a fetch can be
// issued in one step of the march and consumed in the next.
template <int FMT> struct RadBuf;
template <> struct RadBuf<LGX_RGBA16F> {
  static constexpr int kBytes = 8;
  typedef unsigned int Raw __attribute__((ext_vector_type(2)));
  static LGX_D Raw raw(__amdgpu_buffer_rsrc_t r, uint32_t ti) {
    return __builtin_amdgcn_raw_buffer_load_b64(r, (int)(ti * 8u), 0, 0);
  }
  static LGX_D float4 load(__amdgpu_buffer_rsrc_t r, uint32_t ti) { return decode(raw(r, ti)); }
  static LGX_D float4 decode(Raw v) {
    return make_float4(h2f((uint16_t)(v[0] & 0xffffu)), h2f((uint16_t)(v[0] >> 16)), h2f((uint16_t)(v[1] & 0xffffu)),
                       h2f((uint16_t)(v[1] >> 16)));
  }
};
template <> struct RadBuf<LGX_RGBA32F> {
  static constexpr int kBytes = 16;
  typedef unsigned int Raw __attribute__((ext_vector_type(4)));
  static LGX_D Raw raw(__amdgpu_buffer_rsrc_t r, uint32_t ti) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, (int)(ti * 16u), 0, 0);
  }
  static LGX_D float4 load(__amdgpu_buffer_rsrc_t r, uint32_t ti) { return decode(raw(r, ti)); }
  static LGX_D float4 decode(Raw v) {
    // the whole vector is reinterpreted at once: a bit cast of one element
    // of a loaded vector (v[i]) compiles to a read of element 0 here
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    const f32x4 f = __builtin_bit_cast(f32x4, v);
    return make_float4(f.x, f.y, f.z, f.w);
  }
};

// Texel index iy * pitch + ix with the full-rate 24-bit multiply: rows and
// pitches of a level stay below 2^24.  (A speculative index of an
// off-screen step may wrap; the buffer resource's range check covers it.)
LGX_D uint32_t texel_index(int iy, int pitch, int ix) {
  return __umul24((unsigned)iy, (unsigned)pitch) + (uint32_t)ix;
}

// The running horizon maximum of a slice side and the test theta <
// theta_max of :254, without a division per step.  The pseudo-angle
// p(y, x) = 1 - x / s for y >= 0 and x / s - 1 for y < 0, s = |x| + |y|, is a
// monotone stand-in for atan2 on (-pi, pi]; it is copysign(s - x, y) / s, so
// with s > 0, p(y, x) < P is copysign(s - x, y) < s P: a subtraction, a
// sign copy, a product and a compare per step, the division only when the
// maximum P is replaced.  Only the order matters: near-ties decide steps
// whose contribution h(theta, theta_max) vanishes with the gap.  A zero
// vector has pseudo-angle 0 (atan2(0, 0) = 0) and is never below.
template <class T> struct HorizonMax {
  T P;
  LGX_D void set(T y, T x) {
    const T s = fabs(x) + fabs(y);
    P = s > T(0) ? copysign(s - x, y) * fast_rcp(s) : T(0);
  }
  LGX_D bool below(T y, T x) const {
    const T s = fabs(x) + fabs(y);
    return copysign(s - x, y) < s * P;
  }
};

// x / b for several x over one b, correctly rounded: r = RN(1/b) once (an
// IEEE division), then per quotient y = RN(x r) and one fused correction
// with the exact residual x - y b (Markstein), which rounds to the IEEE
// quotient when r is the correctly rounded reciprocal (checked against the
// IEEE division for every significand of b, tests/csrc/check_common_div.c);
// x = 0 keeps its sign as the IEEE quotient does.  Three quotients cost one
// division and three short corrections instead of three divisions.
template <class T> LGX_D T div_by(T x, T b, T r) {
  if (sizeof(T) == 8) return x / b;
  const T y = x * r;
  const T e = fma(-y, b, x);
  const T q = fma(e, r, y);
  return x == T(0) ? x : q;
}
template <class T> LGX_D Vec3<T> normalize3_cr(Vec3<T> v) {
  const T len = sqrt(dot3(v, v));
  const T r = T(1) / len;
  return Vec3<T>{div_by(v.x, len, r), div_by(v.y, len, r), div_by(v.z, len, r)};
}

template <class T, class ARGS> LGX_D Vec3<T> unproject_far(ARGS &a, T sx, T sy) {
  // indirectLighting.frag:20-25 with screenPos.z = 1 (:241)
  const auto &m = a.inv_vp;
  T x = sx * T(2) - T(1), y = sy * T(2) - T(1);
  T vx = m[0] * x + m[4] * y + m[8] * T(1) + m[12] * T(1);
  T vy = m[1] * x + m[5] * y + m[9] * T(1) + m[13] * T(1);
  T vz = m[2] * x + m[6] * y + m[10] * T(1) + m[14] * T(1);
  T vw = m[3] * x + m[7] * y + m[11] * T(1) + m[15] * T(1);
  const T rw = T(1) / vw;
  return Vec3<T>{div_by(vx, vw, rw), div_by(vy, vw, rw), div_by(vz, vw, rw)};
}

template <int FMT> struct RadLoad {
  static LGX_D float4 at(const void *base, size_t i) { return Texel<FMT>::load(base, i); }
};

// Vulkan clamp-to-edge bilinear in one level (u = s*W - 1/2).
template <class T, int FMT>
LGX_D void bilinear(const void *base, int pitch, int lw, int lh, T sx, T sy, T *out, int nch) {
  T u = sx * T(lw) - T(0.5), v = sy * T(lh) - T(0.5);
  T fu = floor(u), fv = floor(v);
  T al = u - fu, be = v - fv;
  int i0 = (int)fu, j0 = (int)fv;
  int x0 = min(max(i0, 0), lw - 1), x1 = min(max(i0 + 1, 0), lw - 1);
  int y0 = min(max(j0, 0), lh - 1), y1 = min(max(j0 + 1, 0), lh - 1);
  float4 p00 = Texel<FMT>::load(base, (size_t)y0 * pitch + x0);
  float4 p10 = Texel<FMT>::load(base, (size_t)y0 * pitch + x1);
  float4 p01 = Texel<FMT>::load(base, (size_t)y1 * pitch + x0);
  float4 p11 = Texel<FMT>::load(base, (size_t)y1 * pitch + x1);
  const float *a00 = &p00.x, *a10 = &p10.x, *a01 = &p01.x, *a11 = &p11.x;
  for (int k = 0; k < nch; k++) {
    T top = T(a00[k]) * (T(1) - al) + T(a10[k]) * al;
    T bot = T(a01[k]) * (T(1) - al) + T(a11[k]) * al;
    out[k] = top * (T(1) - be) + bot * be;
  }
}

template <class T, int FMT, class LV, class PI>
LGX_D void sample_trilinear(const LV &lv, const PI &pitch, const PI &lw, const PI &lh, int nl, T sx, T sy,
                            T lod, T *out, int nch) {
  T q = T(nl - 1);
  T d = lod;
  if (!(d > T(0))) d = T(0);
  if (d > q) d = q;
  int l0 = (int)floor(d);
  int l1 = min(l0 + 1, nl - 1);
  T delta = d - T(l0);
  T a[4], b[4];
  bilinear<T, FMT>(lv[l0], pitch[l0], lw[l0], lh[l0], sx, sy, a, nch);
  if (delta > T(0) && l1 != l0) {
    bilinear<T, FMT>(lv[l1], pitch[l1], lw[l1], lh[l1], sx, sy, b, nch);
    for (int k = 0; k < nch; k++) out[k] = a[k] * (T(1) - delta) + b[k] * delta;
  } else {
    for (int k = 0; k < nch; k++) out[k] = a[k];
  }
}

template <class T> LGX_D int point_level(T lod, int nl) {
  int l = 0;
  if (lod > T(0)) l = (int)min(floor(lod), T(nl - 1));
  return l;
}
template <class T> LGX_D size_t point_index(T sx, T sy, int lw, int lh, int pitch) {
  T u = floor(sx * T(lw)), v = floor(sy * T(lh));
  int i = (int)min(max(u, T(0)), T(lw - 1));
  int j = (int)min(max(v, T(0)), T(lh - 1));
  return (size_t)j * pitch + i;
}

// T: per-step arithmetic; A: accumulation of the slice contributions and of
// the direction sum (float/float, double/double, or float/double).
// One workgroup's share: logical block b (slice k = b % (grid grid_y), tile b / (grid grid_y)).
// ARGS: GatherArgs<T> as the kernel's argument, or in constant memory
// (address space 4, the multi-view kernel's per-view table): either way the
// compiler reads its fields as scalars.
template <class T, class A, int SAMPLER, int RFMT, int NFMT, bool COUNT, bool PIPE, bool SWZ, class ARGS>
LGX_D void gather_block(ARGS &a, const uint32_t b) {
  const int g2 = a.grid * a.grid_y;
  const int k = (int)(b % (uint32_t)g2);
  const int t = (int)(b / (uint32_t)g2);
  const int kx = k % a.grid, ky = k / a.grid;
  const int lx = (t % a.tiles_x) * a.tile_w + (int)(threadIdx.x % (uint32_t)a.tile_w);
  const int ly = (t / a.tiles_x) * a.tile_h + (int)(threadIdx.x / (uint32_t)a.tile_w);
  if (lx >= a.lcells_x || ly >= a.lcells_y) return;
  const int sx = a.sx0 + lx, sy = a.sy0 + ly;
  const int x = sx * a.grid + kx, y = sy * a.grid_y + ky;
  if (x < a.rx0 || x >= a.rx1 || y < a.ry0 || y >= a.ry1) return;

  const Vec3<T> cam{a.cam[0], a.cam[1], a.cam[2]};
  const T cx = T(x) + T(0.5), cy = T(y) + T(0.5);  // gl_FragCoord (:116)
  const T scx = cx / a.W, scy = cy / a.H;           // :118
  const T rc = T(a.depth[(size_t)y * a.depth_pitch + x]);
  const float4 nt = Texel<NFMT>::load(a.normal, (size_t)y * a.normal_pitch + x);
  const Vec3<T> n = oct_decode<T>(T(nt.x), T(nt.y));  // :133
  const Vec3<T> rayc = normalize3_cr(unproject_far(a, scx, scy) - cam);
  const Vec3<T> P = cam + rayc * rc;  // :130 from the radial distance
  const Vec3<T> eye = normalize3_cr(cam - P);  // :182
  const Vec3<T> pdir = normalize3_cr(P - cam);
  const T ambient = a.ambient;
  // Ray through the sample, unnormalised, at the pixel centre (U) and per
  // unit of step offset along a direction (V is formed per direction).
  const Vec3<T> E0{a.e0[0], a.e0[1], a.e0[2]}, EX{a.ex[0], a.ex[1], a.ex[2]}, EY{a.ey[0], a.ey[1], a.ey[2]};
  const Vec3<T> U = E0 + EX * scx + EY * scy;
  const T erc = rc * dot3(eye, rayc);
  const T en = dot3(eye, n);

  const T *dirs = a.dirs + (size_t)k * a.D * 4;
  const T *steps = a.steps + (size_t)k * a.S_tab * 2;
  A sum0 = 0, sum1 = 0, sum2 = 0;
  int step_total = 0, hit_total = 0;

  for (int d = 0; d < a.D; d++) {
    // the direction and its reciprocals (the host's correctly rounded 1 / d)
    const T dx = dirs[4 * d], dy = dirs[4 * d + 1], idx = dirs[4 * d + 2], idy = dirs[4 * d + 3];
    // :180-183 tangent of the slice through the offset point (the quotient
    // by the extent is the IEEE one, as for the step positions)
    const T ox = div_extent(cx + dx * T(10e-1), a.W, a.rW), oy = div_extent(cy + dy * T(10e-1), a.H, a.rH);
    const Vec3<T> offp = cam + normalize3_cr(unproject_far(a, ox, oy) - cam) * rc;
    const Vec3<T> tangent = normalize3_cr(normalize3_cr(offp - cam) - pdir);
    // :193-198 normal-limited initial horizon
    const Vec3<T> c1 = cross3(tangent, eye);
    const Vec3<T> dn = cross3(Vec3<T>{-c1.x, -c1.y, -c1.z}, n);
    const T hy = dot3(dn, tangent), hx = dot3(dn, eye);
    T max_h = atan2_finite(hy, hx);
    HorizonMax<T> hm;
    hm.set(hy, hx);
    const T iq2 = fast_rcp(hx * hx + hy * hy);
    T c2max = (hx * hx - hy * hy) * iq2, s2max = (T(2) * hx * hy) * iq2;
    const T tn = dot3(tangent, n);
    // :202-204 path to the screen border
    T path;
    {
      const T ix = idx, iy = idy;
      T t1 = (T(0) - cx) * ix, t2 = (a.W - cx) * ix, t3 = (T(0) - cy) * iy, t4 = (a.H - cy) * iy;
      path = fabs(min(max(t1, t2), max(t3, t4)));
    }
    // :209 ambient term: h(0, max_h)
    T h0 = T(0.25) * en * (T(1) - c2max) + T(0.25) * tn * (T(2) * max_h - s2max);
    A L0 = A(ambient * h0), L1 = L0, L2 = L0;
    // :212 step count, capped by the table (S)
    int iters = (int)(log(path / a.near_step) / a.log_ratio) + 1;
    iters = min(iters, a.S_tab);
    // per-direction constants of the fast delta evaluation
    const Vec3<T> V = EX * (dx / a.W) + EY * (dy / a.H);
    const T uu = dot3(U, U), uv2 = T(2) * dot3(U, V), vv = dot3(V, V);
    const T tu = dot3(tangent, U), tv = dot3(tangent, V);
    const T eu = dot3(eye, U), ev = dot3(eye, V);
    const T trc = rc * dot3(tangent, rayc);
    int j = 0;
    if (SAMPLER == LGX_SAMPLER_HIZ_POINT) {
      // Level-major march.  The step index j is the same on every active
      // lane (all start at 0 and advance together; lanes only leave), so
      // it is read as a scalar; the Hi-Z level floor(lod) never decreases
      // along a ray, so the level changes at table-given steps and its base,
      // pitch and extent stay in scalar registers for its run of steps.
      // Two step loops with the same arithmetic in the same order (so the
      // same results).  PIPE: fetches overlap arithmetic -- a step's depth
      // texel is requested while the previous step's hit (angle, weights,
      // accumulation) is computed, and a hit's radiance texel is requested
      // before that work; the depth request of the step after a lane's last
      // one is discarded (its offset may lie outside the level, for a ray
      // leaving the screen, where the buffer resource's range check returns
      // 0 instead of reading).  It hides more latency for a few more
      // instructions per step, which pays on long marches (few directions, a
      // wide footprint) and not on short ones; the launch-shape autotuner
      // picks per configuration.  Otherwise each step fetches and then waits.
      const int *lstart = a.level_start + (size_t)k * (LGX_MAX_LEVELS + 1);
      int l = 0;
      int lnext = a.n_levels > 1 ? lstart[1] : INT_MAX;  // first step of the next level
      // the level's depth and radiance as buffer resources: per-step fetches
      // take a 32-bit byte offset against a scalar base
      auto depth_rsrc = [&](int lv) {
        return __builtin_amdgcn_make_buffer_rsrc((void *)a.pd[lv], (short)0, a.lh[lv] * a.pd_pitch[lv] * 4,
                                                 kBufferConfig);
      };
      auto rad_rsrc = [&](int lv) {
        return __builtin_amdgcn_make_buffer_rsrc((void *)a.pr[lv], (short)0,
                                                 a.lh[lv] * a.pr_pitch[lv] * RadBuf<RFMT>::kBytes, kBufferConfig);
      };
      __amdgpu_buffer_rsrc_t drs = depth_rsrc(0), rrs = rad_rsrc(0);
      T flw = T(a.lw[0]), flh = T(a.lh[0]);
      int lw1 = a.lw[0] - 1, lh1 = a.lh[0] - 1, dpitch = a.pd_pitch[0], rpitch = a.pr_pitch[0];
      int smask = 0, sshift = 0, ssw = 0;
      if constexpr (SWZ) {
        smask = a.swz_mask[0];
        sshift = a.swz_shift[0];
        ssw = a.swz_w[0];
      }
      // the storage column of texel column ix on the current level
      auto col = [&](int ix) {
        if constexpr (SWZ) return (int)(__umul24((unsigned)(ix & smask), (unsigned)ssw) + (unsigned)(ix >> sshift));
        return ix;
      };
      // the fetch of step s at offset o on the current level: on-screen flag,
      // q / extent, depth request, radiance texel index
      struct Fetch {
        float rd;
        T sqx, sqy;
        uint32_t rti;
        bool on;
      };
      auto request = [&](int s, T o) {
        Fetch f;
        const T qx = cx + dx * o, qy = cy + dy * o;  // :218
        // :219-228 on screen: q / extent in (0, 1) is q in (0, extent), as
        // the quotient is correctly rounded (it reaches 1 only from q = extent)
        // (bitwise &: one mask, not a chain of divergent branches)
        f.on = (s < iters) & (qx > T(0)) & (qx < a.W) & (qy > T(0)) & (qy < a.H);
        f.sqx = div_extent(qx, a.W, a.rW);  // :219
        f.sqy = div_extent(qy, a.H, a.rH);
        // sq > 0 on screen, so sq * extent >= 0 and truncation is floor; the
        // product can round up to the extent itself, hence the upper clamp
        const int ix = col(min((int)(f.sqx * flw), lw1));
        const int iy = min((int)(f.sqy * flh), lh1);
        f.rti = texel_index(iy, rpitch, ix);
        const uint32_t ti = texel_index(iy, dpitch, ix);
        f.rd = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(drs, (int)(ti * 4u), 0, 0));
        return f;
      };
      // the level of step s (uniform; s never decreases)
      auto to_level = [&](int s) {
        while (s >= lnext) {
          l++;
          lnext = l + 1 < a.n_levels ? lstart[l + 1] : INT_MAX;
          drs = depth_rsrc(l);
          rrs = rad_rsrc(l);
          dpitch = a.pd_pitch[l];
          rpitch = a.pr_pitch[l];
          flw = T(a.lw[l]);
          flh = T(a.lh[l]);
          lw1 = a.lw[l] - 1;
          lh1 = a.lh[l] - 1;
          if constexpr (SWZ) {
            smask = a.swz_mask[l];
            sshift = a.swz_shift[l];
            ssw = a.swz_w[l];
          }
        }
      };
      if constexpr (PIPE) {
        T off = steps[0];
        T off_n = steps[2 * min(1, a.S_tab - 1)];  // the next step's offset (scalar)
        to_level(0);
        Fetch f = request(0, off);
        for (; f.on; j++) {
          const int ju = __builtin_amdgcn_readfirstlane(j);
          // step j: :241-252 sample delta projected on (eye, tangent)
          const T r = T(f.rd);
          const T len2 = fma_(off, fma_(off, vv, uv2), uu);
          const T ri = r * fast_rsqrt(len2);
          const T hyv = fma_(ri, fma_(off, tv, tu), -trc);
          const T hxv = fma_(ri, fma_(off, ev, eu), -erc);
          const bool hit = hm.below(hyv, hxv);  // :254
          typename RadBuf<RFMT>::Raw rv;
          if (hit) rv = RadBuf<RFMT>::raw(rrs, f.rti);  // :256, on step j's level
          const T sqx = f.sqx, sqy = f.sqy;
          // step j + 1: its level, then its depth request
          off = off_n;
          off_n = steps[2 * min(ju + 2, a.S_tab - 1)];  // scalar load, in flight for a step
          to_level(ju + 1);
          f = request(ju + 1, off);
          if (hit) {
            if (COUNT) hit_total++;
            const T th = atan2_finite(hyv, hxv);
            const T iq = fast_rcp(hxv * hxv + hyv * hyv);
            const T c2 = (hxv * hxv - hyv * hyv) * iq, s2 = (T(2) * hxv * hyv) * iq;
            T side = T(1);
            {
              const T width = T(0.1f), inv_width = T(1) / width;  // :222-227
              side *= min(max((T(1) - sqx) * inv_width, T(0)), T(1));
              side *= min(max(sqx * inv_width, T(0)), T(1));
              side *= min(max((T(1) - sqy) * inv_width, T(0)), T(1));
              side *= min(max(sqy * inv_width, T(0)), T(1));
            }
            const T hc =
                (T(0.25) * en * (c2 - c2max) + T(0.25) * tn * (T(2) * max_h - T(2) * th - s2max + s2)) * side;
            const float4 v = RadBuf<RFMT>::decode(rv);
            L0 += A(T(v.x) * hc);  // :261
            L1 += A(T(v.y) * hc);
            L2 += A(T(v.z) * hc);
            L0 -= A(ambient * hc);  // :262
            L1 -= A(ambient * hc);
            L2 -= A(ambient * hc);
            max_h = th;  // :263
            hm.set(hyv, hxv);
            c2max = c2;
            s2max = s2;
          }
        }
      } else {
        T off = steps[0];
        for (;; j++) {
          if (j >= iters) break;
          const int ju = __builtin_amdgcn_readfirstlane(j);
          // the next step's offset: a scalar load in flight during this step
          const T off_next = steps[2 * min(ju + 1, a.S_tab - 1)];
          to_level(ju);
          const T qx = cx + dx * off, qy = cy + dy * off;  // :218
          // :219-228 on screen: q / extent in (0, 1) is q in (0, extent), as
          // the quotient is correctly rounded (it reaches 1 only from q = extent)
          // (bitwise &: one mask, not a chain of divergent branches)
          if (!((qx > T(0)) & (qx < a.W) & (qy > T(0)) & (qy < a.H))) break;
          const T sqx = div_extent(qx, a.W, a.rW), sqy = div_extent(qy, a.H, a.rH);  // :219
          // sq > 0, so sq * extent >= 0 and truncation is floor; the product can
          // round up to the extent itself, hence the upper clamp (in integers)
          const int ix = col(min((int)(sqx * flw), lw1));
          const int iy = min((int)(sqy * flh), lh1);
          const uint32_t ti = texel_index(iy, dpitch, ix);
          const T r = T(__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(drs, (int)(ti * 4u), 0, 0)));
          // :241-252 sample delta projected on (eye, tangent)
          const T len2 = fma_(off, fma_(off, vv, uv2), uu);
          const T ri = r * fast_rsqrt(len2);
          const T hyv = fma_(ri, fma_(off, tv, tu), -trc);
          const T hxv = fma_(ri, fma_(off, ev, eu), -erc);
          if (hm.below(hyv, hxv)) {  // :254
            if (COUNT) hit_total++;
            const T th = atan2_finite(hyv, hxv);
            const T iq = fast_rcp(hxv * hxv + hyv * hyv);
            const T c2 = (hxv * hxv - hyv * hyv) * iq, s2 = (T(2) * hxv * hyv) * iq;
            T side = T(1);
            {
              const T width = T(0.1f), inv_width = T(1) / width;  // :222-227
              side *= min(max((T(1) - sqx) * inv_width, T(0)), T(1));
              side *= min(max(sqx * inv_width, T(0)), T(1));
              side *= min(max((T(1) - sqy) * inv_width, T(0)), T(1));
              side *= min(max(sqy * inv_width, T(0)), T(1));
            }
            const T hc =
                (T(0.25) * en * (c2 - c2max) + T(0.25) * tn * (T(2) * max_h - T(2) * th - s2max + s2)) * side;
            const float4 v = RadBuf<RFMT>::load(rrs, texel_index(iy, rpitch, ix));  // :256
            L0 += A(T(v.x) * hc);  // :261
            L1 += A(T(v.y) * hc);
            L2 += A(T(v.z) * hc);
            L0 -= A(ambient * hc);  // :262
            L1 -= A(ambient * hc);
            L2 -= A(ambient * hc);
            max_h = th;  // :263
            hm.set(hyv, hxv);
            c2max = c2;
            s2max = s2;
          }
          off = off_next;
        }
      }
    } else {
      for (; j < iters; j++) {
        const T off = steps[2 * j], lod = steps[2 * j + 1];
        const T qx = cx + dx * off, qy = cy + dy * off;  // :218
        const T sqx = qx / a.W, sqy = qy / a.H;          // :219
        if (!(sqx > T(0) && sqx < T(1) && sqy > T(0) && sqy < T(1))) break;
        T r;
        if (SAMPLER == LGX_SAMPLER_HIZ_POINT) {
          const int l = point_level(lod, a.n_levels);
          r = T(a.pd[l][point_index(sqx, sqy, a.lw[l], a.lh[l], a.pd_pitch[l])]);
        } else {
          T tmp[4];
          sample_trilinear<T, LGX_R32F>(a.pd, a.pd_pitch, a.lw, a.lh, a.n_levels, sqx, sqy, lod,
                                        tmp, 1);
          r = tmp[0];
        }
        // :241-252 sample delta projected on (eye, tangent)
        const T len2 = fma_(off, fma_(off, vv, uv2), uu);
        const T ri = r * fast_rsqrt(len2);
        const T hyv = ri * fma_(off, tv, tu) - trc;
        const T hxv = ri * fma_(off, ev, eu) - erc;
        if (hm.below(hyv, hxv)) {  // :254
            if (COUNT) hit_total++;
          const T th = atan2(hyv, hxv);
          const T qq = hxv * hxv + hyv * hyv;
          const T c2 = (hxv * hxv - hyv * hyv) / qq, s2 = (T(2) * hxv * hyv) / qq;
          T side = T(1);
          {
            const T width = T(0.1f), inv_width = T(1) / width;  // :222-227
            side *= min(max((T(1) - sqx) * inv_width, T(0)), T(1));
            side *= min(max(sqx * inv_width, T(0)), T(1));
            side *= min(max((T(1) - sqy) * inv_width, T(0)), T(1));
            side *= min(max(sqy * inv_width, T(0)), T(1));
          }
          const T hc = (T(0.25) * en * (c2 - c2max) + T(0.25) * tn * (T(2) * max_h - T(2) * th - s2max + s2)) * side;
          T ls[4];
          if (SAMPLER == LGX_SAMPLER_HIZ_POINT) {
            const int l = point_level(lod, a.n_levels);
            float4 v = Texel<RFMT>::load(a.pr[l], point_index(sqx, sqy, a.lw[l], a.lh[l], a.pr_pitch[l]));
            ls[0] = T(v.x);
            ls[1] = T(v.y);
            ls[2] = T(v.z);
          } else {
            sample_trilinear<T, RFMT>(a.pr, a.pr_pitch, a.lw, a.lh, a.n_levels, sqx, sqy, lod, ls,
                                      3);
          }
          L0 += A(ls[0] * hc);  // :261
          L1 += A(ls[1] * hc);
          L2 += A(ls[2] * hc);
          L0 -= A(ambient * hc);  // :262
          L1 -= A(ambient * hc);
          L2 -= A(ambient * hc);
          max_h = th;  // :263
          hm.set(hyv, hxv);
          c2max = c2;
          s2max = s2;
        }
      }
    }
    if (COUNT) step_total += j;
    // :268; for a power-of-two D the quotient is the product with 1 / D
    // (exact), otherwise the IEEE division
    if (a.d_pow2) {
      sum0 += A(2) * L0 * A(a.inv_d);
      sum1 += A(2) * L1 * A(a.inv_d);
      sum2 += A(2) * L2 * A(a.inv_d);
    } else {
      sum0 += A(2) * L0 / A(a.D);
      sum1 += A(2) * L1 / A(a.D);
      sum2 += A(2) * L2 / A(a.D);
    }
  }

  int ox = x, oy = y;
  if (a.out_slices) {
    ox = kx * a.cells_x + sx;
    oy = ky * a.cells_y + sy;
  }
  const size_t o = (size_t)oy * a.out_pitch + ox;
  if (sizeof(A) == 4) {
    ((float4 *)a.out)[o] = make_float4((float)sum0, (float)sum1, (float)sum2, 1.0f);  // :270
  } else {
    double *po = (double *)a.out + o * 4;
    po[0] = (double)sum0;
    po[1] = (double)sum1;
    po[2] = (double)sum2;
    po[3] = 1.0;
  }
  if (COUNT) {
    a.step_counts[(size_t)y * a.w + x] = step_total;
    if (a.hit_counts) a.hit_counts[(size_t)y * a.w + x] = hit_total;
  }
}

// OCC: minimum waves per SIMD the register allocation must allow (1 = the
// compiler's choice); the launch shapes offer 1 and 8.  PIPE: the step loop
// that overlaps fetches with arithmetic (Hi-Z sampler).
template <class T, class A, int SAMPLER, int RFMT, int NFMT, bool COUNT, int OCC, bool PIPE = false, bool SWZ = false>
__global__ __launch_bounds__(256, OCC) void gather_kernel(const GatherArgs<T> a) {
  gather_block<T, A, SAMPLER, RFMT, NFMT, COUNT, PIPE, SWZ>(a, xcd_remap(blockIdx.x, (uint32_t)a.n_blocks));
}

// Column-swizzled copies of the fine pyramid levels for the SWZ kernels.
// With a pattern of G columns, the lanes of a wave (one slice) sample points
// G pixels apart, so on level l (G >> l >= 2) their texel columns share their
// residue mod F = G >> l: stored as F sub-images of every F-th column
// (column x at (x mod F) * sw + x / F), a wave's texels are contiguous.
struct SwzJob {
  const void *src;
  void *dst;
  int src_pitch, dst_pitch, w, h, mask, shift, sw, bytes;  // pitches in texels; bytes per texel: 4, 8 or 16
};
struct SwzJobs {
  SwzJob j[2 * LGX_MAX_LEVELS];
  int n;
};
// One launch for all jobs: blockIdx.y is the job, workgroups stride over its
// rows, the threads over a row's columns.
LGX_D void swizzle_job(const SwzJob &j) {
  for (int y = blockIdx.x; y < j.h; y += gridDim.x) {
    const size_t srow = (size_t)y * j.src_pitch, drow = (size_t)y * j.dst_pitch;
    for (int x = threadIdx.x; x < j.w; x += 256) {
      const size_t si = srow + x, di = drow + (x & j.mask) * j.sw + (x >> j.shift);
      if (j.bytes == 4)
        ((uint32_t *)j.dst)[di] = ((const uint32_t *)j.src)[si];
      else if (j.bytes == 8)
        ((uint2 *)j.dst)[di] = ((const uint2 *)j.src)[si];
      else
        ((uint4 *)j.dst)[di] = ((const uint4 *)j.src)[si];
    }
  }
}
__global__ __launch_bounds__(256) void swizzle_levels_kernel(const SwzJobs jobs) { swizzle_job(jobs.j[blockIdx.y]); }
// the multi-view gather's jobs, in device memory
__global__ __launch_bounds__(256) void swizzle_jobs_kernel(const SwzJob *__restrict__ jobs) {
  swizzle_job(jobs[blockIdx.y]);
}

// Persistent gather over several views: the logical blocks of all views form
// one list (view-major, first[v] = index of view v's first block, first[V] =
// total).  The list is cut into 8 contiguous chunks, one per XCD (workgroup
// w runs on XCD w % 8, gridDim.x is a multiple of 8), so neighbouring tiles
// and the 16 slices of a tile share one L2.  Within a chunk a workgroup takes
// the next block from the chunk's counter (zeroed before the launch): blocks
// differ in cost (early exits, view content), and a fixed stride would leave
// workgroups idle at the tail.  A workgroup whose own chunk is exhausted
// takes blocks from the other XCDs' chunks in turn (the counters only grow,
// so every block is taken once), and ends when all eight are exhausted.
template <class T, class A, int SAMPLER, int RFMT, int NFMT, int OCC, bool PIPE, bool SWZ>
__global__ __launch_bounds__(256, OCC) void gather_multi_kernel(const GatherArgs<T> *__restrict__ views,
                                                           const uint32_t *__restrict__ first, int n_views,
                                                           uint32_t *next) {
  __shared__ uint32_t s_item;
  // the views' arguments are only read: constant memory, so their fields
  // (with a uniform view index) become scalar loads, as a kernel argument's are
  typedef const GatherArgs<T> __attribute__((address_space(4))) CArgs;
  CArgs *cviews = (CArgs *)views;
  const uint32_t X = 8;
  const uint32_t total = first[n_views];
  const uint32_t xcd = blockIdx.x % X;
  const uint32_t chunk = (total + X - 1) / X;
  for (uint32_t k = 0; k < X; k++) {
    const uint32_t y = (xcd + k) % X;  // own chunk first, then the others
    const uint32_t c0 = min(total, y * chunk), c1 = min(total, c0 + chunk);
    int v = 0;
    for (;;) {
      if (threadIdx.x == 0) s_item = c0 + atomicAdd(&next[y], 1u);
      __syncthreads();
      const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
      __syncthreads();  // every wave has read s_item before lane 0 overwrites it
      if (it >= c1) break;
      while (v + 1 < n_views && it >= first[v + 1]) v++;  // items of one chunk only grow
      gather_block<T, A, SAMPLER, RFMT, NFMT, false, PIPE, SWZ>(cviews[v], it - first[v]);
    }
  }
}

__global__ void zero_u32_kernel(uint32_t *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

// ---------------------------------------------------------------- host side

struct GatherTableShape {
  int g2, D, S_tab;
  size_t elem;
};

static void hammersley(int i, int n, float &aa, float &bb) {
  uint32_t v = ((uint32_t)i << 16u) | ((uint32_t)i >> 16u);
  v = ((v & 0x55555555u) << 1u) | ((v & 0xAAAAAAAAu) >> 1u);
  v = ((v & 0x33333333u) << 2u) | ((v & 0xCCCCCCCCu) >> 2u);
  v = ((v & 0x0F0F0F0Fu) << 4u) | ((v & 0xF0F0F0F0u) >> 4u);
  v = ((v & 0x00FF00FFu) << 8u) | ((v & 0xFF00FF00u) >> 8u);
  aa = (float)i / (float)n;
  bb = (float)v / (float)0xffffffffu;
}

// direction count that sets the step growth and the lod
static int footprint(const lgx_gather_params *p) { return p->footprint_dirs > 0 ? p->footprint_dirs : p->dirs; }

template <class T> static int table_steps(const lgx_gather_params *p) {
  const T pi = T(3.1415f);
  const T W = T(p->width), H = T(p->height);
  const T near_step = W / T(1000);
  const T ratio = T(2) * pi / T(footprint(p)) + T(1);
  const T diag = std::sqrt(W * W + H * H);
  long it = (long)(std::log(diag / near_step) / std::log(ratio)) + 2;
  if (p->max_steps > 0 && it > p->max_steps) it = p->max_steps;
  if (it < 1) it = 1;
  return (int)it;
}

static bool gather_params_ok(const lgx_gather_params *p) {
  return p && p->width > 0 && p->height > 0 && p->dirs >= 1 && p->dirs <= LGX_MAX_DIRS &&
         p->footprint_dirs <= LGX_MAX_DIRS && p->grid >= 1 && p->grid <= 16 && p->grid_y <= 16 &&
         p->grid * (p->grid_y > 0 ? p->grid_y : p->grid) <= 64 &&
         (p->sampler == LGX_SAMPLER_TRILINEAR || p->sampler == LGX_SAMPLER_HIZ_POINT) &&
         (p->precision == LGX_PREC_F32 || p->precision == LGX_PREC_F64 || p->precision == LGX_PREC_F32_ACC64) &&
         (p->out_layout == LGX_OUT_INTERLEAVED || p->out_layout == LGX_OUT_SLICES);
}

static GatherTableShape table_shape(const lgx_gather_params *p) {
  GatherTableShape s;
  s.g2 = p->grid * (p->grid_y > 0 ? p->grid_y : p->grid);
  s.D = p->dirs;
  s.S_tab = p->precision == LGX_PREC_F64 ? table_steps<double>(p) : table_steps<float>(p);
  s.elem = p->precision == LGX_PREC_F64 ? sizeof(double) : sizeof(float);
  return s;
}

// What a workspace's tables were computed for: every parameter the tables
// depend on, normalised (grid_y <= 0 is grid, footprint <= 0 is dirs; F32 and
// F32_ACC64 share the float tables).
static bool uses_swizzle(const lgx_gather_params *p);

struct TableKey {
  int32_t width, height, dirs, max_steps, grid, grid_y, footprint, f64, swz;
  bool operator==(const TableKey &o) const {
    return width == o.width && height == o.height && dirs == o.dirs && max_steps == o.max_steps && grid == o.grid &&
           grid_y == o.grid_y && footprint == o.footprint && f64 == o.f64 && swz == o.swz;
  }
};

static TableKey table_key(const lgx_gather_params *p) {
  TableKey k;
  k.width = p->width;
  k.height = p->height;
  k.dirs = p->dirs;
  k.max_steps = p->max_steps > 0 ? p->max_steps : 0;
  k.grid = p->grid;
  k.grid_y = p->grid_y > 0 ? p->grid_y : p->grid;
  k.footprint = p->footprint_dirs > 0 ? p->footprint_dirs : p->dirs;
  k.f64 = p->precision == LGX_PREC_F64;
  k.swz = uses_swizzle(p);  // the workspace then also holds the swizzled levels
  return k;
}

// Workspaces filled by lgx_gather_prepare, by device address.  lgx_gather
// and lgx_gather_batch_create reject a workspace that was not prepared for
// the call's parameters (a smaller table would be read past its end); one
// prepared for a swizzling shape also serves the others (same tables, more
// room).
static std::mutex g_ws_mu;
static std::unordered_map<const void *, TableKey> g_ws;

static void record_workspace(const void *ws, const lgx_gather_params *p) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  g_ws[ws] = table_key(p);
}

static bool workspace_prepared(const void *ws, const lgx_gather_params *p) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  auto it = g_ws.find(ws);
  if (it == g_ws.end()) return false;
  TableKey need = table_key(p);
  if (it->second.swz) need.swz = 1;
  return it->second == need;
}

// Tables in the shader's own arithmetic and evaluation order
// (indirectLighting.frag:161-177 directions, :217 offsets, :234 lod), then
// per slice the first step of each Hi-Z level floor(lod) (clamped at 0).
// Returns false if a slice's level sequence ever decreases (the level-major
// march requires it non-decreasing; it is, as off and lod grow with j).
template <class T>
static bool fill_tables(const lgx_gather_params *p, const GatherTableShape &s, std::vector<unsigned char> &out) {
  const T pi = T(3.1415f);
  const T W = T(p->width);
  const T near_step = W / T(1000);
  const int D = p->dirs;
  const size_t nd = (size_t)s.g2 * D * 4, ns = (size_t)s.g2 * s.S_tab * 2;
  out.assign((nd + ns) * sizeof(T) + (size_t)s.g2 * (LGX_MAX_LEVELS + 1) * sizeof(int32_t), 0);
  T *dirs = (T *)out.data();
  T *steps = dirs + nd;
  int32_t *lstart = (int32_t *)(steps + ns);
  for (int k = 0; k < s.g2; k++) {
    float angf, linf;
    hammersley(k, s.g2, angf, linf);
    const T ang_offset = T(angf), lin_offset = T(linf);
    const T pixel_ang_offset = T(2) * pi / T(D) * ang_offset;
    for (int d = 0; d < D; d++) {
      T screen_ang = pixel_ang_offset + T(2) * pi / T(D) * T(d);
      const T c = std::cos(screen_ang), sn = std::sin(screen_ang);
      T *e = dirs + ((size_t)k * D + d) * 4;
      e[0] = c;
      e[1] = sn;
      e[2] = T(1) / c;  // the march's box test divides by the direction (:202)
      e[3] = T(1) / sn;
    }
    const T Df = T(footprint(p));
    const T ratio = T(2) * pi / Df + T(1);
    int prev_level = 0;
    for (int l = 0; l <= LGX_MAX_LEVELS; l++) lstart[k * (LGX_MAX_LEVELS + 1) + l] = l == 0 ? 0 : s.S_tab;
    for (int j = 0; j < s.S_tab; j++) {
      T off = near_step * std::pow(ratio, T(j) + lin_offset) + T(1) - near_step;
      T lod = std::log(std::max(T(0), T(2) * pi / Df * (off - T(1)) * T(0.5f))) / std::log(T(2)) + T(-2);
      steps[((size_t)k * s.S_tab + j) * 2 + 0] = off;
      steps[((size_t)k * s.S_tab + j) * 2 + 1] = lod;
      int level = lod > T(0) ? (int)std::min(std::floor(lod), T(LGX_MAX_LEVELS)) : 0;
      if (level < prev_level) return false;
      for (int l = prev_level + 1; l <= level; l++) lstart[k * (LGX_MAX_LEVELS + 1) + l] = j;
      prev_level = std::max(prev_level, level);
    }
  }
  return true;
}

template <class T> static void frame_constants(const lgx_gather_params *p, GatherArgs<T> &a) {
  Mat4 view, proj;
  memcpy(view.m, p->view, 64);
  memcpy(proj.m, p->proj, 64);
  double invvp[16], cam[3];
  if (sizeof(T) == 4) {
    Mat4 vp = mat4_mul(proj, view);
    Mat4 inv = mat4_inverse(vp);
    Mat4 iv = mat4_inverse(view);
    for (int i = 0; i < 16; i++) invvp[i] = inv.m[i];
    for (int i = 0; i < 3; i++) cam[i] = iv.m[12 + i];  // invV * (0,0,0,1)
  } else {
    Mat4d vd, pd;
    for (int i = 0; i < 16; i++) {
      vd.m[i] = p->view[i];
      pd.m[i] = p->proj[i];
    }
    Mat4d inv = mat4d_inverse(mat4d_mul(pd, vd));
    Mat4d iv = mat4d_inverse(vd);
    for (int i = 0; i < 16; i++) invvp[i] = inv.m[i];
    for (int i = 0; i < 3; i++) cam[i] = iv.m[12 + i];
  }
  for (int i = 0; i < 16; i++) a.inv_vp[i] = T(invvp[i]);
  for (int i = 0; i < 3; i++) a.cam[i] = T(cam[i]);
  // v(s) = invVP * (2s-1, 1, 1) = c0*(2sx-1) + c1*(2sy-1) + c2 + c3;
  // ray(s) ~ sign(v.w) * (v.xyz - cam * v.w), affine in s.
  double base[4], c0[4], c1[4];
  for (int r = 0; r < 4; r++) {
    base[r] = invvp[8 + r] + invvp[12 + r] - invvp[0 + r] - invvp[4 + r];
    c0[r] = 2.0 * invvp[0 + r];
    c1[r] = 2.0 * invvp[4 + r];
  }
  double wc = base[3] + 0.5 * c0[3] + 0.5 * c1[3];
  double sg = wc < 0 ? -1.0 : 1.0;
  for (int i = 0; i < 3; i++) {
    a.e0[i] = T(sg * (base[i] - cam[i] * base[3]));
    a.ex[i] = T(sg * (c0[i] - cam[i] * c0[3]));
    a.ey[i] = T(sg * (c1[i] - cam[i] * c1[3]));
  }
}

template <class T, class A, int SAMPLER, int RFMT, int NFMT>
static hipError_t launch_gather(const GatherArgs<T> &a, bool count, hipStream_t s) {
  dim3 grid((unsigned)a.n_blocks), block(256);
  // the FP32 Hi-Z gather has every variant; the others run the default one
  constexpr bool kAll = sizeof(T) == 4 && sizeof(A) == 4 && SAMPLER == LGX_SAMPLER_HIZ_POINT;
  constexpr int O8 = kAll ? 8 : 1;
  if (count)
    hipLaunchKernelGGL((gather_kernel<T, A, SAMPLER, RFMT, NFMT, true, 1>), grid, block, 0, s, a);
  else if (kAll && a.swz) {  // (make_args sets swz only for this gather)
    if (a.occ == 8 && a.pipe)
      hipLaunchKernelGGL((gather_kernel<T, A, SAMPLER, RFMT, NFMT, false, O8, kAll, kAll>), grid, block, 0, s, a);
    else if (a.occ == 8)
      hipLaunchKernelGGL((gather_kernel<T, A, SAMPLER, RFMT, NFMT, false, O8, false, kAll>), grid, block, 0, s, a);
    else if (a.pipe)
      hipLaunchKernelGGL((gather_kernel<T, A, SAMPLER, RFMT, NFMT, false, 1, kAll, kAll>), grid, block, 0, s, a);
    else
      hipLaunchKernelGGL((gather_kernel<T, A, SAMPLER, RFMT, NFMT, false, 1, false, kAll>), grid, block, 0, s, a);
  } else if (kAll && a.occ == 8 && a.pipe)
    hipLaunchKernelGGL((gather_kernel<T, A, SAMPLER, RFMT, NFMT, false, O8, kAll>), grid, block, 0, s, a);
  else if (kAll && a.occ == 8)
    hipLaunchKernelGGL((gather_kernel<T, A, SAMPLER, RFMT, NFMT, false, O8>), grid, block, 0, s, a);
  else if (kAll && a.pipe)
    hipLaunchKernelGGL((gather_kernel<T, A, SAMPLER, RFMT, NFMT, false, 1, kAll>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((gather_kernel<T, A, SAMPLER, RFMT, NFMT, false, 1>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <class T, class A, int SAMPLER>
static hipError_t dispatch_fmt(const GatherArgs<T> &a, int rfmt, int nfmt, bool count, hipStream_t s) {
  if (rfmt == LGX_RGBA16F && nfmt == LGX_RG16F) return launch_gather<T, A, SAMPLER, LGX_RGBA16F, LGX_RG16F>(a, count, s);
  if (rfmt == LGX_RGBA32F && nfmt == LGX_RG32F) return launch_gather<T, A, SAMPLER, LGX_RGBA32F, LGX_RG32F>(a, count, s);
  if (rfmt == LGX_RGBA16F && nfmt == LGX_RG32F) return launch_gather<T, A, SAMPLER, LGX_RGBA16F, LGX_RG32F>(a, count, s);
  if (rfmt == LGX_RGBA32F && nfmt == LGX_RG16F) return launch_gather<T, A, SAMPLER, LGX_RGBA32F, LGX_RG16F>(a, count, s);
  return hipErrorInvalidValue;
}

// Launch shapes offered to the autotuner: tile_w x tile_h slice pixels per
// 256-thread workgroup, the occupancy the registers are allocated for (1: the
// compiler's choice, 6 waves per SIMD; 8: 64 VGPRs with a few spilled set-up
// values), the step loop (1: fetches overlapped with arithmetic), and the
// pyramid layout (1: the fine levels read from column-swizzled copies made in
// the workspace at each call; single-view FP32 Hi-Z gathers on a power-of-two
// pattern, otherwise as 0).  Every shape computes the same arithmetic in the
// same order on the same texels.
static constexpr int kTiles[][5] = {
    {16, 16, 1, 0, 0}, {8, 32, 1, 0, 0}, {32, 8, 1, 0, 0}, {64, 4, 1, 0, 0}, {16, 16, 8, 0, 0}, {8, 32, 8, 0, 0},
    {32, 8, 8, 0, 0},  {64, 4, 8, 0, 0}, {16, 16, 1, 1, 0}, {8, 32, 1, 1, 0}, {32, 8, 1, 1, 0}, {64, 4, 1, 1, 0},
    {16, 16, 8, 1, 0}, {8, 32, 8, 1, 0}, {32, 8, 8, 1, 0},  {64, 4, 8, 1, 0}, {16, 16, 1, 0, 1}, {8, 32, 1, 0, 1},
    {32, 8, 1, 0, 1},  {64, 4, 1, 0, 1}, {16, 16, 8, 0, 1}, {8, 32, 8, 0, 1}, {32, 8, 8, 0, 1}, {64, 4, 8, 0, 1},
    {16, 16, 1, 1, 1}, {8, 32, 1, 1, 1}, {32, 8, 1, 1, 1}, {64, 4, 1, 1, 1}, {16, 16, 8, 1, 1}, {8, 32, 8, 1, 1},
    {32, 8, 8, 1, 1},  {64, 4, 8, 1, 1}};
static constexpr int kNumTiles = 32;
static constexpr int kSwzTiles = 16;  // shapes kSwzTiles + t are shape t reading swizzled levels
constexpr bool swz_twins_ok() {
  for (int t = 0; t < kSwzTiles; t++) {
    for (int i = 0; i < 4; i++)
      if (kTiles[t][i] != kTiles[t + kSwzTiles][i]) return false;
    if (kTiles[t][4] || !kTiles[t + kSwzTiles][4]) return false;
  }
  return kNumTiles == 2 * kSwzTiles && sizeof(kTiles) / sizeof(kTiles[0]) == kNumTiles;
}
static_assert(swz_twins_ok(), "launch shape table: the second half swizzles the first");
int gather_unswizzled_tile(int tile) { return tile >= kSwzTiles && tile < kNumTiles ? tile - kSwzTiles : tile; }

// Levels stored column-swizzled: those with pattern columns / 2^l >= 2.
static int swz_levels(const lgx_gather_params *p) {
  const int g = p->grid;
  if (g < 2 || (g & (g - 1))) return 0;
  int n = 0;
  while ((g >> n) > 1 && (p->width >> n) >= 1 && (p->height >> n) >= 1) n++;
  return std::min(n, LGX_MAX_LEVELS);
}
static bool uses_swizzle(const lgx_gather_params *p) {
  return p->tile >= 0 && p->tile < kNumTiles && kTiles[p->tile][4] && p->precision == LGX_PREC_F32 &&
         p->sampler == LGX_SAMPLER_HIZ_POINT && swz_levels(p) > 0;
}
// The pyramid levels copied column-swizzled by a call with these parameters
// are laid out from offset 0 of its copy area, each 256-byte aligned, with
// room for RGBA32F radiance; returns the area's size.
struct SwzLevel {
  int f, shift, sw, pitch, w, h;
  size_t off_d, off_r;  // byte offsets in the workspace
};
static size_t swz_layout(const lgx_gather_params *p, SwzLevel *lv, int *n) {
  const int nl = swz_levels(p);
  size_t off = 0;
  for (int l = 0; l < nl; l++) {
    SwzLevel &L = lv[l];
    L.f = p->grid >> l;
    L.shift = 0;
    while ((1 << L.shift) < L.f) L.shift++;
    L.w = p->width >> l;  // floor halving, as the pyramid's levels
    L.h = p->height >> l;
    L.sw = (L.w + L.f - 1) / L.f;
    L.pitch = L.f * L.sw;
    L.off_d = off;
    off += ((size_t)L.pitch * L.h * 4 + 255) / 256 * 256;
    L.off_r = off;
    off += ((size_t)L.pitch * L.h * 16 + 255) / 256 * 256;
  }
  *n = nl;
  return off;
}

}  // namespace lgx

using namespace lgx;

static size_t table_bytes(const lgx_gather_params *p) {
  GatherTableShape s = table_shape(p);
  return ((size_t)s.g2 * s.D * 4 + (size_t)s.g2 * s.S_tab * 2) * s.elem +
         (size_t)s.g2 * (LGX_MAX_LEVELS + 1) * sizeof(int32_t);
}
// single view: the copy area follows the tables in the workspace
static size_t swz_offset(const lgx_gather_params *p) { return (table_bytes(p) + 255) / 256 * 256; }

extern "C" size_t lgx_gather_workspace_bytes(const lgx_gather_params *p) {
  if (!gather_params_ok(p)) return 0;
  if (!uses_swizzle(p)) return table_bytes(p);
  SwzLevel lv[LGX_MAX_LEVELS];
  int n = 0;
  return swz_offset(p) + swz_layout(p, lv, &n);
}

extern "C" int lgx_gather_prepare(const lgx_gather_params *p, void *workspace) {
  if (!gather_params_ok(p) || !workspace) return LGX_EINVAL;
  GatherTableShape s = table_shape(p);
  if (s.S_tab > LGX_MAX_STEPS) return LGX_EUNSUPPORTED;
  std::vector<unsigned char> t;
  bool ok = p->precision == LGX_PREC_F64 ? fill_tables<double>(p, s, t) : fill_tables<float>(p, s, t);
  if (!ok) return LGX_EUNSUPPORTED;
  const int rc = hip_status(hipMemcpy(workspace, t.data(), t.size(), hipMemcpyHostToDevice));
  if (rc == LGX_OK) record_workspace(workspace, p);
  return rc;
}

extern "C" int lgx_gather_tiles(void) { return kNumTiles; }
extern "C" int lgx_gather_tile_info(int32_t tile, int32_t *info) {
  if (tile < 0 || tile >= kNumTiles || !info) return LGX_EINVAL;
  for (int i = 0; i < 5; i++) info[i] = kTiles[tile][i];
  return LGX_OK;
}

// jobs: the swizzle copies a swizzling launch needs first, into swz_area
// (swz_layout's size; null: never swizzle)
template <class T>
static int make_args(const lgx_gather_params *p, const void *ws, const lgx_gbuffer *g, const lgx_pyramid *pyr,
                     lgx_image *out, GatherArgs<T> &a, SwzJobs *jobs = nullptr, void *swz_area = nullptr) {
  GatherTableShape s = table_shape(p);
  memset(&a, 0, sizeof(a));
  frame_constants<T>(p, a);
  const T pi = T(3.1415f);
  a.W = T(p->width);
  a.H = T(p->height);
  a.near_step = a.W / T(1000);
  a.log_ratio = std::log(T(2) * pi / T(footprint(p)) + T(1));
  a.ambient = T(0.01f);
  a.w = p->width;
  a.h = p->height;
  a.grid = p->grid;
  a.grid_y = p->grid_y > 0 ? p->grid_y : p->grid;
  a.cells_x = (p->width + a.grid - 1) / a.grid;
  a.cells_y = (p->height + a.grid_y - 1) / a.grid_y;
  a.D = p->dirs;
  a.d_pow2 = (a.D & (a.D - 1)) == 0;
  a.inv_d = 1.0 / a.D;
  a.S_tab = s.S_tab;
  a.dirs = (const T *)ws;
  a.steps = (const T *)ws + (size_t)s.g2 * s.D * 4;
  a.level_start = (const int *)(a.steps + (size_t)s.g2 * s.S_tab * 2);
  a.rW = T(1) / a.W;
  a.rH = T(1) / a.H;
  a.depth = (const float *)g->depth.ptr;
  a.depth_pitch = g->depth.pitch;
  a.normal = g->normal.ptr;
  a.normal_pitch = g->normal.pitch;
  a.n_levels = pyr->n_levels;
  for (int l = 0; l < pyr->n_levels; l++) {
    a.pd[l] = (const float *)pyr->depth[l].ptr;
    a.pd_pitch[l] = pyr->depth[l].pitch;
    a.pr[l] = pyr->radiance[l].ptr;
    a.pr_pitch[l] = pyr->radiance[l].pitch;
    a.lw[l] = pyr->depth[l].w;
    a.lh[l] = pyr->depth[l].h;
  }
  a.out = out->ptr;
  a.out_pitch = out->pitch;
  a.out_slices = p->out_layout == LGX_OUT_SLICES;
  a.step_counts = p->step_counts;
  a.hit_counts = p->step_counts ? p->hit_counts : nullptr;
  int tile = p->tile;
  if (tile < 0 || tile >= kNumTiles) return LGX_EUNSUPPORTED;
  a.tile_w = kTiles[tile][0];
  a.tile_h = kTiles[tile][1];
  a.occ = kTiles[tile][2];
  a.pipe = kTiles[tile][3];
  a.rx0 = 0;
  a.ry0 = 0;
  a.rx1 = p->width;
  a.ry1 = p->height;
  if (p->roi[2] > 0) {
    a.rx0 = p->roi[0];
    a.ry0 = p->roi[1];
    a.rx1 = p->roi[0] + p->roi[2];
    a.ry1 = p->roi[1] + p->roi[3];
  }
  if (jobs) jobs->n = 0;
  if (jobs && swz_area && uses_swizzle(p) && !p->step_counts) {
    SwzLevel lv[LGX_MAX_LEVELS];
    int n = 0;
    swz_layout(p, lv, &n);
    n = std::min(n, pyr->n_levels);
    const int rb = pyr->radiance[0].format == LGX_RGBA16F ? 8 : 16;
    a.swz = 1;
    for (int l = 0; l < n; l++) {
      const SwzLevel &L = lv[l];
      void *dd = (char *)swz_area + L.off_d, *dr = (char *)swz_area + L.off_r;
      jobs->j[jobs->n++] = SwzJob{pyr->depth[l].ptr, dd, pyr->depth[l].pitch, L.pitch, L.w, L.h, L.f - 1, L.shift,
                                  L.sw, 4};
      jobs->j[jobs->n++] = SwzJob{pyr->radiance[l].ptr, dr, pyr->radiance[l].pitch, L.pitch, L.w, L.h, L.f - 1,
                                  L.shift, L.sw, rb};
      a.pd[l] = (const float *)dd;
      a.pr[l] = dr;
      a.pd_pitch[l] = a.pr_pitch[l] = L.pitch;
      a.swz_mask[l] = L.f - 1;
      a.swz_shift[l] = L.shift;
      a.swz_w[l] = L.sw;
    }
  }
  a.sx0 = a.rx0 / a.grid;
  a.sy0 = a.ry0 / a.grid_y;
  a.lcells_x = (a.rx1 + a.grid - 1) / a.grid - a.sx0;
  a.lcells_y = (a.ry1 + a.grid_y - 1) / a.grid_y - a.sy0;
  a.tiles_x = (a.lcells_x + a.tile_w - 1) / a.tile_w;
  int tiles_y = (a.lcells_y + a.tile_h - 1) / a.tile_h;
  a.n_blocks = a.tiles_x * tiles_y * s.g2;
  return LGX_OK;
}

template <class T, class A>
static int gather_impl(const lgx_gather_params *p, void *ws, const lgx_gbuffer *g, const lgx_pyramid *pyr,
                       lgx_image *out, hipStream_t stream) {
  GatherArgs<T> a;
  SwzJobs jobs;
  int rc = make_args<T>(p, ws, g, pyr, out, a, &jobs, (char *)ws + swz_offset(p));
  if (rc != LGX_OK) return rc;
  hipError_t e;
  if (jobs.n) {
    int rows = 1;
    for (int i = 0; i < jobs.n; i++) rows = std::max(rows, jobs.j[i].h);
    const unsigned bx = (unsigned)std::min(rows, 4096);
    hipLaunchKernelGGL(swizzle_levels_kernel, dim3(bx, (unsigned)jobs.n), dim3(256), 0, stream, jobs);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_status(e);
  }
  const int rf = pyr->radiance[0].format, nf = g->normal.format;
  if (p->sampler == LGX_SAMPLER_HIZ_POINT)
    e = dispatch_fmt<T, A, LGX_SAMPLER_HIZ_POINT>(a, rf, nf, p->step_counts != nullptr, stream);
  else
    e = dispatch_fmt<T, A, LGX_SAMPLER_TRILINEAR>(a, rf, nf, p->step_counts != nullptr, stream);
  return hip_status(e);
}

// The argument checks of lgx_gather, apart from the workspace.
static int check_gather_images(const lgx_gather_params *p, const lgx_gbuffer *g, const lgx_pyramid *pyr,
                               const lgx_image *out) {
  if (!gather_params_ok(p) || !g || !pyr || !out) return LGX_EINVAL;
  if (!image_ok(&g->depth) || !image_ok(&g->normal) || !image_ok(out)) return LGX_EINVAL;
  if (g->depth.format != LGX_R32F) return LGX_EFORMAT;
  if (g->normal.format != LGX_RG16F && g->normal.format != LGX_RG32F) return LGX_EFORMAT;
  if (g->depth.w != p->width || g->depth.h != p->height || g->normal.w != p->width || g->normal.h != p->height)
    return LGX_ESIZE;
  if (out->w != p->width || out->h != p->height) return LGX_ESIZE;
  if (out->format != (p->precision == LGX_PREC_F32 ? LGX_RGBA32F : LGX_RGBA64F)) return LGX_EFORMAT;
  if (p->out_layout == LGX_OUT_SLICES &&
      (p->width % p->grid || p->height % (p->grid_y > 0 ? p->grid_y : p->grid)))
    return LGX_ESIZE;
  if (p->roi[2] > 0 && (p->roi[0] < 0 || p->roi[1] < 0 || p->roi[3] <= 0 || p->roi[0] + p->roi[2] > p->width ||
                        p->roi[1] + p->roi[3] > p->height))
    return LGX_ESIZE;
  if (pyr->n_levels < 1 || pyr->n_levels > LGX_MAX_LEVELS) return LGX_EUNSUPPORTED;
  const int rf = pyr->radiance[0].format;
  if (rf != LGX_RGBA16F && rf != LGX_RGBA32F) return LGX_EFORMAT;
  int ew = p->width, eh = p->height;
  for (int l = 0; l < pyr->n_levels; l++) {
    const lgx_image &d = pyr->depth[l], &r = pyr->radiance[l];
    if (!image_ok(&d) || !image_ok(&r)) return LGX_EINVAL;
    if (d.format != LGX_R32F || r.format != rf) return LGX_EFORMAT;
    if (d.w != ew || d.h != eh || r.w != ew || r.h != eh) return LGX_ESIZE;  // floor halving (MipBuilder.h:148-152)
    ew /= 2;
    eh /= 2;
  }
  return LGX_OK;
}

// The argument checks of lgx_gather.
static int check_gather(const lgx_gather_params *p, const void *workspace, const lgx_gbuffer *g,
                        const lgx_pyramid *pyr, const lgx_image *out) {
  if (!gather_params_ok(p) || !g || !pyr || !out) return LGX_EINVAL;
  if (!workspace) return LGX_EWORKSPACE;
  const int rc = check_gather_images(p, g, pyr, out);
  if (rc != LGX_OK) return rc;
  return workspace_prepared(workspace, p) ? LGX_OK : LGX_EWORKSPACE;
}

extern "C" int lgx_gather(const lgx_gather_params *p, void *workspace, const lgx_gbuffer *g,
                          const lgx_pyramid *pyr, lgx_image *out, lgx_stream_t stream) {
  const int rc = check_gather(p, workspace, g, pyr, out);
  if (rc != LGX_OK) return rc;
  if (p->precision == LGX_PREC_F64) return gather_impl<double, double>(p, workspace, g, pyr, out, (hipStream_t)stream);
  if (p->precision == LGX_PREC_F32_ACC64)
    return gather_impl<float, double>(p, workspace, g, pyr, out, (hipStream_t)stream);
  return gather_impl<float, float>(p, workspace, g, pyr, out, (hipStream_t)stream);
}

// ---------------------------------------------------------------- several views, one persistent launch

struct lgx_gather_batch {
  std::vector<lgx_gather_params> p;
  std::vector<const void *> ws;
  std::vector<lgx_gbuffer> g;
  std::vector<lgx_pyramid> pyr;
  std::vector<lgx_image> out;
  int sampler = 0, rf = 0, nf = 0, tile = 0;
  uint32_t total = 0;
  void *dev = nullptr;  // GatherArgs<float>[n], uint32 first[n + 1], uint32 next[8] (work counters)
  // swizzling shapes: the views' swizzle jobs, then their copy areas
  // (allocated by the first set_tile to such a shape)
  void *swz = nullptr;
  size_t swz_bytes = 0;
  int njobs = 0, job_rows = 0;
  ~lgx_gather_batch() {
    if (dev) (void)hipFree(dev);
    if (swz) (void)hipFree(swz);
  }
};

namespace lgx {

static int gather_batch_upload(lgx_gather_batch *b, int tile) {
  const int n = (int)b->p.size();
  std::vector<GatherArgs<float>> args((size_t)n);
  std::vector<uint32_t> first((size_t)n + 1, 0);
  // copy areas of the views (one each: views may share a workspace), after
  // room for every view's jobs
  std::vector<size_t> area((size_t)n, 0);
  size_t off = ((size_t)n * 2 * LGX_MAX_LEVELS * sizeof(SwzJob) + 255) / 256 * 256;
  bool swz = false;
  for (int i = 0; i < n; i++) {
    lgx_gather_params q = b->p[i];
    q.tile = tile;
    if (!uses_swizzle(&q)) continue;
    SwzLevel lv[LGX_MAX_LEVELS];
    int nl = 0;
    area[i] = off;
    off += swz_layout(&q, lv, &nl);
    swz = true;
  }
  // no launch of this batch may still read the previous arguments or copies
  int rc0 = hip_status(hipDeviceSynchronize());
  if (rc0 != LGX_OK) return rc0;
  if (swz && off > b->swz_bytes) {  // the same for every swizzling shape: allocated once
    if (b->swz) (void)hipFree(b->swz);
    b->swz_bytes = 0;
    rc0 = hip_status(hipMalloc(&b->swz, off));
    if (rc0 != LGX_OK) {
      b->swz = nullptr;
      return rc0;
    }
    b->swz_bytes = off;
  }
  std::vector<SwzJob> jobs;
  int rows = 1;
  for (int i = 0; i < n; i++) {
    lgx_gather_params q = b->p[i];
    q.tile = tile;
    SwzJobs vj;
    int rc = make_args<float>(&q, b->ws[i], &b->g[i], &b->pyr[i], &b->out[i], args[i], swz ? &vj : nullptr,
                              swz ? (char *)b->swz + area[i] : nullptr);
    if (rc != LGX_OK) return rc;
    if (swz)
      for (int k = 0; k < vj.n; k++) {
        jobs.push_back(vj.j[k]);
        rows = std::max(rows, vj.j[k].h);
      }
    first[i + 1] = first[i] + (uint32_t)args[i].n_blocks;
  }
  if (!jobs.empty()) {
    rc0 = hip_status(hipMemcpy(b->swz, jobs.data(), jobs.size() * sizeof(SwzJob), hipMemcpyHostToDevice));
    if (rc0 != LGX_OK) return rc0;
  }
  const size_t ab = args.size() * sizeof(GatherArgs<float>);
  std::vector<unsigned char> blob(ab + first.size() * sizeof(uint32_t));
  memcpy(blob.data(), args.data(), ab);
  memcpy(blob.data() + ab, first.data(), first.size() * sizeof(uint32_t));
  int rc = hip_status(hipMemcpy(b->dev, blob.data(), blob.size(), hipMemcpyHostToDevice));
  if (rc != LGX_OK) return rc;
  b->tile = tile;
  b->total = first[n];
  b->njobs = (int)jobs.size();
  b->job_rows = std::min(rows, 4096);
  return LGX_OK;
}

template <int SAMPLER, int RFMT, int NFMT>
static hipError_t launch_multi(const lgx_gather_batch *b, unsigned grid, hipStream_t s) {
  const bool occ8 = kTiles[b->tile][2] == 8, pipe = kTiles[b->tile][3] == 1;
  const int n = (int)b->p.size();
  const GatherArgs<float> *args = (const GatherArgs<float> *)b->dev;
  const uint32_t *first = (const uint32_t *)((const unsigned char *)b->dev + (size_t)n * sizeof(GatherArgs<float>));
  uint32_t *next = (uint32_t *)(first + n + 1);
  // a kernel, not a memset, so the launch can be captured into a graph
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, next, 8);
  // the Hi-Z sampler has the swizzled variants (make_args sets no jobs otherwise)
  constexpr bool kSwz = SAMPLER == LGX_SAMPLER_HIZ_POINT;
  if (kSwz && b->njobs) {
    hipLaunchKernelGGL(swizzle_jobs_kernel, dim3((unsigned)b->job_rows, (unsigned)b->njobs), dim3(256), 0, s,
                       (const SwzJob *)b->swz);
    if (occ8 && pipe)
      hipLaunchKernelGGL((gather_multi_kernel<float, float, SAMPLER, RFMT, NFMT, 8, true, kSwz>), dim3(grid), dim3(256),
                         0, s, args, first, n, next);
    else if (occ8)
      hipLaunchKernelGGL((gather_multi_kernel<float, float, SAMPLER, RFMT, NFMT, 8, false, kSwz>), dim3(grid),
                         dim3(256), 0, s, args, first, n, next);
    else if (pipe)
      hipLaunchKernelGGL((gather_multi_kernel<float, float, SAMPLER, RFMT, NFMT, 1, true, kSwz>), dim3(grid), dim3(256),
                         0, s, args, first, n, next);
    else
      hipLaunchKernelGGL((gather_multi_kernel<float, float, SAMPLER, RFMT, NFMT, 1, false, kSwz>), dim3(grid),
                         dim3(256), 0, s, args, first, n, next);
  } else if (occ8 && pipe)
    hipLaunchKernelGGL((gather_multi_kernel<float, float, SAMPLER, RFMT, NFMT, 8, true, false>), dim3(grid), dim3(256),
                       0, s, args, first, n, next);
  else if (occ8)
    hipLaunchKernelGGL((gather_multi_kernel<float, float, SAMPLER, RFMT, NFMT, 8, false, false>), dim3(grid),
                       dim3(256), 0, s, args, first, n, next);
  else if (pipe)
    hipLaunchKernelGGL((gather_multi_kernel<float, float, SAMPLER, RFMT, NFMT, 1, true, false>), dim3(grid), dim3(256),
                       0, s, args, first, n, next);
  else
    hipLaunchKernelGGL((gather_multi_kernel<float, float, SAMPLER, RFMT, NFMT, 1, false, false>), dim3(grid),
                       dim3(256), 0, s, args, first, n, next);
  return hipGetLastError();
}

}  // namespace lgx

extern "C" int lgx_gather_batch_create(const lgx_gather_params *p, const void *const *workspaces,
                                       const lgx_gbuffer *g, const lgx_pyramid *pyr, lgx_image *out, int32_t n,
                                       lgx_gather_batch **res) {
  if (!p || !workspaces || !g || !pyr || !out || !res || n < 1) return LGX_EINVAL;
  *res = nullptr;
  if (p[0].tile < 0 || p[0].tile >= kNumTiles) return LGX_EUNSUPPORTED;
  if ((size_t)n * 2 * LGX_MAX_LEVELS > 65535) return LGX_EUNSUPPORTED;  // swizzle jobs on the grid's y
  for (int i = 0; i < n; i++) {
    // the batch holds its views' swizzled copies itself: the workspaces
    // need the tables only
    lgx_gather_params q = p[i];
    q.tile = gather_unswizzled_tile(q.tile);
    int rc = check_gather(&q, workspaces[i], &g[i], &pyr[i], &out[i]);
    if (rc != LGX_OK) return rc;
    if (p[i].precision != LGX_PREC_F32 || p[i].step_counts || p[i].hit_counts) return LGX_EUNSUPPORTED;
    if (p[i].sampler != p[0].sampler || p[i].tile != p[0].tile ||
        pyr[i].radiance[0].format != pyr[0].radiance[0].format || g[i].normal.format != g[0].normal.format)
      return LGX_EINVAL;
  }
  const int rf = pyr[0].radiance[0].format, nf = g[0].normal.format;
  if (!((rf == LGX_RGBA16F && nf == LGX_RG16F) || (rf == LGX_RGBA32F && nf == LGX_RG32F))) return LGX_EFORMAT;
  lgx_gather_batch *b = new (std::nothrow) lgx_gather_batch;
  if (!b) return LGX_EINVAL;
  b->p.assign(p, p + n);
  b->ws.assign(workspaces, workspaces + n);
  b->g.assign(g, g + n);
  b->pyr.assign(pyr, pyr + n);
  b->out.assign(out, out + n);
  b->sampler = p[0].sampler;
  b->rf = rf;
  b->nf = nf;
  int rc = hip_status(
      hipMalloc(&b->dev, (size_t)n * sizeof(GatherArgs<float>) + ((size_t)n + 1 + 8) * sizeof(uint32_t)));
  if (rc == LGX_OK) rc = gather_batch_upload(b, p[0].tile);
  if (rc != LGX_OK) {
    delete b;
    return rc;
  }
  *res = b;
  return LGX_OK;
}

extern "C" int lgx_gather_batch_set_tile(lgx_gather_batch *b, int32_t tile) {
  if (!b) return LGX_EINVAL;
  if (tile < 0 || tile >= kNumTiles) return LGX_EUNSUPPORTED;
  if (tile == b->tile) return LGX_OK;
  return gather_batch_upload(b, tile);
}

extern "C" int64_t lgx_gather_batch_blocks(const lgx_gather_batch *b) { return b ? (int64_t)b->total : -1; }

extern "C" int lgx_gather_batch_launch(const lgx_gather_batch *b, int32_t workgroups, lgx_stream_t stream) {
  if (!b || workgroups < 0) return LGX_EINVAL;
  uint32_t grid = workgroups > 0 ? (uint32_t)workgroups : b->total;
  grid = (grid + 7) / 8 * 8;  // the XCD chunking needs a multiple of 8
  if (grid == 0 || grid > (1u << 24)) return LGX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (b->sampler == LGX_SAMPLER_HIZ_POINT)
    e = b->rf == LGX_RGBA16F ? launch_multi<LGX_SAMPLER_HIZ_POINT, LGX_RGBA16F, LGX_RG16F>(b, grid, s)
                             : launch_multi<LGX_SAMPLER_HIZ_POINT, LGX_RGBA32F, LGX_RG32F>(b, grid, s);
  else
    e = b->rf == LGX_RGBA16F ? launch_multi<LGX_SAMPLER_TRILINEAR, LGX_RGBA16F, LGX_RG16F>(b, grid, s)
                             : launch_multi<LGX_SAMPLER_TRILINEAR, LGX_RGBA32F, LGX_RG32F>(b, grid, s);
  return hip_status(e);
}

extern "C" void lgx_gather_batch_destroy(lgx_gather_batch *b) { delete b; }

// ---------------------------------------------------------------- the gather without allocations
//
// lgx_gather_noalloc is lgx_gather with its workspace (the direction, step
// and level tables) held in `scratch`, a file-scope device array, instead of
// caller memory.  The array is read through the address hipGetSymbolAddress
// returns for the current device (its host-side name is not a device
// address), once per device on the first call there.  The tables are
// uploaded when the call's table parameters differ from the last upload on
// that device; a call whose tables are already resident launches only the
// gather kernel.  Every call on a device shares its array.
static __device__ __align__(256) unsigned char scratch[LGX_GATHER_SCRATCH_BYTES];

namespace lgx {

struct ScratchSlot {
  void *ptr = nullptr;
  bool loaded = false;
  TableKey key{};
};
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static ScratchSlot g_scratch[kMaxDevices];

}  // namespace lgx

extern "C" int lgx_gather_noalloc(const lgx_gather_params *p, const lgx_gbuffer *g, const lgx_pyramid *pyr,
                                  lgx_image *out, lgx_stream_t stream) {
  const int rc = check_gather_images(p, g, pyr, out);
  if (rc != LGX_OK) return rc;
  if (uses_swizzle(p)) return LGX_EUNSUPPORTED;  // the scratch array holds the tables only
  const GatherTableShape s = table_shape(p);
  if (s.S_tab > LGX_MAX_STEPS) return LGX_EUNSUPPORTED;
  if (lgx_gather_workspace_bytes(p) > (size_t)LGX_GATHER_SCRATCH_BYTES) return LGX_EUNSUPPORTED;
  int dev = 0;
  int hrc = hip_status(hipGetDevice(&dev));
  if (hrc != LGX_OK) return hrc;
  if (dev < 0 || dev >= kMaxDevices) return LGX_EUNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  ScratchSlot &slot = g_scratch[dev];
  if (!slot.ptr) {
    void *ptr = nullptr;
    hrc = hip_status(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (hrc != LGX_OK) return hrc;
    if (!ptr) return LGX_EINVAL;
    slot.ptr = ptr;
  }
  const TableKey key = table_key(p);
  if (!slot.loaded || !(slot.key == key)) {
    std::vector<unsigned char> t;
    const bool ok = p->precision == LGX_PREC_F64 ? fill_tables<double>(p, s, t) : fill_tables<float>(p, s, t);
    if (!ok) return LGX_EUNSUPPORTED;
    // earlier gathers on this stream read the old tables: let them finish,
    // then upload on the stream and wait, so the host buffer can go
    slot.loaded = false;
    hrc = hip_status(hipStreamSynchronize(st));
    if (hrc == LGX_OK) hrc = hip_status(hipMemcpyAsync(slot.ptr, t.data(), t.size(), hipMemcpyHostToDevice, st));
    if (hrc == LGX_OK) hrc = hip_status(hipStreamSynchronize(st));
    if (hrc != LGX_OK) return hrc;
    slot.key = key;
    slot.loaded = true;
  }
  if (p->precision == LGX_PREC_F64) return gather_impl<double, double>(p, slot.ptr, g, pyr, out, st);
  if (p->precision == LGX_PREC_F32_ACC64) return gather_impl<float, double>(p, slot.ptr, g, pyr, out, st);
  return gather_impl<float, float>(p, slot.ptr, g, pyr, out, st);
}
