/*
 * CPU oracle for the screen-space GI path — TEST INFRASTRUCTURE, not product.
 *
 * Scalar C++17 transliteration of LegitEngine's SSVGI GLSL passes, templated
 * on float / double, parallelised over rows with OpenMP.  Used only as the
 * checker of the HIP path (tests/, smoke) and as the CPU baseline (bench.py).
 *
 * Reference files followed (paths relative to the reference root):
 *   bin/data/Shaders/glsl/SSVGI/indirectLighting.frag   gather
 *   bin/data/Shaders/glsl/Common/mipLevelBuilder.frag   avg / depth mips
 *   bin/data/Shaders/glsl/Common/blurLayerBuilder.frag  box pre-blur
 *   bin/data/Shaders/glsl/Common/denoiser.frag          regression + bilateral weights
 *   bin/data/Shaders/glsl/Common/{interleave,deinterleave}.frag, interleavedRendering.decl
 *   bin/data/Shaders/glsl/Common/finalGatherer.frag     composite
 *   src/Render/Common/MipBuilder.h, BlurBuilder.h, InterleaveBuilder.h
 *   src/Render/Renderers/SSVGIRenderer.h, src/Scene/Scene.h
 *   dependencies/glm/include/glm/ext/matrix_clip_space.inl (perspectiveRH_NO)
 *
 * Pinning (what fixes each function independently of its own formula; the
 * tests live in tests/test_oracle_kat.py):
 *   hammersley            exact table k/16 for i<16 and brute-force bit reversal
 *   mat4 inverse/persp.   M^-1 M = I, Project(Unproject(s)) = s, glm closed form
 *   horizon contribution  numerical quadrature of the slice integral; full
 *                         unoccluded slice with n = eye gives 1/2
 *   f16 conversion        exhaustive round trip of all 65536 halves, ties-to-even
 *                         cases, numpy's float16 cast on random floats
 *   mip avg / max         brute force over 2^l x 2^l footprints (power-of-two
 *                         inputs), parent >= children, constant image fixed point
 *   box blur              r=0 identity, constant fixed point, brute force on 8x8
 *   (de)interleave        closed-form permutation, inverse pair = identity,
 *                         GetDPixel/GetIPixel cross-check
 *   samplers              texel-centre exactness, linear ramp reproduction,
 *                         clamp-to-edge, level selection
 *   gather                planar KAT (uniform radiance plane gives
 *                         ambient * sum_d (2/D) h(0, theta_max_d), exactly ambient
 *                         where n = eye: the no-hit branch), zero-radiance =
 *                         ambient-only, linearity of the (L - ambient) part in
 *                         radiance, float vs double; the hit branch (:254-264)
 *                         by a sphere of constant radial depth around the camera,
 *                         where every on-screen sample rises above the horizon at
 *                         theta = atan2(cos((phi - delta)/2), sin(phi/2)) in
 *                         closed form: the hit sum telescopes (ring of ambient
 *                         radiance) or is summed hit by hit with the side weights,
 *                         the slice integral by quadrature (tests/kat_scenes.py)
 *   bilateral             constant image fixed point, r=0 identity, weights
 *                         normalised (output inside the tap range); a two-region
 *                         depth/normal field whose cross-region taps weigh
 *                         exp(-10.5 |dd|) exp(-20 |dn|^2) = 0.266, every output
 *                         column a closed-form ratio of tap counts
 *   regression denoiser   affine colour-in-depth reproduced, radius 0 passthrough
 * Parity with the reference's own GPU output is unpinned: the reference runs
 * only under Vulkan, which this environment lacks (no loader, no GPU, empty
 * LegitVulkan submodule), and it ships no golden images.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <class T> struct V2 { T x, y; };
template <class T> struct V3 { T x, y, z; };
template <class T> struct V4 { T x, y, z, w; };

template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
/* GLSL normalize(v) = v * inversesqrt(dot(v, v)); written as a division by
 * the length, the form the SPIR-V GLSL.std.450 Normalize is specified by. */
template <class T> inline V3<T> normalize(V3<T> a) {
  T len = std::sqrt(dot(a, a));
  return {a.x / len, a.y / len, a.z / len};
}
template <class T> inline T saturate(T v) { return std::min(std::max(v, T(0)), T(1)); }

/* Column-major 4x4, m[col*4 + row], matching the glm uniform layout the
 * shaders receive (SSVGIRenderer.h:242-244). */
template <class T> struct M4 { T m[16]; };

template <class T> M4<T> load(const float *f) {
  M4<T> r;
  for (int i = 0; i < 16; i++) r.m[i] = T(f[i]);
  return r;
}
template <class T> V4<T> mul(const M4<T> &a, V4<T> v) {
  V4<T> r;
  r.x = a.m[0] * v.x + a.m[4] * v.y + a.m[8] * v.z + a.m[12] * v.w;
  r.y = a.m[1] * v.x + a.m[5] * v.y + a.m[9] * v.z + a.m[13] * v.w;
  r.z = a.m[2] * v.x + a.m[6] * v.y + a.m[10] * v.z + a.m[14] * v.w;
  r.w = a.m[3] * v.x + a.m[7] * v.y + a.m[11] * v.z + a.m[15] * v.w;
  return r;
}
template <class T> M4<T> mul(const M4<T> &a, const M4<T> &b) {
  M4<T> r;
  for (int c = 0; c < 4; c++)
    for (int row = 0; row < 4; row++) {
      T s = 0;
      for (int k = 0; k < 4; k++) s += a.m[k * 4 + row] * b.m[c * 4 + k];
      r.m[c * 4 + row] = s;
    }
  return r;
}
/* inverse(mat4) by the adjugate: cofactors of the 3x3 minors / determinant.
 * (GLSL inverse() at indirectLighting.frag:124-126; the driver's algorithm is
 * not specified, any exact-arithmetic-equivalent form is faithful.) */
template <class T> M4<T> inverse(const M4<T> &a) {
  auto at = [&](int r, int c) { return a.m[c * 4 + r]; };
  T cof[4][4];
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) {
      int rs[3], cs[3], ri = 0, ci = 0;
      for (int k = 0; k < 4; k++) {
        if (k != r) rs[ri++] = k;
        if (k != c) cs[ci++] = k;
      }
      T d = at(rs[0], cs[0]) * (at(rs[1], cs[1]) * at(rs[2], cs[2]) - at(rs[1], cs[2]) * at(rs[2], cs[1])) -
            at(rs[0], cs[1]) * (at(rs[1], cs[0]) * at(rs[2], cs[2]) - at(rs[1], cs[2]) * at(rs[2], cs[0])) +
            at(rs[0], cs[2]) * (at(rs[1], cs[0]) * at(rs[2], cs[1]) - at(rs[1], cs[1]) * at(rs[2], cs[0]));
      cof[r][c] = ((r + c) & 1) ? -d : d;
    }
  T det = 0;
  for (int c = 0; c < 4; c++) det += at(0, c) * cof[0][c];
  M4<T> inv;
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) inv.m[c * 4 + r] = cof[c][r] / det; /* adjugate = cofactor^T */
  return inv;
}

/* indirectLighting.frag:20-25 Unproject */
template <class T> V3<T> unproject(V3<T> s, const M4<T> &inv) {
  V4<T> v = mul(inv, V4<T>{s.x * T(2) - T(1), s.y * T(2) - T(1), s.z, T(1)});
  return {v.x / v.w, v.y / v.w, v.z / v.w};
}

/* indirectLighting.frag:44-49 / horizonContribution.decl:1-6 */
template <class T> T horizon_contribution(V3<T> eye, V3<T> tangent, V3<T> n, T mn, T mx) {
  return T(0.25) * dot(eye, n) * (-std::cos(T(2) * mx) + std::cos(T(2) * mn)) +
         T(0.25) * dot(tangent, n) * (T(2) * mx - T(2) * mn - std::sin(T(2) * mx) + std::sin(T(2) * mn));
}

/* indirectLighting.frag:83-99 BoxRayCast */
template <class T> void box_ray_cast(V2<T> start, V2<T> dir, V2<T> bmin, V2<T> bmax, T &tmin, T &tmax) {
  V2<T> inv{T(1) / dir.x, T(1) / dir.y};
  T t1 = (bmin.x - start.x) * inv.x;
  T t2 = (bmax.x - start.x) * inv.x;
  T t3 = (bmin.y - start.y) * inv.y;
  T t4 = (bmax.y - start.y) * inv.y;
  tmin = std::max(std::min(t1, t2), std::min(t3, t4));
  tmax = std::min(std::max(t1, t2), std::max(t3, t4));
}

/* indirectLighting.frag:101-112 HammersleyNorm (host twin src/Scene/Mesh.h:114-124) */
inline void hammersley(int i, int n, float &a, float &b) {
  uint32_t v = (uint32_t(i) << 16u) | (uint32_t(i) >> 16u);
  v = ((v & 0x55555555u) << 1u) | ((v & 0xAAAAAAAAu) >> 1u);
  v = ((v & 0x33333333u) << 2u) | ((v & 0xCCCCCCCCu) >> 2u);
  v = ((v & 0x0F0F0F0Fu) << 4u) | ((v & 0xF0F0F0F0u) >> 4u);
  v = ((v & 0x00FF00FFu) << 8u) | ((v & 0xFF00FF00u) >> 8u);
  a = float(i) / float(n);
  b = float(v) / float(0xffffffffu); /* uint -> float rounds 2^32-1 to 2^32 */
}

/* Octahedral normal decode (G-buffer format of the configs; the reference
 * stores a raw RGBA16F normal, gBufferBuilder.frag:35). */
template <class T> V3<T> oct_decode(T ex, T ey) {
  V3<T> n{ex, ey, T(1) - std::fabs(ex) - std::fabs(ey)};
  if (n.z < T(0)) {
    T ox = (T(1) - std::fabs(n.y)) * (n.x >= T(0) ? T(1) : T(-1));
    T oy = (T(1) - std::fabs(n.x)) * (n.y >= T(0) ? T(1) : T(-1));
    n.x = ox;
    n.y = oy;
  }
  return normalize(n);
}

struct Level {
  const float *data;
  int w, h;
};

/* Vulkan clamp-to-edge bilinear at normalized (sx, sy) in one level:
 * u = s*W - 1/2, taps floor(u), floor(u)+1 clamped, weights frac(u). */
template <class T> void bilinear(const Level &L, int c, T sx, T sy, T *out) {
  T u = sx * T(L.w) - T(0.5), v = sy * T(L.h) - T(0.5);
  T fu = std::floor(u), fv = std::floor(v);
  T a = u - fu, b = v - fv;
  long i0 = (long)fu, j0 = (long)fv;
  long i1 = i0 + 1, j1 = j0 + 1;
  auto cl = [](long k, int n) { return (int)std::min<long>(std::max<long>(k, 0), n - 1); };
  int x0 = cl(i0, L.w), x1 = cl(i1, L.w), y0 = cl(j0, L.h), y1 = cl(j1, L.h);
  const float *p00 = L.data + ((size_t)y0 * L.w + x0) * c;
  const float *p10 = L.data + ((size_t)y0 * L.w + x1) * c;
  const float *p01 = L.data + ((size_t)y1 * L.w + x0) * c;
  const float *p11 = L.data + ((size_t)y1 * L.w + x1) * c;
  for (int k = 0; k < c; k++) {
    T top = T(p00[k]) * (T(1) - a) + T(p10[k]) * a;
    T bot = T(p01[k]) * (T(1) - a) + T(p11[k]) * a;
    out[k] = top * (T(1) - b) + bot * b;
  }
}

/* textureLod with ClampToEdge / eLinear / eLinear (SSVGIRenderer.h:17):
 * lod clamped to [0, levels-1] (lod <= 0, including -inf, is magnification at
 * level 0 with the linear filter), linear blend of floor(lod) and floor(lod)+1. */
template <class T> void sample_trilinear(const Level *lv, int nl, int c, T sx, T sy, T lod, T *out) {
  T q = T(nl - 1);
  T d = lod;
  if (!(d > T(0))) d = T(0); /* also maps -inf and NaN to level 0 */
  if (d > q) d = q;
  int l0 = (int)std::floor(d);
  int l1 = std::min(l0 + 1, nl - 1);
  T delta = d - T(l0);
  T a[4], b[4];
  bilinear(lv[l0], c, sx, sy, a);
  if (delta > T(0) && l1 != l0) {
    bilinear(lv[l1], c, sx, sy, b);
    for (int k = 0; k < c; k++) out[k] = a[k] * (T(1) - delta) + b[k] * delta;
  } else {
    for (int k = 0; k < c; k++) out[k] = a[k];
  }
}

/* Hi-Z point fetch: level clamp(floor(lod), 0, levels-1), nearest texel
 * floor(s * W_l) clamped to the level (the Hi-Z point-fetch mode). */
template <class T> void sample_point(const Level *lv, int nl, int c, T sx, T sy, T lod, T *out) {
  int l = 0;
  if (lod > T(0)) l = (int)std::min<T>(std::floor(lod), T(nl - 1));
  const Level &L = lv[l];
  T u = std::floor(sx * T(L.w)), v = std::floor(sy * T(L.h));
  long i = (long)std::min<T>(std::max<T>(u, T(0)), T(L.w - 1));
  long j = (long)std::min<T>(std::max<T>(v, T(0)), T(L.h - 1));
  const float *p = L.data + ((size_t)j * L.w + i) * c;
  for (int k = 0; k < c; k++) out[k] = T(p[k]);
}

uint16_t f32_to_f16(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u | ((ax >> 13) & 0x3ffu) : 0u));
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 rounds to inf */
  if (ax < 0x38800000u) {                                     /* below 2^-14: half subnormal or zero */
    uint32_t e = ax >> 23;
    if (e < 102) return (uint16_t)sign; /* < 2^-25 rounds to 0 */
    uint32_t m = (ax & 0x7fffffu) | 0x800000u;
    uint32_t shift = 126 - e; /* 14..24 */
    uint32_t q = m >> shift, rem = m & ((1u << shift) - 1), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint16_t)(sign | q);
  }
  uint32_t r = ax + 0xfffu + ((ax >> 13) & 1u);
  return (uint16_t)(sign | ((r - 0x38000000u) >> 13));
}
float f16_to_f32(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu, x;
  if (e == 0) {
    float v = std::ldexp((float)m, -24);
    std::memcpy(&x, &v, 4);
    x |= sign;
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}
inline float round16(float v, int on) { return on ? f16_to_f32(f32_to_f16(v)) : v; }

int nthreads(int t) {
#ifdef _OPENMP
  return t > 0 ? t : omp_get_max_threads();
#else
  (void)t;
  return 1;
#endif
}

/* One pixel of indirectLighting.frag main (:114-272), in reference mode
 * (sampler 0, D from params, no cap) or the Hi-Z mode (point fetch,
 * step cap S).  The centre is reconstructed from the radial distance the
 * G-buffer stores (gBufferBuilder.frag:31,37): P = cam + ray(s) * r, the
 * same reconstruction the shader applies to every sample (:241).  The
 * offset point of :181 is taken at the same radial distance; only its
 * direction from the camera enters the tangent (:183). */
template <class T>
void gather_pixel(const orc_gather_params &p, const M4<T> &invVP, V3<T> cam, int px, int py, const float *depth,
                  const float *oct, const Level *ld, const Level *lr, T *out_rgb, int32_t *steps_out,
                  int32_t *hits_out = nullptr) {
  const T W = T(p.width), H = T(p.height);
  const int D = p.dirs;
  V2<T> c{T(px) + T(0.5), T(py) + T(0.5)}; /* gl_FragCoord.xy, :116 */
  V2<T> s{c.x / W, c.y / H};                /* :118 */
  size_t idx = (size_t)py * p.width + px;
  T rc = T(depth[idx]);
  V3<T> n = oct_decode<T>(T(oct[idx * 2]), T(oct[idx * 2 + 1])); /* :133 */
  V3<T> rayc = normalize(unproject(V3<T>{s.x, s.y, T(1)}, invVP) - cam);
  V3<T> P = cam + rayc * rc; /* :130 */

  const int g = p.grid, gy = p.grid_y > 0 ? p.grid_y : p.grid; /* square in the shader */
  int pix = px % g, piy = py % gy; /* :155 */
  int index = pix + piy * g;       /* :161 */
  float angf, linf;
  hammersley(index, g * gy, angf, linf); /* :162-164 */
  const T ang_offset = T(angf), lin_offset = T(linf);
  const T pi = T(3.1415f);             /* :169, verbatim */
  const T near_step = W / T(1000);     /* :167 */
  const T ambient = T(0.01f);          /* :154 */
  const T pixel_ang_offset = T(2) * pi / T(D) * ang_offset; /* :170 */
  /* the shader uses D for the step growth and the lod as well; a separate
     footprint count lets a few directions march with a finer footprint */
  const T Df = T(p.footprint_dirs > 0 ? p.footprint_dirs : D);
  const T ratio = T(2) * pi / Df + T(1);

  T sum[3] = {0, 0, 0};
  int32_t steps_total = 0, hits_total = 0;
  for (int d = 0; d < D; d++) { /* :174 */
    T screen_ang = pixel_ang_offset + T(2) * pi / T(D) * T(d); /* :177 */
    V2<T> dir{std::cos(screen_ang), std::sin(screen_ang)};
    const T eps = T(10e-1); /* :180 */
    V2<T> oc{(c.x + dir.x * eps) / W, (c.y + dir.y * eps) / H};
    V3<T> offset_pos = cam + normalize(unproject(V3<T>{oc.x, oc.y, T(1)}, invVP) - cam) * rc; /* :181 */
    V3<T> eye = normalize(cam - P);                                                            /* :182 */
    V3<T> tangent = normalize(normalize(offset_pos - cam) - normalize(P - cam));              /* :183 */
    T max_h;
    {
      V3<T> dn = cross(-cross(tangent, eye), n); /* :193 */
      max_h = std::atan2(dot(dn, tangent), dot(dn, eye)); /* :194-198 */
    }
    T tmin, tmax;
    box_ray_cast(c, dir, V2<T>{T(0), T(0)}, V2<T>{W, H}, tmin, tmax); /* :202 */
    T path = std::fabs(tmax);
    T L[3];
    {
      T h0 = horizon_contribution(eye, tangent, n, T(0), max_h); /* :209 */
      for (int k = 0; k < 3; k++) L[k] = ambient * h0;
    }
    T iters_f = std::log(path / near_step) / std::log(ratio); /* :212 */
    long iters = (long)iters_f + 1; /* int() truncates toward zero */
    if (p.max_steps > 0 && iters > p.max_steps) iters = p.max_steps;
    for (long j = 0; j < iters; j++) {
      T off = near_step * std::pow(ratio, T(j) + lin_offset) + T(1) - near_step; /* :217 */
      V2<T> q{c.x + dir.x * off, c.y + dir.y * off};
      V2<T> sq{q.x / W, q.y / H};
      T side = T(1);
      {
        const T width = T(0.1f), inv_width = T(1) / width; /* :222-227 */
        side *= saturate((T(1) - sq.x) * inv_width);
        side *= saturate(sq.x * inv_width);
        side *= saturate((T(1) - sq.y) * inv_width);
        side *= saturate(sq.y * inv_width);
      }
      if (side > T(0)) steps_total++;
      const T blur_lod_offset = T(-2), lod_mult = T(0.5f); /* :231-233 */
      T lod = std::log(std::max(T(0), T(2) * pi / Df * (off - T(1)) * lod_mult)) / std::log(T(2)) + blur_lod_offset;
      T ds[4];
      if (p.sampler == 0)
        sample_trilinear<T>(ld, p.n_levels, 1, sq.x, sq.y, lod, ds); /* :240 */
      else
        sample_point<T>(ld, p.n_levels, 1, sq.x, sq.y, lod, ds);
      V3<T> sp = cam + normalize(unproject(V3<T>{sq.x, sq.y, T(1)}, invVP) - cam) * ds[0]; /* :241 */
      V3<T> delta = sp - P;                                                                /* :249 */
      T sh = std::atan2(dot(tangent, delta), dot(eye, delta));                            /* :250-252 */
      if (sh < max_h) {                                                                    /* :254 */
        if (side > T(0)) hits_total++; /* an on-screen step that fetches radiance (:256) */
        T ls[4];
        if (p.sampler == 0)
          sample_trilinear<T>(lr, p.n_levels, 4, sq.x, sq.y, lod, ls); /* :256 */
        else
          sample_point<T>(lr, p.n_levels, 4, sq.x, sq.y, lod, ls);
        T hc = horizon_contribution(eye, tangent, n, sh, max_h) * side; /* :258 */
        for (int k = 0; k < 3; k++) {
          L[k] += ls[k] * hc;  /* :261 */
          L[k] -= ambient * hc; /* :262 */
        }
        max_h = sh; /* :263 */
      }
    }
    for (int k = 0; k < 3; k++) sum[k] += T(2) * L[k] / T(D); /* :268 */
  }
  for (int k = 0; k < 3; k++) out_rgb[k] = sum[k];
  if (steps_out) *steps_out = steps_total;
  if (hits_out) *hits_out = hits_total;
}

template <class T, class O>
int gather(const orc_gather_params *p, const float *depth, const float *oct, const float *const *pd,
           const float *const *pr, const int32_t *lw, const int32_t *lh, O *out, int32_t *steps, int x0, int y0,
           int cw, int ch, int threads) {
  if (!p || !depth || !oct || !pd || !pr || !out || p->n_levels < 1 || p->n_levels > 16 || p->dirs < 1 ||
      p->grid < 1)
    return -1;
  if (cw <= 0) cw = p->width;
  if (ch <= 0) ch = p->height;
  if (x0 < 0 || y0 < 0 || x0 + cw > p->width || y0 + ch > p->height) return -2;
  Level ld[16], lr[16];
  for (int l = 0; l < p->n_levels; l++) {
    ld[l] = Level{pd[l], lw[l], lh[l]};
    lr[l] = Level{pr[l], lw[l], lh[l]};
  }
  M4<T> view = load<T>(p->view), proj = load<T>(p->proj);
  M4<T> vp = mul(proj, view);                    /* :123 */
  M4<T> invVP = inverse(vp);                     /* :124 */
  M4<T> invV = inverse(view);                    /* :125 */
  V4<T> cam4 = mul(invV, V4<T>{0, 0, 0, 1});     /* :128 */
  V3<T> cam{cam4.x, cam4.y, cam4.z};
#pragma omp parallel for schedule(dynamic) num_threads(nthreads(threads))
  for (int yy = 0; yy < ch; yy++) {
    for (int xx = 0; xx < cw; xx++) {
      T rgb[3];
      int32_t st = 0;
      gather_pixel<T>(*p, invVP, cam, x0 + xx, y0 + yy, depth, oct, ld, lr, rgb, &st);
      size_t o = (size_t)yy * cw + xx;
      out[o * 4 + 0] = O(rgb[0]);
      out[o * 4 + 1] = O(rgb[1]);
      out[o * 4 + 2] = O(rgb[2]);
      out[o * 4 + 3] = O(1); /* :270 */
      if (steps) steps[o] = st;
    }
  }
  return 0;
}

/* The gather on a list of whole rows (stratified samples of a frame);
 * out holds n_rows x width RGBA in list order. */
template <class T, class O>
int gather_rows(const orc_gather_params *p, const float *depth, const float *oct, const float *const *pd,
                const float *const *pr, const int32_t *lw, const int32_t *lh, const int32_t *rows, int n_rows,
                O *out, int32_t *steps, int32_t *hits, int threads) {
  if (!p || !rows || n_rows < 1 || !depth || !oct || !pd || !pr || !out || p->n_levels < 1 || p->n_levels > 16 ||
      p->dirs < 1 || p->grid < 1)
    return -1;
  for (int i = 0; i < n_rows; i++)
    if (rows[i] < 0 || rows[i] >= p->height) return -2;
  Level ld[16], lr[16];
  for (int l = 0; l < p->n_levels; l++) {
    ld[l] = Level{pd[l], lw[l], lh[l]};
    lr[l] = Level{pr[l], lw[l], lh[l]};
  }
  M4<T> view = load<T>(p->view), proj = load<T>(p->proj);
  M4<T> invVP = inverse(mul(proj, view)); /* :123-124 */
  M4<T> invV = inverse(view);             /* :125 */
  V4<T> cam4 = mul(invV, V4<T>{0, 0, 0, 1}); /* :128 */
  V3<T> cam{cam4.x, cam4.y, cam4.z};
  const int w = p->width;
#pragma omp parallel for schedule(dynamic) collapse(2) num_threads(nthreads(threads))
  for (int i = 0; i < n_rows; i++) {
    for (int xb = 0; xb < (w + 63) / 64; xb++) {
      for (int x = xb * 64; x < std::min(w, xb * 64 + 64); x++) {
        T rgb[3];
        int32_t st = 0, hi = 0;
        gather_pixel<T>(*p, invVP, cam, x, rows[i], depth, oct, ld, lr, rgb, &st, &hi);
        size_t o = (size_t)i * w + x;
        out[o * 4 + 0] = O(rgb[0]);
        out[o * 4 + 1] = O(rgb[1]);
        out[o * 4 + 2] = O(rgb[2]);
        out[o * 4 + 3] = O(1);
        if (steps) steps[o] = st;
        if (hits) hits[o] = hi;
      }
    }
  }
  return 0;
}

} // namespace

extern "C" {

uint32_t orc_version(void) { return 1; }

void orc_hammersley(int32_t i, int32_t n, float out[2]) { hammersley(i, n, out[0], out[1]); }

void orc_mat4_inverse(const float m[16], float out[16]) {
  M4<float> r = inverse(load<float>(m));
  for (int i = 0; i < 16; i++) out[i] = r.m[i];
}
void orc_mat4_mul(const float a[16], const float b[16], float out[16]) {
  M4<float> r = mul(load<float>(a), load<float>(b));
  for (int i = 0; i < 16; i++) out[i] = r.m[i];
}

/* glm perspectiveRH_NO (matrix_clip_space.inl:240-252) followed by
 * scale(1,-1,-1) as at SSVGIRenderer.h:58. */
void orc_perspective_rh_no(float fovy, float aspect, float znear, float zfar, float out[16]) {
  float t = std::tan(fovy / 2.0f);
  float p[16] = {0};
  p[0] = 1.0f / (aspect * t);
  p[5] = 1.0f / t;
  p[10] = -(zfar + znear) / (zfar - znear);
  p[11] = -1.0f;
  p[14] = -(2.0f * zfar * znear) / (zfar - znear);
  float s[16] = {1, 0, 0, 0, 0, -1, 0, 0, 0, 0, -1, 0, 0, 0, 0, 1};
  orc_mat4_mul(p, s, out);
}

/* Camera::GetTransformMatrix (Scene.h:30-33): translate(pos) * rotY(hor) * rotX(vert). */
void orc_camera_transform(const float pos[3], float hor, float vert, float out[16]) {
  float t[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, pos[0], pos[1], pos[2], 1};
  float ch = std::cos(hor), sh = std::sin(hor), cv = std::cos(vert), sv = std::sin(vert);
  float ry[16] = {ch, 0, -sh, 0, 0, 1, 0, 0, sh, 0, ch, 0, 0, 0, 0, 1};
  float rx[16] = {1, 0, 0, 0, 0, cv, sv, 0, 0, -sv, cv, 0, 0, 0, 0, 1};
  float a[16];
  orc_mat4_mul(t, ry, a);
  orc_mat4_mul(a, rx, out);
}

float orc_horizon_contribution(const float e[3], const float t[3], const float n[3], float mn, float mx) {
  return horizon_contribution<float>({e[0], e[1], e[2]}, {t[0], t[1], t[2]}, {n[0], n[1], n[2]}, mn, mx);
}
double orc_horizon_contribution_f64(const double e[3], const double t[3], const double n[3], double mn, double mx) {
  return horizon_contribution<double>({e[0], e[1], e[2]}, {t[0], t[1], t[2]}, {n[0], n[1], n[2]}, mn, mx);
}
void orc_oct_decode(float ex, float ey, float out[3]) {
  V3<float> n = oct_decode<float>(ex, ey);
  out[0] = n.x;
  out[1] = n.y;
  out[2] = n.z;
}
uint16_t orc_f32_to_f16(float f) { return f32_to_f16(f); }
float orc_f16_to_f32(uint16_t h) { return f16_to_f32(h); }

/* mipLevelBuilder.frag:22-28 (Avg): sum of children (0,0),(1,0),(0,1),(1,1)
 * in that order, divided by 4; destination size floor(w/2) x floor(h/2)
 * (MipBuilder.h:148-152). */
int orc_mip_avg(const float *src, int32_t w, int32_t h, int32_t c, float *dst, int32_t round_fp16) {
  int dw = w / 2, dh = h / 2;
  if (!src || !dst || dw <= 0 || dh <= 0 || c < 1 || c > 4) return -1;
  static const int off[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
#pragma omp parallel for schedule(static)
  for (int y = 0; y < dh; y++)
    for (int x = 0; x < dw; x++)
      for (int k = 0; k < c; k++) {
        float s = 0.0f;
        for (int i = 0; i < 4; i++) s += src[((size_t)(2 * y + off[i][1]) * w + 2 * x + off[i][0]) * c + k];
        dst[((size_t)y * dw + x) * c + k] = round16(s / 4.0f, round_fp16);
      }
  return 0;
}

/* mipLevelBuilder.frag:29-42 (Depth filter), max channel only: the max-Z
 * Hi-Z keeps max over the 2x2 children and drops the mass term. */
int orc_mip_max(const float *src, int32_t w, int32_t h, float *dst) {
  int dw = w / 2, dh = h / 2;
  if (!src || !dst || dw <= 0 || dh <= 0) return -1;
  static const int off[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
#pragma omp parallel for schedule(static)
  for (int y = 0; y < dh; y++)
    for (int x = 0; x < dw; x++) {
      float m = -1e5f; /* :32 */
      for (int i = 0; i < 4; i++) m = std::max(m, src[(size_t)(2 * y + off[i][1]) * w + 2 * x + off[i][0]]);
      dst[(size_t)y * dw + x] = m;
    }
  return 0;
}

/* blurLayerBuilder.frag:17-35: x outer, y inner over [-r, r), clamp to edge;
 * radius 0 copies. */
int orc_box_blur(const float *src, int32_t w, int32_t h, int32_t c, int32_t radius, float *dst, int32_t round_fp16) {
  if (!src || !dst || w <= 0 || h <= 0 || c < 1 || c > 4 || radius < 0) return -1;
#pragma omp parallel for schedule(static)
  for (int py = 0; py < h; py++)
    for (int px = 0; px < w; px++) {
      float sum[4] = {0, 0, 0, 0}, tw = 0.0f;
      for (int x = -radius; x < radius; x++)
        for (int y = -radius; y < radius; y++) {
          int sx = std::min(std::max(px + x, 0), w - 1), sy = std::min(std::max(py + y, 0), h - 1);
          for (int k = 0; k < c; k++) sum[k] += src[((size_t)sy * w + sx) * c + k];
          tw += 1.0f;
        }
      for (int k = 0; k < c; k++) {
        float v = radius == 0 ? src[((size_t)py * w + px) * c + k] : sum[k] / tw;
        dst[((size_t)py * w + px) * c + k] = round16(v, round_fp16);
      }
    }
  return 0;
}

/* deinterleave.frag:16-28 InterleavePixel: dst (deinterleaved) pixel d reads
 * src (interleaved) pixel (d % cell) * grid + d / cell, cell = size / grid. */
int orc_deinterleave(const void *src, int32_t w, int32_t h, int32_t eb, int32_t gx, int32_t gy, void *dst) {
  if (!src || !dst || gx < 1 || gy < 1 || w % gx || h % gy || eb < 1) return -1;
  int cx = w / gx, cy = h / gy;
  const char *s = (const char *)src;
  char *d = (char *)dst;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      int pxi = x / cx, pyi = y / cy;
      int sx = (x % cx) * gx + pxi, sy = (y % cy) * gy + pyi;
      std::memcpy(d + ((size_t)y * w + x) * eb, s + ((size_t)sy * w + sx) * eb, eb);
    }
  return 0;
}

/* interleave.frag:15-27 DeinterleavePixel: dst (interleaved) pixel i reads
 * src (deinterleaved) pixel (i % grid) * cell + i / grid. */
int orc_interleave(const void *src, int32_t w, int32_t h, int32_t eb, int32_t gx, int32_t gy, void *dst) {
  if (!src || !dst || gx < 1 || gy < 1 || w % gx || h % gy || eb < 1) return -1;
  int cx = w / gx, cy = h / gy;
  const char *s = (const char *)src;
  char *d = (char *)dst;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      int sx = (x % gx) * cx + x / gx, sy = (y % gy) * cy + y / gy;
      std::memcpy(d + ((size_t)y * w + x) * eb, s + ((size_t)sy * w + sx) * eb, eb);
    }
  return 0;
}

void orc_sample_trilinear(const float *const *levels, const int32_t *lw, const int32_t *lh, int32_t nl, int32_t c,
                          float sx, float sy, float lod, float *out) {
  Level lv[16];
  for (int l = 0; l < nl && l < 16; l++) lv[l] = Level{levels[l], lw[l], lh[l]};
  sample_trilinear<float>(lv, nl, c, sx, sy, lod, out);
}
void orc_sample_point(const float *const *levels, const int32_t *lw, const int32_t *lh, int32_t nl, int32_t c,
                      float sx, float sy, float lod, float *out) {
  Level lv[16];
  for (int l = 0; l < nl && l < 16; l++) lv[l] = Level{levels[l], lw[l], lh[l]};
  sample_point<float>(lv, nl, c, sx, sy, lod, out);
}

int orc_gather_f32(const orc_gather_params *p, const float *depth, const float *oct, const float *const *pd,
                   const float *const *pr, const int32_t *lw, const int32_t *lh, float *out, int32_t *steps,
                   int32_t x0, int32_t y0, int32_t cw, int32_t ch, int32_t threads) {
  return gather<float, float>(p, depth, oct, pd, pr, lw, lh, out, steps, x0, y0, cw, ch, threads);
}
int orc_gather_rows_f32(const orc_gather_params *p, const float *depth, const float *oct, const float *const *pd,
                        const float *const *pr, const int32_t *lw, const int32_t *lh, const int32_t *rows,
                        int32_t n_rows, float *out, int32_t *steps, int32_t *hits, int32_t threads) {
  return gather_rows<float, float>(p, depth, oct, pd, pr, lw, lh, rows, n_rows, out, steps, hits, threads);
}
int orc_gather_f64(const orc_gather_params *p, const float *depth, const float *oct, const float *const *pd,
                   const float *const *pr, const int32_t *lw, const int32_t *lh, double *out, int32_t *steps,
                   int32_t x0, int32_t y0, int32_t cw, int32_t ch, int32_t threads) {
  return gather<double, double>(p, depth, oct, pd, pr, lw, lh, out, steps, x0, y0, cw, ch, threads);
}

/* Geometry-aware bilateral: the weights of the commented block at
 * denoiser.frag:87-110 (exp(-|d - dc| * 10.5) * exp(-|n - nc|^2 * 20)),
 * y outer / x inner over [-r, r) (:90-92), taps at pixel centres (exact
 * texels, clamp to edge), colour / (sumW + 1e-7) with alpha 1 (:110).
 * radius 0 passes the input through (:82-86). */
int orc_bilateral(const float *in, const float *depth, const float *oct, int32_t w, int32_t h, int32_t radius,
                  float ds, float ns, float *out, int32_t threads) {
  if (!in || !depth || !oct || !out || w <= 0 || h <= 0 || radius < 0) return -1;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads(threads))
  for (int py = 0; py < h; py++)
    for (int px = 0; px < w; px++) {
      size_t ci = (size_t)py * w + px;
      if (radius == 0) {
        for (int k = 0; k < 4; k++) out[ci * 4 + k] = in[ci * 4 + k];
        continue;
      }
      float dc = depth[ci];
      V3<float> nc = oct_decode<float>(oct[ci * 2], oct[ci * 2 + 1]);
      float sc[3] = {0, 0, 0}, sw = 0.0f;
      for (int y = -radius; y < radius; y++)
        for (int x = -radius; x < radius; x++) {
          int sx = std::min(std::max(px + x, 0), w - 1), sy = std::min(std::max(py + y, 0), h - 1);
          size_t si = (size_t)sy * w + sx;
          V3<float> nn = oct_decode<float>(oct[si * 2], oct[si * 2 + 1]);
          V3<float> nd = nn - nc;
          float wgt = 1.0f;
          wgt *= std::exp(-std::fabs(depth[si] - dc) * ds);
          wgt *= std::exp(-dot(nd, nd) * ns);
          for (int k = 0; k < 3; k++) sc[k] += in[si * 4 + k] * wgt;
          sw += wgt;
        }
      for (int k = 0; k < 3; k++) out[ci * 4 + k] = sc[k] / (sw + 1e-7f);
      out[ci * 4 + 3] = 1.0f;
    }
  return 0;
}

/* denoiser.frag:72-185: least-squares colour = a * depth + b over the 16
 * taps [-2,2)^2 (y outer, x inner), Gram matrix + 1e-7 ridge, evaluated at
 * the centre.  Taps at pixel centres, clamp to edge.  The 2x2 block of the
 * mat4 inverse (:148) is solved in closed form. */
int orc_regression_denoise(const float *in, const float *depth, int32_t w, int32_t h, int32_t radius, float *out) {
  if (!in || !depth || !out || w <= 0 || h <= 0) return -1;
#pragma omp parallel for schedule(static)
  for (int py = 0; py < h; py++)
    for (int px = 0; px < w; px++) {
      size_t ci = (size_t)py * w + px;
      if (radius == 0) {
        for (int k = 0; k < 4; k++) out[ci * 4 + k] = in[ci * 4 + k];
        continue;
      }
      float dep[16], col[16][3];
      int i = 0;
      for (int y = -2; y < 2; y++)
        for (int x = -2; x < 2; x++) {
          int sx = std::min(std::max(px + x, 0), w - 1), sy = std::min(std::max(py + y, 0), h - 1);
          size_t si = (size_t)sy * w + sx;
          dep[i] = depth[si];
          for (int k = 0; k < 3; k++) col[i][k] = in[si * 4 + k];
          i++;
        }
      float g00 = 0, g01 = 0, g11 = 0;
      for (int k = 0; k < 16; k++) g00 += dep[k] * dep[k];
      for (int k = 0; k < 16; k++) g01 += dep[k] * 1.0f;
      for (int k = 0; k < 16; k++) g11 += 1.0f * 1.0f;
      g00 += 1e-7f;
      g11 += 1e-7f;
      float det = g00 * g11 - g01 * g01;
      float i00 = g11 / det, i01 = -g01 / det, i11 = g00 / det;
      float dcen = depth[ci];
      for (int ch = 0; ch < 3; ch++) {
        float m0 = 0, m1 = 0;
        for (int k = 0; k < 16; k++) m0 += dep[k] * col[k][ch];
        for (int k = 0; k < 16; k++) m1 += 1.0f * col[k][ch];
        float c0 = i00 * m0 + i01 * m1;
        float c1 = i01 * m0 + i11 * m1;
        out[ci * 4 + ch] = c0 * dcen + c1 * 1.0f;
      }
      out[ci * 4 + 3] = 1.0f;
    }
  return 0;
}

/* finalGatherer.frag:42-58: direct + indirect * albedo (the blurred-direct
 * terms carry weight 0 and total weight stays 1). */
int orc_composite(const float *direct, const float *albedo, const float *indirect, int32_t w, int32_t h, float *out) {
  if (!direct || !albedo || !indirect || !out) return -1;
  size_t n = (size_t)w * h * 4;
  for (size_t i = 0; i < n; i++) out[i] = direct[i] + indirect[i] * albedo[i];
  return 0;
}

double orc_rel_l2(const float *x, const float *y, int64_t n, int32_t stride, int32_t channels) {
  double num = 0, den = 0;
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < channels; k++) {
      double a = x[i * stride + k], b = y[i * stride + k];
      num += (a - b) * (a - b);
      den += b * b;
    }
  return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}
double orc_rel_l2_f64(const float *x, const double *y, int64_t n, int32_t stride, int32_t channels) {
  double num = 0, den = 0;
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < channels; k++) {
      double a = x[i * stride + k], b = y[i * stride + k];
      num += (a - b) * (a - b);
      den += b * b;
    }
  return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}
int32_t orc_max_threads(void) { return nthreads(0); }

} // extern "C"
