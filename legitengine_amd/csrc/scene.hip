// Seeded analytic scene and G-buffer producer.
//
// Replaces the reference's ShadowPass / GBufferPass / LightPass
// (src/Render/Renderers/SSVGIRenderer.h:61-205) with a closed-form ray
// caster, so the GPU path and the CPU restatement consume byte-identical
// inputs.  The G-buffer keeps the reference's contract: depth is the radial
// distance ||P - cam|| (gBufferBuilder.frag:31,37), the normal is the world
// normal, radiance is diffuse direct light with hard shadows
// (directLighting.frag:45-83, here a directional light).
//
// Scene: the inside of the unit cube (no background pixels), per-wall
// albedo U[0.1,0.9]^3, axis-aligned boxes with extents U[0.05,0.3] per axis
// fully inside the room, spheres of radius U[0.05,0.15], a 60 degree pinhole
// camera at a free point, one directional light whose shadow rays test boxes
// and spheres only (a closed room would otherwise shadow everything).
#include <math.h>

#include <thread>
#include <vector>

#include "lgx_internal.h"

namespace lgx {

// ------------------------------------------------------------ host maths

Mat4 mat4_mul(const Mat4 &a, const Mat4 &b) {
  Mat4 r;
  for (int c = 0; c < 4; c++)
    for (int row = 0; row < 4; row++) {
      float s = 0.0f;
      for (int k = 0; k < 4; k++) s += a.m[k * 4 + row] * b.m[c * 4 + k];
      r.m[c * 4 + row] = s;
    }
  return r;
}
Mat4d mat4d_mul(const Mat4d &a, const Mat4d &b) {
  Mat4d r;
  for (int c = 0; c < 4; c++)
    for (int row = 0; row < 4; row++) {
      double s = 0.0;
      for (int k = 0; k < 4; k++) s += a.m[k * 4 + row] * b.m[c * 4 + k];
      r.m[c * 4 + row] = s;
    }
  return r;
}

// Inverse by the adjugate (cofactors of the 3x3 minors over the determinant).
template <class T> static void inverse4(const T *m, T *out) {
  auto at = [&](int r, int c) { return m[c * 4 + r]; };
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
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) out[c * 4 + r] = cof[c][r] / det;
}
Mat4 mat4_inverse(const Mat4 &a) {
  Mat4 r;
  inverse4<float>(a.m, r.m);
  return r;
}
Mat4d mat4d_inverse(const Mat4d &a) {
  Mat4d r;
  inverse4<double>(a.m, r.m);
  return r;
}

int hip_status(hipError_t e) { return e == hipSuccess ? LGX_OK : (int)e; }

// ------------------------------------------------------------ seeded scene

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  float uni() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }  // [0,1), 24 bits
  float uni(float a, float b) { return a + (b - a) * uni(); }
};

struct Ray {
  float ox, oy, oz, dx, dy, dz;
};

LGX_HD bool hit_box(const float *mn, const float *mx, const Ray &r, float &t, int &axis) {
  float o[3] = {r.ox, r.oy, r.oz}, d[3] = {r.dx, r.dy, r.dz};
  float tn = -1e30f, tf = 1e30f;
  int ax = 0;
  for (int i = 0; i < 3; i++) {
    float inv = 1.0f / d[i];
    float t0 = (mn[i] - o[i]) * inv, t1 = (mx[i] - o[i]) * inv;
    if (t0 > t1) {
      float tmp = t0;
      t0 = t1;
      t1 = tmp;
    }
    if (t0 > tn) {
      tn = t0;
      ax = i;
    }
    if (t1 < tf) tf = t1;
  }
  if (tn <= tf && tn > 1e-6f) {
    t = tn;
    axis = ax;
    return true;
  }
  return false;
}

LGX_HD bool hit_sphere(const float *c, float rad, const Ray &r, float &t) {
  float lx = r.ox - c[0], ly = r.oy - c[1], lz = r.oz - c[2];
  float b = lx * r.dx + ly * r.dy + lz * r.dz;
  float cc = lx * lx + ly * ly + lz * lz - rad * rad;
  float disc = b * b - cc;
  if (disc < 0.0f) return false;
  float sq = sqrtf(disc);
  float t0 = -b - sq;
  if (t0 > 1e-6f) {
    t = t0;
    return true;
  }
  float t1 = -b + sq;
  if (t1 > 1e-6f) {
    t = t1;
    return true;
  }
  return false;
}

struct PixelOut {
  float depth, nx, ny, nz, ar, ag, ab, lr, lg, lb;
};

LGX_HD PixelOut shade(const lgx_scene &s, const float *inv_vp, const float *cam, int x, int y, int w, int h) {
  float sx = ((float)x + 0.5f) / (float)w, sy = ((float)y + 0.5f) / (float)h;
  float nx_ = sx * 2.0f - 1.0f, ny_ = sy * 2.0f - 1.0f;
  float v[4];
  for (int r = 0; r < 4; r++) v[r] = inv_vp[r] * nx_ + inv_vp[4 + r] * ny_ + inv_vp[8 + r] + inv_vp[12 + r];
  float px = v[0] / v[3] - cam[0], py = v[1] / v[3] - cam[1], pz = v[2] / v[3] - cam[2];
  float il = 1.0f / sqrtf(px * px + py * py + pz * pz);
  Ray ray{cam[0], cam[1], cam[2], px * il, py * il, pz * il};
  // room: inside of [0,1]^3
  float best = 1e30f;
  float n[3] = {0, 0, 0}, alb[3] = {0.5f, 0.5f, 0.5f};
  {
    float o[3] = {ray.ox, ray.oy, ray.oz}, d[3] = {ray.dx, ray.dy, ray.dz};
    for (int i = 0; i < 3; i++) {
      if (d[i] == 0.0f) continue;
      float t = d[i] > 0.0f ? (1.0f - o[i]) / d[i] : (0.0f - o[i]) / d[i];
      if (t > 0.0f && t < best) {
        best = t;
        n[0] = n[1] = n[2] = 0.0f;
        n[i] = d[i] > 0.0f ? -1.0f : 1.0f;
        int wall = i * 2 + (d[i] > 0.0f ? 1 : 0);
        alb[0] = s.wall_albedo[wall][0];
        alb[1] = s.wall_albedo[wall][1];
        alb[2] = s.wall_albedo[wall][2];
      }
    }
  }
  for (int b = 0; b < s.n_boxes; b++) {
    float t;
    int ax;
    if (hit_box(s.box_min[b], s.box_max[b], ray, t, ax) && t < best) {
      best = t;
      float d[3] = {ray.dx, ray.dy, ray.dz};
      n[0] = n[1] = n[2] = 0.0f;
      n[ax] = d[ax] > 0.0f ? -1.0f : 1.0f;
      alb[0] = s.box_albedo[b][0];
      alb[1] = s.box_albedo[b][1];
      alb[2] = s.box_albedo[b][2];
    }
  }
  for (int k = 0; k < s.n_spheres; k++) {
    float t;
    if (hit_sphere(s.sph_center[k], s.sph_radius[k], ray, t) && t < best) {
      best = t;
      float hx = ray.ox + ray.dx * t - s.sph_center[k][0];
      float hy = ray.oy + ray.dy * t - s.sph_center[k][1];
      float hz = ray.oz + ray.dz * t - s.sph_center[k][2];
      float inv = 1.0f / sqrtf(hx * hx + hy * hy + hz * hz);
      n[0] = hx * inv;
      n[1] = hy * inv;
      n[2] = hz * inv;
      alb[0] = s.sph_albedo[k][0];
      alb[1] = s.sph_albedo[k][1];
      alb[2] = s.sph_albedo[k][2];
    }
  }
  PixelOut o;
  o.depth = best;
  o.nx = n[0];
  o.ny = n[1];
  o.nz = n[2];
  o.ar = alb[0];
  o.ag = alb[1];
  o.ab = alb[2];
  float ndl = n[0] * s.light_dir[0] + n[1] * s.light_dir[1] + n[2] * s.light_dir[2];
  float vis = 0.0f;
  if (ndl > 0.0f) {
    vis = 1.0f;
    Ray sr{ray.ox + ray.dx * best + n[0] * 1e-4f, ray.oy + ray.dy * best + n[1] * 1e-4f,
           ray.oz + ray.dz * best + n[2] * 1e-4f, s.light_dir[0], s.light_dir[1], s.light_dir[2]};
    for (int b = 0; b < s.n_boxes && vis > 0.0f; b++) {
      float t;
      int ax;
      if (hit_box(s.box_min[b], s.box_max[b], sr, t, ax)) vis = 0.0f;
    }
    for (int k = 0; k < s.n_spheres && vis > 0.0f; k++) {
      float t;
      if (hit_sphere(s.sph_center[k], s.sph_radius[k], sr, t)) vis = 0.0f;
    }
  }
  float e = s.light_intensity * (ndl > 0.0f ? ndl : 0.0f) * vis;
  o.lr = alb[0] * e;
  o.lg = alb[1] * e;
  o.lb = alb[2] * e;
  return o;
}

struct GbufPtrs {
  void *depth, *normal, *albedo, *radiance;
  int dp, np, ap, rp;
  int nf, af, rf;
};

LGX_HD void write_pixel(const GbufPtrs &g, int x, int y, const PixelOut &o) {
  ((float *)g.depth)[(size_t)y * g.dp + x] = o.depth;
  float ex, ey;
  oct_encode<float>(Vec3<float>{o.nx, o.ny, o.nz}, ex, ey);
  size_t ni = (size_t)y * g.np + x;
  if (g.nf == LGX_RG16F)
    ((uint32_t *)g.normal)[ni] = (uint32_t)f2h(ex) | ((uint32_t)f2h(ey) << 16);
  else
    ((float2 *)g.normal)[ni] = make_float2(ex, ey);
  size_t ai = (size_t)y * g.ap + x;
  if (g.af == LGX_RGBA8) {
    uint32_t r = (uint32_t)(o.ar * 255.0f + 0.5f), gg = (uint32_t)(o.ag * 255.0f + 0.5f),
             b = (uint32_t)(o.ab * 255.0f + 0.5f);
    ((uint32_t *)g.albedo)[ai] = r | (gg << 8) | (b << 16) | (255u << 24);
  } else {
    ((float4 *)g.albedo)[ai] = make_float4(o.ar, o.ag, o.ab, 1.0f);
  }
  size_t ri = (size_t)y * g.rp + x;
  if (g.rf == LGX_RGBA16F)
    ((uint2 *)g.radiance)[ri] = make_uint2((uint32_t)f2h(o.lr) | ((uint32_t)f2h(o.lg) << 16),
                                           (uint32_t)f2h(o.lb) | ((uint32_t)f2h(1.0f) << 16));
  else
    ((float4 *)g.radiance)[ri] = make_float4(o.lr, o.lg, o.lb, 1.0f);
}

struct SceneKernelArgs {
  lgx_scene s;
  float inv_vp[16];
  float cam[3];
  int w, h;
  GbufPtrs g;
};

__global__ __launch_bounds__(256) void scene_kernel(const SceneKernelArgs a) {
  int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= a.w || y >= a.h) return;
  write_pixel(a.g, x, y, shade(a.s, a.inv_vp, a.cam, x, y, a.w, a.h));
}

static bool gbuf_ok(const lgx_gbuffer *g) {
  if (!g || !image_ok(&g->depth) || !image_ok(&g->normal) || !image_ok(&g->albedo) || !image_ok(&g->radiance))
    return false;
  int w = g->depth.w, h = g->depth.h;
  const lgx_image *ims[4] = {&g->depth, &g->normal, &g->albedo, &g->radiance};
  for (auto *im : ims)
    if (im->w != w || im->h != h) return false;
  return g->depth.format == LGX_R32F && (g->normal.format == LGX_RG16F || g->normal.format == LGX_RG32F) &&
         (g->albedo.format == LGX_RGBA8 || g->albedo.format == LGX_RGBA32F) &&
         (g->radiance.format == LGX_RGBA16F || g->radiance.format == LGX_RGBA32F);
}

static void setup_args(const lgx_scene *s, const lgx_gbuffer *g, SceneKernelArgs &a) {
  a.s = *s;
  a.w = g->depth.w;
  a.h = g->depth.h;
  float view[16], proj[16];
  lgx_scene_matrices(s, a.w, a.h, view, proj);
  Mat4 v, p;
  memcpy(v.m, view, 64);
  memcpy(p.m, proj, 64);
  Mat4 inv = mat4_inverse(mat4_mul(p, v));
  Mat4 iv = mat4_inverse(v);
  memcpy(a.inv_vp, inv.m, 64);
  a.cam[0] = iv.m[12];
  a.cam[1] = iv.m[13];
  a.cam[2] = iv.m[14];
  a.g = GbufPtrs{g->depth.ptr, g->normal.ptr, g->albedo.ptr, g->radiance.ptr, g->depth.pitch, g->normal.pitch,
                 g->albedo.pitch, g->radiance.pitch, g->normal.format, g->albedo.format, g->radiance.format};
}

// Camera at a free point (outside every box and sphere by a margin), aimed
// at a point in the lower middle of the room.
static void sample_camera(lgx_scene *out, SplitMix64 &r) {
  const int n_boxes = out->n_boxes, n_spheres = out->n_spheres;
  float pos[3] = {0.5f, 0.5f, 0.5f};
  for (int tries = 0; tries < 4096; tries++) {
    float c[3] = {r.uni(0.08f, 0.92f), r.uni(0.08f, 0.92f), r.uni(0.08f, 0.92f)};
    bool free_ = true;
    const float m = 0.03f;
    for (int b = 0; b < n_boxes && free_; b++) {
      bool inside = true;
      for (int i = 0; i < 3; i++)
        if (c[i] < out->box_min[b][i] - m || c[i] > out->box_max[b][i] + m) inside = false;
      if (inside) free_ = false;
    }
    for (int k = 0; k < n_spheres && free_; k++) {
      float dx = c[0] - out->sph_center[k][0], dy = c[1] - out->sph_center[k][1], dz = c[2] - out->sph_center[k][2];
      float rr = out->sph_radius[k] + m;
      if (dx * dx + dy * dy + dz * dz < rr * rr) free_ = false;
    }
    if (free_) {
      pos[0] = c[0];
      pos[1] = c[1];
      pos[2] = c[2];
      break;
    }
  }
  out->cam_pos[0] = pos[0];
  out->cam_pos[1] = pos[1];
  out->cam_pos[2] = pos[2];
  // aim at a point in the lower middle of the room: forward =
  // rotY(hor) * rotX(vert) * (0,0,1) = (sin hor cos vert, -sin vert, cos hor cos vert)
  {
    float tx = r.uni(0.25f, 0.75f), ty = r.uni(0.1f, 0.45f), tz = r.uni(0.25f, 0.75f);
    float fx = tx - pos[0], fy = ty - pos[1], fz = tz - pos[2];
    float fl = sqrtf(fx * fx + fy * fy + fz * fz);
    if (fl < 1e-3f) {
      fx = 0.0f;
      fy = -0.3f;
      fz = 1.0f;
      fl = sqrtf(fy * fy + fz * fz);
    }
    out->cam_hor = atan2f(fx, fz);
    out->cam_vert = -asinf(fy / fl);
  }
}

}  // namespace lgx

using namespace lgx;

extern "C" int lgx_scene_make(uint64_t seed, int32_t n_boxes, int32_t n_spheres, lgx_scene *out) {
  if (!out || n_boxes < 0 || n_boxes > 32 || n_spheres < 0 || n_spheres > 8) return LGX_EINVAL;
  memset(out, 0, sizeof(*out));
  SplitMix64 r{seed};
  for (int i = 0; i < 6; i++)
    for (int c = 0; c < 3; c++) out->wall_albedo[i][c] = r.uni(0.1f, 0.9f);
  out->n_boxes = n_boxes;
  out->n_spheres = n_spheres;
  for (int b = 0; b < n_boxes; b++) {
    for (int i = 0; i < 3; i++) {
      float ext = r.uni(0.05f, 0.3f);
      float mn = r.uni(0.0f, 1.0f - ext);
      out->box_min[b][i] = mn;
      out->box_max[b][i] = mn + ext;
    }
    for (int c = 0; c < 3; c++) out->box_albedo[b][c] = r.uni(0.1f, 0.9f);
  }
  for (int k = 0; k < n_spheres; k++) {
    float rad = r.uni(0.05f, 0.15f);
    out->sph_radius[k] = rad;
    for (int i = 0; i < 3; i++) out->sph_center[k][i] = r.uni(rad, 1.0f - rad);
    for (int c = 0; c < 3; c++) out->sph_albedo[k][c] = r.uni(0.1f, 0.9f);
  }
  {
    float lx = r.uni(-0.8f, 0.8f), lz = r.uni(-0.8f, 0.8f), ly = 1.0f;
    float il = 1.0f / sqrtf(lx * lx + ly * ly + lz * lz);
    out->light_dir[0] = lx * il;
    out->light_dir[1] = ly * il;
    out->light_dir[2] = lz * il;
    out->light_intensity = 3.0f;
  }
  sample_camera(out, r);
  out->fovy = 1.04719755f;  // 60 degrees
  out->znear = 0.01f;       // SSVGIRenderer.h:58
  out->zfar = 1000.0f;
  return LGX_OK;
}

extern "C" int lgx_scene_pose(lgx_scene *s, uint64_t pose_seed) {
  if (!s) return LGX_EINVAL;
  SplitMix64 r{pose_seed ^ 0x5DEECE66Dull};
  sample_camera(s, r);
  return LGX_OK;
}

extern "C" int lgx_scene_matrices(const lgx_scene *s, int32_t width, int32_t height, float view[16], float proj[16]) {
  if (!s || !view || !proj || width <= 0 || height <= 0) return LGX_EINVAL;
  // Camera::GetTransformMatrix (Scene.h:30-33)
  float ch = cosf(s->cam_hor), sh = sinf(s->cam_hor), cv = cosf(s->cam_vert), sv = sinf(s->cam_vert);
  Mat4 t = {{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, s->cam_pos[0], s->cam_pos[1], s->cam_pos[2], 1}};
  Mat4 ry = {{ch, 0, -sh, 0, 0, 1, 0, 0, sh, 0, ch, 0, 0, 0, 0, 1}};
  Mat4 rx = {{1, 0, 0, 0, 0, cv, sv, 0, 0, -sv, cv, 0, 0, 0, 0, 1}};
  Mat4 v = mat4_inverse(mat4_mul(mat4_mul(t, ry), rx));  // SSVGIRenderer.h:54
  // perspectiveRH_NO * scale(1,-1,-1) (SSVGIRenderer.h:58)
  float aspect = (float)width / (float)height;
  float th = tanf(s->fovy / 2.0f);
  Mat4 p = {{0}};
  p.m[0] = 1.0f / (aspect * th);
  p.m[5] = 1.0f / th;
  p.m[10] = -(s->zfar + s->znear) / (s->zfar - s->znear);
  p.m[11] = -1.0f;
  p.m[14] = -(2.0f * s->zfar * s->znear) / (s->zfar - s->znear);
  Mat4 sc = {{1, 0, 0, 0, 0, -1, 0, 0, 0, 0, -1, 0, 0, 0, 0, 1}};
  Mat4 pr = mat4_mul(p, sc);
  memcpy(view, v.m, 64);
  memcpy(proj, pr.m, 64);
  return LGX_OK;
}

extern "C" int lgx_scene_render_cpu(const lgx_scene *s, const lgx_gbuffer *g) {
  if (!s || !gbuf_ok(g)) return LGX_EINVAL;
  SceneKernelArgs a;
  setup_args(s, g, a);
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t]() {
      for (int y = (int)t; y < a.h; y += (int)nt)
        for (int x = 0; x < a.w; x++) write_pixel(a.g, x, y, shade(a.s, a.inv_vp, a.cam, x, y, a.w, a.h));
    });
  for (auto &t : th) t.join();
  return LGX_OK;
}

extern "C" int lgx_scene_render_gpu(const lgx_scene *s, const lgx_gbuffer *g, lgx_stream_t stream) {
  if (!s || !gbuf_ok(g)) return LGX_EINVAL;
  SceneKernelArgs a;
  setup_args(s, g, a);
  dim3 grid((a.w + 15) / 16, (a.h + 15) / 16), block(256);
  hipLaunchKernelGGL(scene_kernel, grid, block, 0, (hipStream_t)stream, a);
  return hip_status(hipGetLastError());
}

extern "C" int lgx_copy_image(const lgx_image *src, lgx_image *dst, lgx_stream_t stream) {
  if (!image_ok(src) || !image_ok(dst)) return LGX_EINVAL;
  if (src->format != dst->format) return LGX_EFORMAT;
  if (src->w != dst->w || src->h != dst->h) return LGX_ESIZE;
  const size_t eb = (size_t)format_bytes(src->format);
  return hip_status(hipMemcpy2DAsync(dst->ptr, (size_t)dst->pitch * eb, src->ptr, (size_t)src->pitch * eb,
                                     (size_t)src->w * eb, (size_t)src->h, hipMemcpyDeviceToDevice,
                                     (hipStream_t)stream));
}

extern "C" const char *lgx_strerror(int code) {
  switch (code) {
    case LGX_OK: return "ok";
    case LGX_EINVAL: return "invalid argument";
    case LGX_EFORMAT: return "unsupported image format for this entry point";
    case LGX_ESIZE: return "inconsistent image extents";
    case LGX_EUNSUPPORTED: return "parameter outside the supported range";
    case LGX_EWORKSPACE: return "workspace missing or not prepared";
    case LGX_ENODEVICE: return "no HIP device";
    default: break;
  }
  if (code > 0) return hipGetErrorString((hipError_t)code);
  return "unknown lgx error";
}

extern "C" int32_t lgx_version(void) { return 1; }
