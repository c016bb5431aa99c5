"""Closed-form scenes shared by the oracle known-answer tests and the GPU tests.

Sphere-around-the-camera gather scene (pins the gather's hit branch,
indirectLighting.frag:254-264):

  The camera sits at the origin with the identity view, and every depth
  texel holds the same radial distance r. Every sample then reconstructs
  (cam + ray(sq) * r, :241) onto the sphere of radius r around the camera.
  Each G-buffer normal faces the camera (n = -ray(c)).

  In the slice plane spanned by the camera rays of one screen direction,
  take coordinates with the origin at the camera and the x axis along the
  centre ray. Let phi be the angle between the centre ray and the ray of a
  sample, and delta the angle to the ray of the one-pixel offset point
  (:180-181). Then:
    P = r (1, 0)               omega0 = (-1, 0)
    tau = (-sin(delta/2), cos(delta/2))        (unit chord of the two rays, :183)
    Q(phi) = r (cos phi, sin phi)
    Q - P = 2 r sin(phi/2) (-sin(phi/2), cos(phi/2))
  so the sample's horizon angle (:249-252) is
    theta(phi) = atan2(cos((phi - delta)/2), sin(phi/2)),
  with tan(theta) = cot(phi/2) cos(delta/2) + sin(delta/2), which falls
  strictly as phi grows.

  The normal-limited initial horizon (:193-198) with n = omega0 is pi/2.
  The component of tau perpendicular to n has no omega0 part. So every
  on-screen sample rises above the running horizon, and the hits come in
  marching order.

  The slice integral h(a, b) of the cosine-weighted measure,
    int_a^b sin(t) n.(cos(t) omega0 + sin(t) tau) dt,
  has omega0.n = 1 and tau.n = sin(delta/2) here. This module evaluates it
  by Gauss-Legendre quadrature, not by the oracle's closed form.
"""
import ctypes as C
import math

import numpy as np

AMBIENT = float(np.float32(0.01))  # indirectLighting.frag:154, a float constant
PI_F = float(np.float32(3.1415))    # :169, verbatim and rounded to float


def projection(w, h, fovy=1.0):
    """perspectiveRH_NO(fovy, w/h, 0.01, 1000) * scale(1,-1,-1), column-major float32."""
    t = math.tan(fovy / 2)
    zn, zf = 0.01, 1000.0
    aspect = w / h
    ref = np.array([[1 / (aspect * t), 0, 0, 0], [0, 1 / t, 0, 0],
                    [0, 0, -(zf + zn) / (zf - zn), -2 * zf * zn / (zf - zn)], [0, 0, -1, 0]])
    ref = ref @ np.diag([1, -1, -1, 1])
    return ref.T.ravel().astype(np.float32)


_INV = {}
_GL = {}


def rays(proj, sx, sy):
    """Unit camera rays through normalized screen points (identity view, camera at 0)."""
    key = np.asarray(proj, np.float32).tobytes()
    if key not in _INV:
        _INV[key] = np.linalg.inv(np.asarray(proj, np.float64).reshape(4, 4).T)
    inv = _INV[key]
    sx = np.asarray(sx, np.float64)
    sy = np.asarray(sy, np.float64)
    v = np.einsum("ij,j...->i...", inv, np.stack([sx * 2 - 1, sy * 2 - 1, np.ones_like(sx), np.ones_like(sx)]))
    d = v[:3] / v[3]
    return d / np.linalg.norm(d, axis=0)


def oct_encode(n):
    """Octahedral encoding (Cigolle et al. 2014) of unit normals [..., 3]."""
    n = np.asarray(n, np.float64)
    p = n[..., :2] / np.abs(n).sum(-1, keepdims=True)
    neg = n[..., 2] < 0
    fx = (1 - np.abs(p[..., 1])) * np.where(p[..., 0] >= 0, 1.0, -1.0)
    fy = (1 - np.abs(p[..., 0])) * np.where(p[..., 1] >= 0, 1.0, -1.0)
    out = p.copy()
    out[neg, 0] = fx[neg]
    out[neg, 1] = fy[neg]
    return out.astype(np.float32)


def slice_h(tau_n, a, b, order=64):
    """int_a^b sin(t) (cos(t) + sin(t) tau_n) dt (omega0.n = 1) by Gauss-Legendre quadrature."""
    if order not in _GL:
        _GL[order] = np.polynomial.legendre.leggauss(order)
    xs, ws = _GL[order]
    t = 0.5 * (b - a) * xs + 0.5 * (a + b)
    f = np.sin(t) * (np.cos(t) + np.sin(t) * tau_n)
    return 0.5 * (b - a) * float((ws * f).sum())


def sphere_scene(w=80, h=60, r=2.0, radiance=0.5, ring_ambient=True):
    """G-buffer and one-level pyramid of the sphere scene.

    radiance: one value for rgb, or three (the closed form is per channel).
    ring_ambient: the radiance is `radiance` on the texels whose samples keep
    the full side weight (x in [0.1 w, 0.9 w), y in [0.1 h, 0.9 h)) and the
    ambient value on the outer ring, where the side fade (:220-228) is below
    one. Ring hits then add exactly nothing, and the hit sum telescopes to
    h(theta_last_interior, pi/2). Otherwise the radiance is uniform and every
    ring hit counts with its side weight.
    """
    assert (w % 10) == 0 and (h % 10) == 0
    proj = projection(w, h)
    view = np.eye(4, dtype=np.float32).T.ravel().copy()
    ys, xs = np.mgrid[0:h, 0:w]
    rc = rays(proj, (xs + 0.5) / w, (ys + 0.5) / h)
    normal = np.moveaxis(-rc, 0, -1)
    octn = oct_encode(normal)
    depth = np.full((h, w), r, np.float32)
    rad = np.full((h, w, 4), AMBIENT, np.float32)
    if not ring_ambient:
        rad[..., :3] = radiance
    else:
        rad[h // 10:h - h // 10, w // 10:w - w // 10, :3] = radiance
    rad[..., 3] = 1.0
    return dict(view=view, proj=proj, depth=depth, oct=octn, rad=rad, w=w, h=h, r=r, radiance=radiance,
                ring_ambient=ring_ambient)


def sphere_expected(sc, px, py, dirs, footprint_dirs):
    """Gather output of pixel (px, py) on the 1x1 pattern (Hammersley(0, 1)
    = (0, 0): directions 2 pi d / D, no linear offset), from the closed
    forms above: a float for one radiance value, rgb for three."""
    if np.ndim(sc["radiance"]):
        return np.array([sphere_expected(dict(sc, radiance=c), px, py, dirs, footprint_dirs)
                         for c in sc["radiance"]])
    w, h, proj = sc["w"], sc["h"], sc["proj"]
    L, amb = float(np.float32(sc["radiance"])), AMBIENT
    cx, cy = px + 0.5, py + 0.5
    rc = rays(proj, cx / w, cy / h)
    near = w / 1000.0
    ratio = 2 * PI_F / footprint_dirs + 1
    total = 0.0
    for d in range(dirs):
        ang = 2 * PI_F / dirs * d
        dx, dy = math.cos(ang), math.sin(ang)
        ro = rays(proj, (cx + dx * 1.0) / w, (cy + dy * 1.0) / h)
        delta = 2 * math.asin(np.linalg.norm(ro - rc) / 2)
        tau_n = math.sin(delta / 2)
        ts = [(0 - cx) / dx if dx else math.inf, (w - cx) / dx if dx else math.inf,
              (0 - cy) / dy if dy else math.inf, (h - cy) / dy if dy else math.inf]
        path = abs(min(max(ts[0], ts[1]), max(ts[2], ts[3])))
        iters = int(math.log(path / near) / math.log(ratio)) + 1
        acc = amb * slice_h(tau_n, 0.0, math.pi / 2)
        prev = math.pi / 2
        for j in range(iters):
            off = near * ratio ** j + 1 - near
            qx, qy = cx + dx * off, cy + dy * off
            sx, sy = qx / w, qy / h
            if not (0 < sx < 1 and 0 < sy < 1):
                break  # off screen: side weight 0, and the ray never comes back
            rq = rays(proj, sx, sy)
            phi = 2 * math.asin(np.linalg.norm(rq - rc) / 2)
            theta = math.atan2(math.cos((phi - delta) / 2), math.sin(phi / 2))
            side = 1.0
            for v in (sx, 1 - sx, sy, 1 - sy):
                side *= min(max(v / 0.1, 0.0), 1.0)
            if sc["ring_ambient"]:
                tx, ty = int(math.floor(qx)), int(math.floor(qy))
                inner = w // 10 <= tx < w - w // 10 and h // 10 <= ty < h - h // 10
                lj = L if inner else amb
            else:
                lj = L
            acc += (lj - amb) * side * slice_h(tau_n, theta, prev)
            prev = theta
        total += 2 * acc / dirs
    return total


def slice_angles(sc, px, py, dirs):
    """theta of every on-screen sample per direction (for the monotonicity precondition)."""
    w, h, proj = sc["w"], sc["h"], sc["proj"]
    cx, cy = px + 0.5, py + 0.5
    rc = rays(proj, cx / w, cy / h)
    out = []
    for d in range(dirs):
        ang = 2 * PI_F / dirs * d
        dx, dy = math.cos(ang), math.sin(ang)
        ro = rays(proj, (cx + dx) / w, (cy + dy) / h)
        delta = 2 * math.asin(np.linalg.norm(ro - rc) / 2)
        th = []
        for off in np.linspace(1.0, 200.0, 400):
            sx, sy = (cx + dx * off) / w, (cy + dy * off) / h
            if not (0 < sx < 1 and 0 < sy < 1):
                break
            phi = 2 * math.asin(np.linalg.norm(rays(proj, sx, sy) - rc) / 2)
            th.append(math.atan2(math.cos((phi - delta) / 2), math.sin(phi / 2)))
        out.append(th)
    return out


def two_region_field(w=20, h=16, split=10, d=(1.0, 1.05), alpha=math.acos(0.98)):
    """Bilateral scene: columns x < split hold depth d[0] and normal (0, 0, 1);
    the others hold depth d[1] and the normal tilted by alpha about y. Colour
    0.2 on the left and 0.8 on the right."""
    depth = np.full((h, w), d[0], np.float32)
    depth[:, split:] = d[1]
    n = np.zeros((h, w, 3))
    n[..., 2] = 1
    n[:, split:] = [math.sin(alpha), 0, math.cos(alpha)]
    octn = oct_encode(n)
    col = np.zeros((h, w, 4), np.float32)
    col[:, :split, :3] = 0.2
    col[:, split:, :3] = 0.8
    col[..., 3] = 1
    return depth, octn, col


def bilateral_expected(depth, octn, col, radius, split, ds=10.5, ns=20.0, decode=None):
    """Two-region bilateral known answer, per column.

    A tap in the pixel's own region weighs 1. A tap in the other region weighs
    w = exp(-|d1 - d0| ds) exp(-|n1 - n0|^2 ns) (denoiser.frag:102-105). Of the
    2r columns x + [-r, r), clamped to the edge, k_own fall in the own region
    and k_oth in the other. Each column contributes 2r rows, so the output is
      (k_own c_own + k_oth w c_oth) / (k_own + k_oth w + 1e-7 / (2r)).
    """
    h, w = depth.shape
    nrm = decode(octn)
    dd = abs(float(depth[0, split]) - float(depth[0, 0]))
    dn = nrm[0, split].astype(np.float64) - nrm[0, 0]
    wc = math.exp(-dd * ds) * math.exp(-float(dn @ dn) * ns)
    out = np.zeros((h, w, 3))
    for x in range(w):
        cols = [min(max(x + t, 0), w - 1) for t in range(-radius, radius)]
        left = sum(c < split for c in cols)
        right = len(cols) - left
        own, oth = (left, right) if x < split else (right, left)
        c_own = float(col[0, x, 0])
        c_oth = float(col[0, split if x < split else 0, 0])
        out[:, x, :] = (own * c_own + oth * wc * c_oth) / (own + oth * wc + 1e-7 / (2 * radius))
    return out, wc


# ---------------------------------------------------------------- scene generator pins

def known_scene(lgx, with_objects=True):
    """A hand-placed scene: the unit room, one floating box, one sphere, a
    camera looking down into the room and a directional light from above."""
    s = lgx.Scene()
    for i in range(6):
        for k in range(3):
            s.wall_albedo[i][k] = (0.2, 0.35, 0.5, 0.65, 0.8, 0.3)[i] + 0.05 * k
    s.n_boxes, s.n_spheres = (1, 1) if with_objects else (0, 0)
    for k, (a, b) in enumerate([(0.40, 0.62), (0.30, 0.50), (0.35, 0.55)]):
        s.box_min[0][k], s.box_max[0][k] = a, b
        s.box_albedo[0][k] = (0.7, 0.4, 0.2)[k]
    for k, c in enumerate((0.25, 0.12, 0.70)):
        s.sph_center[0][k] = c
        s.sph_albedo[0][k] = (0.3, 0.6, 0.9)[k]
    s.sph_radius[0] = 0.12
    ld = np.array([0.25, 1.0, 0.15])
    ld /= np.linalg.norm(ld)
    for k in range(3):
        s.light_dir[k] = ld[k]
    s.light_intensity = 2.0
    for k, c in enumerate((0.5, 0.8, 0.05)):
        s.cam_pos[k] = c
    s.cam_hor, s.cam_vert = 0.0, 0.6  # looking along +z, tilted down
    s.fovy, s.znear, s.zfar = math.radians(60.0), 0.01, 1000.0
    return s


def _slabs(o, d, mn, mx):
    """Slab test per ray (columns of d; o one point or one per ray): the
    entry and exit parameters, the entry axis, and the gap between the two
    largest entry parameters (small near a box edge)."""
    o = o if o.ndim == 2 else o[:, None]
    dd = np.where(d == 0, 1e-300, d)
    t0 = (np.asarray(mn, np.float64)[:, None] - o) / dd
    t1 = (np.asarray(mx, np.float64)[:, None] - o) / dd
    lo, hi = np.minimum(t0, t1), np.maximum(t0, t1)
    slo = np.sort(lo, 0)
    return lo.max(0), hi.min(0), lo.argmax(0), slo[2] - slo[1]


def analytic_gbuffer(lgx, s, w, h, margin=1e-3):
    """Float64 ray caster of the scene: per pixel the radial depth, world
    normal, albedo and direct radiance E max(0, n.l) vis (shadow rays test
    the boxes and spheres, not the room), plus a mask of the pixels that sit
    clear of every silhouette, edge and shadow boundary by `margin` in the
    ray parameters, where float32 rendering cannot decide differently."""
    view, proj = lgx.scene_matrices(s, w, h)
    V = np.asarray(view, np.float64).reshape(4, 4).T
    P = np.asarray(proj, np.float64).reshape(4, 4).T
    inv = np.linalg.inv(P @ V)
    cam = np.linalg.inv(V)[:3, 3]
    ys, xs = np.mgrid[0:h, 0:w]
    sx, sy = ((xs + 0.5) / w).ravel(), ((ys + 0.5) / h).ravel()
    v = inv @ np.stack([sx * 2 - 1, sy * 2 - 1, np.ones_like(sx), np.ones_like(sx)])
    d = v[:3] / v[3] - cam[:, None]
    d /= np.linalg.norm(d, axis=0)
    n_px = d.shape[1]
    idx = np.arange(n_px)
    # room: the inside of the unit cube
    safe = np.where(d == 0, 1.0, d)
    tw = np.where(d > 0, (1 - cam[:, None]) / safe, (0 - cam[:, None]) / safe)
    tw = np.where(d == 0, np.inf, tw)
    axis = tw.argmin(0)
    best = tw.min(0)
    srt = np.sort(tw, 0)
    clear = (srt[1] - srt[0]) > margin
    normal = np.zeros((3, n_px))
    normal[axis, idx] = np.where(d[axis, idx] > 0, -1.0, 1.0)
    wall = axis * 2 + (d[axis, idx] > 0)
    alb = np.array([[s.wall_albedo[i][k] for k in range(3)] for i in range(6)], np.float64)[wall].T.copy()
    boxes = [([s.box_min[b][k] for k in range(3)], [s.box_max[b][k] for k in range(3)],
              [s.box_albedo[b][k] for k in range(3)]) for b in range(s.n_boxes)]
    spheres = [(np.array([s.sph_center[i][k] for k in range(3)], np.float64), float(s.sph_radius[i]),
                [s.sph_albedo[i][k] for k in range(3)]) for i in range(s.n_spheres)]
    for mn, mx, ab in boxes:
        tn, tf, ax, gap = _slabs(cam, d, mn, mx)
        hit = (tn <= tf) & (tn > 1e-6) & (tn < best)
        clear &= np.abs(tf - tn) > margin                  # silhouette
        clear &= ~hit | (gap > margin)                    # box edge between two faces
        clear &= ~((tn <= tf) & (np.abs(tn - best) < margin))  # depth tie with what is behind
        best = np.where(hit, tn, best)
        nb = np.zeros((3, n_px))
        nb[ax, idx] = np.where(d[ax, idx] > 0, -1.0, 1.0)
        normal = np.where(hit, nb, normal)
        alb = np.where(hit, np.array(ab, np.float64)[:, None], alb)
    for c, r, ab in spheres:
        lv = cam - c
        b = lv @ d
        disc = b * b - (lv @ lv - r * r)
        clear &= np.abs(disc) > margin * r
        t0 = -b - np.sqrt(np.maximum(disc, 0))
        hit = (disc >= 0) & (t0 > 1e-6) & (t0 < best)
        clear &= ~((disc >= 0) & (np.abs(t0 - best) < margin))
        best = np.where(hit, t0, best)
        hp = cam[:, None] + d * best
        normal = np.where(hit, (hp - c[:, None]) / r, normal)
        alb = np.where(hit, np.array(ab, np.float64)[:, None], alb)
    l = np.array([s.light_dir[k] for k in range(3)], np.float64)
    ld = np.repeat(l[:, None], n_px, 1)
    o = cam[:, None] + d * best + normal * 1e-4
    ndl = (normal * l[:, None]).sum(0)
    vis = np.ones(n_px)
    lit_side = ndl > 0
    for mn, mx, _ in boxes:
        tn, tf, _, _ = _slabs(o, ld, mn, mx)
        vis = np.where((tn <= tf) & (tn > 1e-6), 0.0, vis)
        clear &= ~lit_side | (np.abs(tf - tn) > margin)
    for c, r, _ in spheres:
        lv = o - c[:, None]
        b = (lv * ld).sum(0)
        disc = b * b - ((lv * lv).sum(0) - r * r)
        t1 = -b + np.sqrt(np.maximum(disc, 0))
        vis = np.where((disc >= 0) & (t1 > 1e-6), 0.0, vis)
        clear &= ~lit_side | (np.abs(disc) > margin * r)
    clear &= np.abs(ndl) > margin
    e = float(s.light_intensity) * np.maximum(ndl, 0) * vis
    img = lambda a: np.moveaxis(a.reshape(-1, h, w), 0, -1)
    return dict(depth=best.reshape(h, w), normal=img(normal), albedo=img(alb), radiance=img(alb * e),
                clear=clear.reshape(h, w), lit=(e > 0).reshape(h, w), shadowed=(lit_side & (vis == 0)).reshape(h, w))


def decode_oct(o):
    """Octahedral decode in float64 (both hemispheres)."""
    o = np.asarray(o, np.float64)
    x, y = o[..., 0], o[..., 1]
    z = 1 - np.abs(x) - np.abs(y)
    fx = np.where(z < 0, (1 - np.abs(y)) * np.where(x >= 0, 1, -1), x)
    fy = np.where(z < 0, (1 - np.abs(x)) * np.where(y >= 0, 1, -1), y)
    v = np.stack([fx, fy, z], -1)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)
