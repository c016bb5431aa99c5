"""ctypes binding of the CPU oracle (test infrastructure only).

Imported only by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg.  Wraps liboracle.so (oracle/oracle.cpp) with numpy-level
helpers that widen storage formats exactly (FP16 -> float32, RGBA8 / 255)
before the restated passes run.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


class GatherParams(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("proj", C.c_float * 16), ("width", C.c_int32), ("height", C.c_int32),
                ("dirs", C.c_int32), ("max_steps", C.c_int32), ("grid", C.c_int32), ("sampler", C.c_int32),
                ("n_levels", C.c_int32), ("footprint_dirs", C.c_int32), ("grid_y", C.c_int32)]


FP = C.POINTER(C.c_float)
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int32)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_gather_f32.argtypes = [C.POINTER(GatherParams), FP, FP, C.POINTER(FP), C.POINTER(FP), IP, IP, FP, IP,
                                     C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        L.orc_gather_f64.argtypes = [C.POINTER(GatherParams), FP, FP, C.POINTER(FP), C.POINTER(FP), IP, IP, DP, IP,
                                     C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        L.orc_gather_rows_f32.argtypes = [C.POINTER(GatherParams), FP, FP, C.POINTER(FP), C.POINTER(FP), IP, IP, IP,
                                          C.c_int32, FP, IP, IP, C.c_int32]
        L.orc_mip_avg.argtypes = [FP, C.c_int32, C.c_int32, C.c_int32, FP, C.c_int32]
        L.orc_mip_max.argtypes = [FP, C.c_int32, C.c_int32, FP]
        L.orc_box_blur.argtypes = [FP, C.c_int32, C.c_int32, C.c_int32, C.c_int32, FP, C.c_int32]
        L.orc_deinterleave.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
        L.orc_interleave.argtypes = L.orc_deinterleave.argtypes
        L.orc_bilateral.argtypes = [FP, FP, FP, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float, FP, C.c_int32]
        L.orc_regression_denoise.argtypes = [FP, FP, C.c_int32, C.c_int32, C.c_int32, FP]
        L.orc_composite.argtypes = [FP, FP, FP, C.c_int32, C.c_int32, FP]
        L.orc_hammersley.argtypes = [C.c_int32, C.c_int32, FP]
        L.orc_mat4_inverse.argtypes = [FP, FP]
        L.orc_mat4_mul.argtypes = [FP, FP, FP]
        L.orc_perspective_rh_no.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, FP]
        L.orc_camera_transform.argtypes = [FP, C.c_float, C.c_float, FP]
        L.orc_horizon_contribution.argtypes = [FP, FP, FP, C.c_float, C.c_float]
        L.orc_horizon_contribution.restype = C.c_float
        L.orc_horizon_contribution_f64.argtypes = [DP, DP, DP, C.c_double, C.c_double]
        L.orc_horizon_contribution_f64.restype = C.c_double
        L.orc_oct_decode.argtypes = [C.c_float, C.c_float, FP]
        L.orc_f32_to_f16.argtypes = [C.c_float]
        L.orc_f32_to_f16.restype = C.c_uint16
        L.orc_f16_to_f32.argtypes = [C.c_uint16]
        L.orc_f16_to_f32.restype = C.c_float
        L.orc_sample_trilinear.argtypes = [C.POINTER(FP), IP, IP, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                           C.c_float, FP]
        L.orc_sample_point.argtypes = L.orc_sample_trilinear.argtypes
        L.orc_max_threads.restype = C.c_int32
        _lib = L
    return _lib


def fp(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(FP)


def widen(a):
    """Storage format -> float32 exactly (FP16 exact, RGBA8 as v/255 in float32)."""
    if a.dtype == np.float16:
        return np.ascontiguousarray(a.astype(np.float32))
    if a.dtype == np.uint8:
        return np.ascontiguousarray((a.astype(np.float32) * np.float32(1.0 / 255.0)).astype(np.float32))
    return np.ascontiguousarray(a.astype(np.float32))


def mip_avg(src, round_fp16=False):
    h, w = src.shape[:2]
    c = 1 if src.ndim == 2 else src.shape[2]
    dst = np.zeros((h // 2, w // 2) + (() if src.ndim == 2 else (c,)), np.float32)
    assert lib().orc_mip_avg(fp(np.ascontiguousarray(src, np.float32)), w, h, c, fp(dst), int(round_fp16)) == 0
    return dst


def mip_max(src):
    h, w = src.shape
    dst = np.zeros((h // 2, w // 2), np.float32)
    assert lib().orc_mip_max(fp(np.ascontiguousarray(src, np.float32)), w, h, fp(dst)) == 0
    return dst


def box_blur(src, radius, round_fp16=False):
    h, w = src.shape[:2]
    c = 1 if src.ndim == 2 else src.shape[2]
    dst = np.zeros_like(src, dtype=np.float32)
    assert lib().orc_box_blur(fp(np.ascontiguousarray(src, np.float32)), w, h, c, radius, fp(dst),
                              int(round_fp16)) == 0
    return dst


def deinterleave(src, gx=4, gy=4):
    src = np.ascontiguousarray(src)
    h, w = src.shape[:2]
    eb = src.itemsize * (int(np.prod(src.shape[2:])) if src.ndim > 2 else 1)
    dst = np.zeros_like(src)
    assert lib().orc_deinterleave(src.ctypes.data, w, h, eb, gx, gy, dst.ctypes.data) == 0
    return dst


def interleave(src, gx=4, gy=4):
    src = np.ascontiguousarray(src)
    h, w = src.shape[:2]
    eb = src.itemsize * (int(np.prod(src.shape[2:])) if src.ndim > 2 else 1)
    dst = np.zeros_like(src)
    assert lib().orc_interleave(src.ctypes.data, w, h, eb, gx, gy, dst.ctypes.data) == 0
    return dst


def build_pyramids(depth, radiance, n_levels, sampler, radiance_fp16):
    """Reference pass order (SSVGIRenderer.h:207-221): avg mips of radiance and
    depth, then box pre-blur per level (r=0 at level 0, r=2 above).  In Hi-Z
    mode the depth pyramid is the unblurred max-Z chain."""
    rad = [np.ascontiguousarray(radiance, np.float32)]
    for _ in range(1, n_levels):
        rad.append(mip_avg(rad[-1], round_fp16=radiance_fp16))
    rad_b = [box_blur(r, 0 if i == 0 else 2, round_fp16=radiance_fp16) for i, r in enumerate(rad)]
    dep = [np.ascontiguousarray(depth, np.float32)]
    if sampler == 1:
        for _ in range(1, n_levels):
            dep.append(mip_max(dep[-1]))
        dep_b = dep
    else:
        for _ in range(1, n_levels):
            dep.append(mip_avg(dep[-1]))
        dep_b = [box_blur(d, 0 if i == 0 else 2) for i, d in enumerate(dep)]
    return dep_b, rad_b, dep, rad


def gather(view, proj, depth, oct_normal, pyr_depth, pyr_rad, dirs, max_steps, sampler, grid=4,
           precision="f32", crop=None, threads=0, with_steps=False, footprint_dirs=0, grid_y=0):
    h, w = depth.shape
    p = GatherParams()
    p.footprint_dirs = footprint_dirs
    p.view[:] = [float(v) for v in view]
    p.proj[:] = [float(v) for v in proj]
    gx, gy = (grid, grid_y) if isinstance(grid, int) else grid
    p.width, p.height, p.dirs, p.max_steps, p.grid, p.grid_y, p.sampler = w, h, dirs, max_steps, gx, gy, sampler
    p.n_levels = len(pyr_depth)
    pd = (FP * len(pyr_depth))(*[fp(a) for a in pyr_depth])
    pr = (FP * len(pyr_rad))(*[fp(a) for a in pyr_rad])
    lw = (C.c_int32 * len(pyr_depth))(*[a.shape[1] for a in pyr_depth])
    lh = (C.c_int32 * len(pyr_depth))(*[a.shape[0] for a in pyr_depth])
    x0, y0, cw, ch = crop if crop else (0, 0, w, h)
    steps = np.zeros((ch, cw), np.int32)
    dep = np.ascontiguousarray(depth, np.float32)
    octn = np.ascontiguousarray(oct_normal, np.float32)
    if precision == "f32":
        out = np.zeros((ch, cw, 4), np.float32)
        rc = lib().orc_gather_f32(C.byref(p), fp(dep), fp(octn), pd, pr, lw, lh, fp(out), steps.ctypes.data_as(IP),
                                  x0, y0, cw, ch, threads)
    else:
        out = np.zeros((ch, cw, 4), np.float64)
        rc = lib().orc_gather_f64(C.byref(p), fp(dep), fp(octn), pd, pr, lw, lh, out.ctypes.data_as(DP),
                                  steps.ctypes.data_as(IP), x0, y0, cw, ch, threads)
    assert rc == 0, rc
    return (out, steps) if with_steps else out


def gather_rows(view, proj, depth, oct_normal, pyr_depth, pyr_rad, dirs, max_steps, sampler, rows, grid=4,
                threads=0, footprint_dirs=0, grid_y=0, with_hits=False):
    """FP32 gather of whole rows (a stratified sample of the frame): returns
    (out [len(rows), W, 4], steps [len(rows), W]) and, with_hits, the
    per-pixel count of steps that fetched radiance."""
    h, w = depth.shape
    p = GatherParams()
    p.footprint_dirs = footprint_dirs
    p.view[:] = [float(v) for v in view]
    p.proj[:] = [float(v) for v in proj]
    gx, gy = (grid, grid_y) if isinstance(grid, int) else grid
    p.width, p.height, p.dirs, p.max_steps, p.grid, p.grid_y, p.sampler = w, h, dirs, max_steps, gx, gy, sampler
    p.n_levels = len(pyr_depth)
    pd = (FP * len(pyr_depth))(*[fp(a) for a in pyr_depth])
    pr = (FP * len(pyr_rad))(*[fp(a) for a in pyr_rad])
    lw = (C.c_int32 * len(pyr_depth))(*[a.shape[1] for a in pyr_depth])
    lh = (C.c_int32 * len(pyr_depth))(*[a.shape[0] for a in pyr_depth])
    r = np.ascontiguousarray(rows, np.int32)
    out = np.zeros((len(r), w, 4), np.float32)
    steps = np.zeros((len(r), w), np.int32)
    hits = np.zeros((len(r), w), np.int32)
    rc = lib().orc_gather_rows_f32(C.byref(p), fp(np.ascontiguousarray(depth, np.float32)),
                                   fp(np.ascontiguousarray(oct_normal, np.float32)), pd, pr, lw, lh,
                                   r.ctypes.data_as(IP), len(r), fp(out), steps.ctypes.data_as(IP),
                                   hits.ctypes.data_as(IP), threads)
    assert rc == 0, rc
    return (out, steps, hits) if with_hits else (out, steps)


def bilateral(inp, depth, oct_normal, radius=4, ds=10.5, ns=20.0, threads=0):
    h, w = depth.shape
    out = np.zeros((h, w, 4), np.float32)
    assert lib().orc_bilateral(fp(np.ascontiguousarray(inp, np.float32)), fp(np.ascontiguousarray(depth, np.float32)),
                               fp(np.ascontiguousarray(oct_normal, np.float32)), w, h, radius, ds, ns, fp(out),
                               threads) == 0
    return out


def regression_denoise(inp, depth, radius=2):
    h, w = depth.shape
    out = np.zeros((h, w, 4), np.float32)
    assert lib().orc_regression_denoise(fp(np.ascontiguousarray(inp, np.float32)),
                                        fp(np.ascontiguousarray(depth, np.float32)), w, h, radius, fp(out)) == 0
    return out


def composite(direct, albedo, indirect):
    h, w = direct.shape[:2]
    out = np.zeros((h, w, 4), np.float32)
    assert lib().orc_composite(fp(np.ascontiguousarray(direct, np.float32)), fp(np.ascontiguousarray(albedo, np.float32)),
                               fp(np.ascontiguousarray(indirect, np.float32)), w, h, fp(out)) == 0
    return out


def rel_l2(x, y):
    x = np.asarray(x, np.float64)[..., :3]
    y = np.asarray(y, np.float64)[..., :3]
    den = np.sqrt((y * y).sum())
    return float(np.sqrt(((x - y) ** 2).sum()) / den) if den > 0 else float(np.sqrt(((x - y) ** 2).sum()))


def plain_path(gb, view, proj, dirs, max_steps, n_levels, sampler, passes=2, radius=4, threads=0,
               with_steps=False):
    """The whole plain path on host: pyramids, pre-blur, gather, bilateral passes."""
    rad_fp16 = gb["radiance"].dtype == np.float16
    depth = np.ascontiguousarray(gb["depth"], np.float32)
    octn = widen(gb["normal"])
    radiance = widen(gb["radiance"])
    dep_b, rad_b, _, _ = build_pyramids(depth, radiance, n_levels, sampler, rad_fp16)
    g = gather(view, proj, depth, octn, dep_b, rad_b, dirs, max_steps, sampler, threads=threads,
               with_steps=with_steps)
    steps = None
    if with_steps:
        g, steps = g
    out = g
    for _ in range(passes):
        out = bilateral(out, depth, octn, radius, threads=threads)
    return dict(gather=g, out=out, steps=steps, pyr_depth=dep_b, pyr_rad=rad_b)
