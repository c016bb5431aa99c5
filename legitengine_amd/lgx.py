"""ctypes binding of liblgx.so (the C ABI declared in include/lgx.h).

This is plumbing for tests and benchmarks: torch tensors supply device
memory and streams, every computation runs in the HIP kernels behind the C
ABI.  Loading fails loudly when the library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblgx.so")

LGX_MAX_LEVELS = 10
R32F, RG32F, RGBA32F, RG16F, RGBA16F, RGBA8, RGBA64F, R16F = 1, 2, 3, 4, 5, 6, 7, 8
FORMAT_BYTES = {R32F: 4, RG32F: 8, RGBA32F: 16, RG16F: 4, RGBA16F: 8, RGBA8: 4, RGBA64F: 32, R16F: 2}
MIP_AVG, MIP_MAX = 0, 1
SAMPLER_TRILINEAR, SAMPLER_HIZ_POINT = 0, 1
PREC_F32, PREC_F64, PREC_F32_ACC64 = 0, 1, 2
OUT_INTERLEAVED, OUT_SLICES = 0, 1
UNET_LOSS_L1, UNET_LOSS_L2 = 0, 1
OK, EINVAL, EFORMAT, ESIZE, EUNSUPPORTED, EWORKSPACE, ENODEVICE = 0, -1, -2, -3, -4, -5, -6


def pattern(grid):
    """Interleave pattern (width, height) from an int (square) or a pair."""
    return (int(grid), int(grid)) if isinstance(grid, int) else (int(grid[0]), int(grid[1]))


class Image(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("w", C.c_int32), ("h", C.c_int32), ("pitch", C.c_int32),
                ("format", C.c_int32), ("reserved", C.c_int32)]


class GBuffer(C.Structure):
    _fields_ = [("depth", Image), ("normal", Image), ("albedo", Image), ("radiance", Image)]


class Pyramid(C.Structure):
    _fields_ = [("n_levels", C.c_int32), ("depth", Image * LGX_MAX_LEVELS), ("radiance", Image * LGX_MAX_LEVELS)]


class GatherParams(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("proj", C.c_float * 16), ("width", C.c_int32), ("height", C.c_int32),
                ("dirs", C.c_int32), ("max_steps", C.c_int32), ("grid", C.c_int32), ("sampler", C.c_int32),
                ("precision", C.c_int32), ("out_layout", C.c_int32), ("step_counts", C.c_void_p),
                ("tile", C.c_int32), ("roi", C.c_int32 * 4), ("footprint_dirs", C.c_int32), ("grid_y", C.c_int32),
                ("hit_counts", C.c_void_p)]


class BilateralParams(C.Structure):
    _fields_ = [("radius", C.c_int32), ("depth_sharpness", C.c_float), ("normal_sharpness", C.c_float),
                ("reserved", C.c_int32)]


class Scene(C.Structure):
    _fields_ = [("wall_albedo", (C.c_float * 3) * 6), ("n_boxes", C.c_int32), ("n_spheres", C.c_int32),
                ("box_min", (C.c_float * 3) * 32), ("box_max", (C.c_float * 3) * 32),
                ("box_albedo", (C.c_float * 3) * 32), ("sph_center", (C.c_float * 3) * 8),
                ("sph_radius", C.c_float * 8), ("sph_albedo", (C.c_float * 3) * 8), ("light_dir", C.c_float * 3),
                ("light_intensity", C.c_float), ("cam_pos", C.c_float * 3), ("cam_hor", C.c_float),
                ("cam_vert", C.c_float), ("fovy", C.c_float), ("znear", C.c_float), ("zfar", C.c_float)]


class RendererConfig(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("dirs", C.c_int32), ("max_steps", C.c_int32),
                ("levels", C.c_int32), ("grid", C.c_int32), ("sampler", C.c_int32), ("denoiser", C.c_int32),
                ("denoise_passes", C.c_int32), ("denoise_radius", C.c_int32), ("depth_sharpness", C.c_float),
                ("normal_sharpness", C.c_float), ("normal_format", C.c_int32), ("albedo_format", C.c_int32),
                ("radiance_format", C.c_int32), ("tile", C.c_int32), ("composite", C.c_int32),
                ("profile", C.c_int32), ("validate", C.c_int32), ("footprint_dirs", C.c_int32),
                ("grid_y", C.c_int32)]


DENOISER_NONE, DENOISER_BILATERAL, DENOISER_REGRESSION, DENOISER_UNET = 0, 1, 2, 3


class BatchView(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("gbuffer", GBuffer), ("view", C.c_float * 16),
                ("proj", C.c_float * 16)]


class BatchConfig(C.Structure):
    _fields_ = [("dirs", C.c_int32), ("max_steps", C.c_int32), ("levels", C.c_int32), ("grid", C.c_int32),
                ("sampler", C.c_int32), ("footprint_dirs", C.c_int32), ("tile", C.c_int32),
                ("workgroups", C.c_int32), ("grid_y", C.c_int32), ("chunks", C.c_int32)]

# Every entry point declared in include/lgx.h, with its ctypes signature.
SIGNATURES = {
    "lgx_gather_workspace_bytes": (C.c_size_t, [C.POINTER(GatherParams)]),
    "lgx_gather_prepare": (C.c_int, [C.POINTER(GatherParams), C.c_void_p]),
    "lgx_gather": (C.c_int, [C.POINTER(GatherParams), C.c_void_p, C.POINTER(GBuffer), C.POINTER(Pyramid),
                             C.POINTER(Image), C.c_void_p]),
    "lgx_gather_tiles": (C.c_int, []),
    "lgx_gather_tile_info": (C.c_int, [C.c_int32, C.POINTER(C.c_int32)]),
    "lgx_gather_noalloc": (C.c_int, [C.POINTER(GatherParams), C.POINTER(GBuffer), C.POINTER(Pyramid),
                                     C.POINTER(Image), C.c_void_p]),
    "lgx_mip_build": (C.c_int, [C.POINTER(Image), C.c_int32, C.c_int32, C.c_void_p]),
    "lgx_box_blur": (C.c_int, [C.POINTER(Image), C.POINTER(Image), C.POINTER(C.c_int32), C.c_int32, C.c_void_p]),
    "lgx_deinterleave": (C.c_int, [C.POINTER(Image), C.POINTER(Image), C.c_int32, C.c_int32, C.c_void_p]),
    "lgx_interleave": (C.c_int, [C.POINTER(Image), C.POINTER(Image), C.c_int32, C.c_int32, C.c_void_p]),
    "lgx_bilateral": (C.c_int, [C.POINTER(BilateralParams), C.POINTER(Image), C.POINTER(GBuffer),
                                C.POINTER(Image), C.c_void_p]),
    "lgx_regression_denoise": (C.c_int, [C.POINTER(Image), C.POINTER(GBuffer), C.c_int32, C.POINTER(Image),
                                         C.c_void_p]),
    "lgx_composite": (C.c_int, [C.POINTER(GBuffer), C.POINTER(Image), C.POINTER(Image), C.c_void_p]),
    "lgx_scene_make": (C.c_int, [C.c_uint64, C.c_int32, C.c_int32, C.POINTER(Scene)]),
    "lgx_scene_matrices": (C.c_int, [C.POINTER(Scene), C.c_int32, C.c_int32, C.POINTER(C.c_float),
                                     C.POINTER(C.c_float)]),
    "lgx_scene_render_cpu": (C.c_int, [C.POINTER(Scene), C.POINTER(GBuffer)]),
    "lgx_scene_pose": (C.c_int, [C.POINTER(Scene), C.c_uint64]),
    "lgx_scene_render_gpu": (C.c_int, [C.POINTER(Scene), C.POINTER(GBuffer), C.c_void_p]),
    "lgx_renderer_create": (C.c_int, [C.POINTER(RendererConfig), C.POINTER(C.c_void_p)]),
    "lgx_renderer_render": (C.c_int, [C.c_void_p, C.POINTER(GBuffer), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                      C.c_void_p]),
    "lgx_renderer_capture": (C.c_int, [C.c_void_p, C.POINTER(GBuffer), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                       C.c_void_p]),
    "lgx_renderer_replay": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lgx_renderer_attach_unet": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lgx_renderer_intermediate": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.POINTER(Image)]),
    "lgx_renderer_pass_times": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, C.c_size_t]),
    "lgx_renderer_last_error": (C.c_char_p, [C.c_void_p]),
    "lgx_renderer_destroy": (None, [C.c_void_p]),
    "lgx_copy_image": (C.c_int, [C.POINTER(Image), C.POINTER(Image), C.c_void_p]),
    "lgx_unet_param_count": (C.c_int64, []),
    "lgx_unet_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "lgx_unet_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]),
    "lgx_unet_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                    C.c_void_p]),
    "lgx_unet_adam_step": (C.c_int, [C.c_void_p, C.c_float, C.c_void_p]),
    "lgx_unet_set_loss": (C.c_int, [C.c_void_p, C.c_int32]),
    "lgx_unet_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                      C.c_void_p, C.c_void_p]),
    "lgx_unet_output": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "lgx_unet_get_state": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.POINTER(C.c_int32)]),
    "lgx_unet_set_state": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_int32]),
    "lgx_unet_denoise": (C.c_int, [C.c_void_p, C.POINTER(GBuffer), C.POINTER(Image), C.POINTER(Image), C.c_void_p]),
    "lgx_unet_conv_tiles": (C.c_int, [C.c_int32]),
    "lgx_unet_set_tiles": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "lgx_unet_get_tiles": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "lgx_unet_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_char_p,
                                    C.c_size_t]),
    "lgx_unet_mfma_probe": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "lgx_unet_destroy": (None, [C.c_void_p]),
    "lgx_gather_batch_create": (C.c_int, [C.POINTER(GatherParams), C.POINTER(C.c_void_p), C.POINTER(GBuffer),
                                          C.POINTER(Pyramid), C.POINTER(Image), C.c_int32, C.POINTER(C.c_void_p)]),
    "lgx_gather_batch_set_tile": (C.c_int, [C.c_void_p, C.c_int32]),
    "lgx_gather_batch_blocks": (C.c_int64, [C.c_void_p]),
    "lgx_gather_batch_launch": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "lgx_gather_batch_destroy": (None, [C.c_void_p]),
    "lgx_batch_create": (C.c_int, [C.POINTER(BatchConfig), C.POINTER(BatchView), C.c_int32, C.c_void_p,
                                   C.POINTER(C.c_void_p)]),
    "lgx_batch_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_char_p, C.c_size_t]),
    "lgx_batch_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lgx_batch_capture": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lgx_batch_replay": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lgx_batch_latencies": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int32]),
    "lgx_batch_output": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(Image)]),
    "lgx_batch_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "lgx_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "lgx_batch_destroy": (None, [C.c_void_p]),
    "lgx_strerror": (C.c_char_p, [C.c_int]),
    "lgx_version": (C.c_int32, []),
}

_lib = None


def lib():
    """Loads liblgx.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (or legitengine_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class LgxError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: lgx error {code} ({lib().lgx_strerror(code).decode()})")
        self.code = code


def check(code, what="lgx call"):
    if code != OK:
        raise LgxError(code, what)
    return code


# ---------------------------------------------------------------- helpers

def image(ptr, w, h, fmt, pitch=None):
    return Image(C.c_void_p(ptr), w, h, pitch if pitch is not None else w, fmt, 0)


def image_np(arr, fmt):
    """Descriptor over a host numpy array (host entry points only)."""
    h, w = arr.shape[0], arr.shape[1]
    return Image(C.c_void_p(arr.ctypes.data), w, h, w, fmt, 0)


def image_t(t, fmt):
    """Descriptor over a contiguous torch tensor of shape (H, W, ...) in device memory."""
    assert t.is_contiguous()
    return Image(C.c_void_p(t.data_ptr()), t.shape[1], t.shape[0], t.shape[1], fmt, 0)


def make_scene(seed, n_boxes, n_spheres):
    s = Scene()
    check(lib().lgx_scene_make(C.c_uint64(seed), n_boxes, n_spheres, C.byref(s)), "lgx_scene_make")
    return s


def scene_pose(scene, pose_seed):
    """Re-samples the camera of `scene` (same geometry) from pose_seed, in place."""
    check(lib().lgx_scene_pose(C.byref(scene), C.c_uint64(pose_seed)), "lgx_scene_pose")
    return scene


def scene_matrices(scene, w, h):
    view = (C.c_float * 16)()
    proj = (C.c_float * 16)()
    check(lib().lgx_scene_matrices(C.byref(scene), w, h, view, proj), "lgx_scene_matrices")
    return np.array(view, dtype=np.float32), np.array(proj, dtype=np.float32)


GBUF_FORMATS = {
    # config 1: all FP32 (albedo/radiance padded to RGBA)
    "f32": dict(normal=RG32F, albedo=RGBA32F, radiance=RGBA32F),
    # config 2: normals and radiance FP16, albedo RGBA8
    "mixed": dict(normal=RG16F, albedo=RGBA8, radiance=RGBA16F),
}

_NP_DTYPE = {R32F: (np.float32, 1), RG32F: (np.float32, 2), RGBA32F: (np.float32, 4), RG16F: (np.float16, 2),
             RGBA16F: (np.float16, 4), RGBA8: (np.uint8, 4), RGBA64F: (np.float64, 4)}


def np_image(w, h, fmt):
    dt, c = _NP_DTYPE[fmt]
    return np.zeros((h, w, c), dtype=dt)


def render_gbuffer_cpu(scene, w, h, kind="mixed"):
    """Renders the G-buffer on host threads; returns numpy arrays in storage formats."""
    f = GBUF_FORMATS[kind]
    arrs = dict(depth=np_image(w, h, R32F), normal=np_image(w, h, f["normal"]),
                albedo=np_image(w, h, f["albedo"]), radiance=np_image(w, h, f["radiance"]))
    g = GBuffer(image_np(arrs["depth"], R32F), image_np(arrs["normal"], f["normal"]),
                image_np(arrs["albedo"], f["albedo"]), image_np(arrs["radiance"], f["radiance"]))
    check(lib().lgx_scene_render_cpu(C.byref(scene), C.byref(g)), "lgx_scene_render_cpu")
    arrs["depth"] = arrs["depth"][..., 0]
    arrs["formats"] = dict(depth=R32F, **f)
    return arrs


def level_sizes(w, h, n):
    out = []
    for _ in range(n):
        out.append((w, h))
        w //= 2
        h //= 2
    return out
