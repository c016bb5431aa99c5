"""Device-memory plumbing for the lgx C ABI: torch tensors as HIP buffers.

Torch supplies allocations and the current stream; every pass runs in a
HIP kernel behind liblgx.so.  Raises when no HIP device is present.
"""
import ctypes as C

import numpy as np
import torch

from . import lgx

_TORCH_DTYPE = {
    lgx.R32F: (torch.float32, 1), lgx.RG32F: (torch.float32, 2), lgx.RGBA32F: (torch.float32, 4),
    lgx.RG16F: (torch.float16, 2), lgx.RGBA16F: (torch.float16, 4), lgx.RGBA8: (torch.uint8, 4),
    lgx.RGBA64F: (torch.float64, 4), lgx.R16F: (torch.float16, 1),
}


def require_device():
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device: the lgx passes run only on the GPU")
    lgx.lib()


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class DImage:
    """A device image: torch storage + its lgx_image descriptor."""

    def __init__(self, w, h, fmt, tensor=None, device="cuda"):
        dt, c = _TORCH_DTYPE[fmt]
        shape = (h, w) if c == 1 else (h, w, c)
        self.t = tensor if tensor is not None else torch.empty(shape, dtype=dt, device=device)
        assert tuple(self.t.shape) == shape and self.t.dtype == dt and self.t.is_contiguous()
        self.w, self.h, self.fmt = w, h, fmt

    @property
    def desc(self):
        return lgx.Image(C.c_void_p(self.t.data_ptr()), self.w, self.h, self.w, self.fmt, 0)

    @staticmethod
    def from_numpy(a, fmt):
        dt, c = _TORCH_DTYPE[fmt]
        t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
        h, w = a.shape[0], a.shape[1]
        if c == 1:
            t = t.reshape(h, w)
        return DImage(w, h, fmt, tensor=t.to(dt).contiguous())

    def numpy(self):
        return self.t.cpu().numpy()


class GBufferDev:
    def __init__(self, depth, normal, albedo, radiance):
        self.depth, self.normal, self.albedo, self.radiance = depth, normal, albedo, radiance
        self.w, self.h = depth.w, depth.h

    @property
    def desc(self):
        return lgx.GBuffer(self.depth.desc, self.normal.desc, self.albedo.desc, self.radiance.desc)

    @staticmethod
    def from_host(gb):
        f = gb["formats"]
        return GBufferDev(DImage.from_numpy(gb["depth"], lgx.R32F), DImage.from_numpy(gb["normal"], f["normal"]),
                          DImage.from_numpy(gb["albedo"], f["albedo"]),
                          DImage.from_numpy(gb["radiance"], f["radiance"]))

    @staticmethod
    def alloc(w, h, kind="mixed"):
        f = lgx.GBUF_FORMATS[kind]
        return GBufferDev(DImage(w, h, lgx.R32F), DImage(w, h, f["normal"]), DImage(w, h, f["albedo"]),
                          DImage(w, h, f["radiance"]))

    def render(self, scene, stream=None):
        g = self.desc
        lgx.check(lgx.lib().lgx_scene_render_gpu(C.byref(scene), C.byref(g), stream_handle(stream)),
                  "lgx_scene_render_gpu")
        return self


class PyramidDev:
    """Depth (R32F) and radiance pyramids plus their pre-blurred copies."""

    def __init__(self, gb, n_levels, sampler):
        self.n, self.sampler = n_levels, sampler
        sizes = lgx.level_sizes(gb.w, gb.h, n_levels)
        rf = gb.radiance.fmt
        self.rad = [gb.radiance] + [DImage(w, h, rf) for (w, h) in sizes[1:]]
        self.rad_b = [gb.radiance] + [DImage(w, h, rf) for (w, h) in sizes[1:]]  # r=0 at level 0 is a copy
        self.dep = [gb.depth] + [DImage(w, h, lgx.R32F) for (w, h) in sizes[1:]]
        if sampler == lgx.SAMPLER_TRILINEAR:
            self.dep_b = [gb.depth] + [DImage(w, h, lgx.R32F) for (w, h) in sizes[1:]]
        else:
            self.dep_b = self.dep  # max-Z Hi-Z is not blurred

    def build(self, stream=None):
        L, s = lgx.lib(), stream_handle(stream)
        n = self.n
        if n > 1:
            rl = (lgx.Image * n)(*[d.desc for d in self.rad])
            lgx.check(L.lgx_mip_build(rl, n, lgx.MIP_AVG, s), "lgx_mip_build radiance")
            dl = (lgx.Image * n)(*[d.desc for d in self.dep])
            mode = lgx.MIP_MAX if self.sampler == lgx.SAMPLER_HIZ_POINT else lgx.MIP_AVG
            lgx.check(L.lgx_mip_build(dl, n, mode, s), "lgx_mip_build depth")
            src = [d.desc for d in self.rad[1:]]
            dst = [d.desc for d in self.rad_b[1:]]
            if self.sampler == lgx.SAMPLER_TRILINEAR:
                src_d = [d.desc for d in self.dep[1:]]
                dst_d = [d.desc for d in self.dep_b[1:]]
                lgx.check(L.lgx_box_blur((lgx.Image * (n - 1))(*src_d), (lgx.Image * (n - 1))(*dst_d),
                                         (C.c_int32 * (n - 1))(*([2] * (n - 1))), n - 1, s), "lgx_box_blur depth")
            lgx.check(L.lgx_box_blur((lgx.Image * (n - 1))(*src), (lgx.Image * (n - 1))(*dst),
                                     (C.c_int32 * (n - 1))(*([2] * (n - 1))), n - 1, s), "lgx_box_blur radiance")
        return self

    @property
    def desc(self):
        p = lgx.Pyramid()
        p.n_levels = self.n
        for l in range(self.n):
            p.depth[l] = self.dep_b[l].desc
            p.radiance[l] = self.rad_b[l].desc
        return p


class Gather:
    """Prepared gather (workspace tables for one resolution / parameter set)."""

    def __init__(self, w, h, dirs, max_steps, sampler, grid=4, precision=lgx.PREC_F32, out_layout=lgx.OUT_INTERLEAVED,
                 tile=0, footprint_dirs=0):
        p = lgx.GatherParams()
        p.width, p.height, p.dirs, p.max_steps = w, h, dirs, max_steps
        p.grid, p.grid_y = lgx.pattern(grid)
        p.footprint_dirs = footprint_dirs
        p.sampler, p.precision, p.out_layout, p.tile = sampler, precision, out_layout, tile
        self.p = p
        nbytes = lgx.lib().lgx_gather_workspace_bytes(C.byref(p))
        if nbytes == 0:
            raise lgx.LgxError(lgx.EINVAL, "lgx_gather_workspace_bytes")
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        lgx.check(lgx.lib().lgx_gather_prepare(C.byref(p), C.c_void_p(self.ws.data_ptr())), "lgx_gather_prepare")
        self.out_fmt = lgx.RGBA32F if precision == lgx.PREC_F32 else lgx.RGBA64F

    def __call__(self, view, proj, gb, pyr, out=None, step_counts=None, stream=None, roi=None, hit_counts=None):
        p = self.p
        p.roi[:] = list(roi) if roi is not None else [0, 0, 0, 0]
        p.view[:] = [float(v) for v in view]
        p.proj[:] = [float(v) for v in proj]
        p.step_counts = C.c_void_p(step_counts.data_ptr()) if step_counts is not None else None
        p.hit_counts = C.c_void_p(hit_counts.data_ptr()) if hit_counts is not None else None
        if out is None:
            out = DImage(p.width, p.height, self.out_fmt)
        g, pd, o = gb.desc, pyr.desc, out.desc
        lgx.check(lgx.lib().lgx_gather(C.byref(p), C.c_void_p(self.ws.data_ptr()), C.byref(g), C.byref(pd),
                                       C.byref(o), stream_handle(stream)), "lgx_gather")
        return out


def autotune_gather(gb, pyr, view, proj, dirs, max_steps, sampler=lgx.SAMPLER_HIZ_POINT, footprint_dirs=0, reps=3,
                    grid=4):
    """Times every gather launch shape on this frame (hipEvents on the current
    stream) and returns (fastest tile, {tile: ms}); all shapes give
    bit-identical results."""
    out = DImage(gb.w, gb.h, lgx.RGBA32F)
    times = {}
    for t in range(lgx.lib().lgx_gather_tiles()):
        g = Gather(gb.w, gb.h, dirs, max_steps, sampler, grid=grid, tile=t, footprint_dirs=footprint_dirs)
        g(view, proj, gb, pyr, out=out)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            g(view, proj, gb, pyr, out=out)
        e1.record()
        e1.synchronize()
        times[t] = e0.elapsed_time(e1) / reps
    return min(times, key=times.get), times


def bilateral(inp, gb, out=None, radius=4, ds=10.5, ns=20.0, stream=None):
    if out is None:
        out = DImage(inp.w, inp.h, lgx.RGBA32F)
    bp = lgx.BilateralParams(radius, ds, ns, 0)
    i, g, o = inp.desc, gb.desc, out.desc
    lgx.check(lgx.lib().lgx_bilateral(C.byref(bp), C.byref(i), C.byref(g), C.byref(o), stream_handle(stream)),
              "lgx_bilateral")
    return out


def permute(inp, out=None, inverse=False, gx=4, gy=4, stream=None):
    if out is None:
        out = DImage(inp.w, inp.h, inp.fmt)
    i, o = inp.desc, out.desc
    fn = lgx.lib().lgx_interleave if inverse else lgx.lib().lgx_deinterleave
    lgx.check(fn(C.byref(i), C.byref(o), gx, gy, stream_handle(stream)), "lgx_(de)interleave")
    return out


class Renderer:
    """The SSVGI renderer on the C++ RenderGraph (lgx_renderer_*)."""

    def __init__(self, w, h, dirs=32, max_steps=64, levels=6, sampler=lgx.SAMPLER_HIZ_POINT, grid=4,
                 denoiser=lgx.DENOISER_BILATERAL, denoise_passes=2, denoise_radius=4, kind="mixed", tile=0,
                 composite=False, profile=False, validate=False, footprint_dirs=0):
        f = lgx.GBUF_FORMATS[kind]
        c = lgx.RendererConfig()
        c.width, c.height, c.dirs, c.max_steps, c.levels, c.sampler = w, h, dirs, max_steps, levels, sampler
        c.grid, c.grid_y = lgx.pattern(grid)
        c.denoiser, c.denoise_passes, c.denoise_radius = denoiser, denoise_passes, denoise_radius
        c.depth_sharpness, c.normal_sharpness = 10.5, 20.0
        c.normal_format, c.albedo_format, c.radiance_format = f["normal"], f["albedo"], f["radiance"]
        c.tile, c.composite, c.profile, c.validate = tile, int(composite), int(profile), int(validate)
        c.footprint_dirs = footprint_dirs
        self.cfg = c
        h_ = C.c_void_p()
        lgx.check(lgx.lib().lgx_renderer_create(C.byref(c), C.byref(h_)), "lgx_renderer_create")
        self.h = h_

    def _err(self, code, what):
        if code != lgx.OK:
            raise lgx.LgxError(code, f"{what}: {lgx.lib().lgx_renderer_last_error(self.h).decode()}")

    def render(self, gb, view, proj, stream=None):
        g = gb.desc
        v = (C.c_float * 16)(*[float(x) for x in view])
        p = (C.c_float * 16)(*[float(x) for x in proj])
        self._err(lgx.lib().lgx_renderer_render(self.h, C.byref(g), v, p, stream_handle(stream)), "render")

    def capture(self, gb, view, proj, stream=None):
        g = gb.desc
        v = (C.c_float * 16)(*[float(x) for x in view])
        p = (C.c_float * 16)(*[float(x) for x in proj])
        self._err(lgx.lib().lgx_renderer_capture(self.h, C.byref(g), v, p, stream_handle(stream)), "capture")

    def attach_unet(self, net):
        self._unet = net  # keep the network alive as long as the renderer
        self._err(lgx.lib().lgx_renderer_attach_unet(self.h, net.h), "attach_unet")

    def replay(self, stream=None):
        self._err(lgx.lib().lgx_renderer_replay(self.h, stream_handle(stream)), "replay")

    def intermediate(self, name, mip=0):
        """Copy of a named intermediate as a numpy array (synchronises)."""
        im = lgx.Image()
        self._err(lgx.lib().lgx_renderer_intermediate(self.h, name.encode(), mip, C.byref(im)), "intermediate")
        dst = DImage(im.w, im.h, im.format)
        d = dst.desc
        lgx.check(lgx.lib().lgx_copy_image(C.byref(im), C.byref(d), stream_handle()), "lgx_copy_image")
        torch.cuda.synchronize()
        return dst.numpy()

    def output(self):
        return self.intermediate("final" if self.cfg.composite else "denoised")

    def pass_times(self, mode=0):
        import json
        buf = C.create_string_buffer(1 << 16)
        self._err(lgx.lib().lgx_renderer_pass_times(self.h, mode, buf, len(buf)), "pass_times")
        return json.loads(buf.value.decode()) if mode != 2 else None

    def close(self):
        if self.h:
            lgx.lib().lgx_renderer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class UNet:
    """The BF16 MFMA U-Net denoiser (lgx_unet_*)."""

    def __init__(self, max_n, max_h, max_w, seed=0):
        h = C.c_void_p()
        lgx.check(lgx.lib().lgx_unet_create(max_n, max_h, max_w, C.c_uint64(seed), C.byref(h)), "lgx_unet_create")
        self.h = h
        self.n_params = int(lgx.lib().lgx_unet_param_count())

    def forward(self, x12, out=None, stream=None):
        n, h, w, _ = x12.shape
        if out is None:
            out = torch.empty((n, h, w, 4), dtype=torch.float32, device="cuda")
        lgx.check(lgx.lib().lgx_unet_forward(self.h, C.c_void_p(x12.data_ptr()), n, h, w, C.c_void_p(out.data_ptr()),
                                             stream_handle(stream)), "lgx_unet_forward")
        return out

    def backward(self, x12, target, loss=None, stream=None):
        n, h, w, _ = x12.shape
        lp = C.c_void_p(loss.data_ptr()) if loss is not None else None
        lgx.check(lgx.lib().lgx_unet_backward(self.h, C.c_void_p(x12.data_ptr()), C.c_void_p(target.data_ptr()), n, h,
                                              w, lp, stream_handle(stream)), "lgx_unet_backward")

    def set_loss(self, kind):
        """Training loss of later backward calls: "l1" (mean absolute error) or "l2" (mean squared error)."""
        code = {"l1": lgx.UNET_LOSS_L1, "l2": lgx.UNET_LOSS_L2}[kind]
        lgx.check(lgx.lib().lgx_unet_set_loss(self.h, code), "lgx_unet_set_loss")

    def adam_step(self, lr, stream=None):
        lgx.check(lgx.lib().lgx_unet_adam_step(self.h, C.c_float(lr), stream_handle(stream)), "lgx_unet_adam_step")

    def train_step(self, x12, target, lr, loss=None, stream=None):
        n, h, w, _ = x12.shape
        lp = C.c_void_p(loss.data_ptr()) if loss is not None else None
        lgx.check(lgx.lib().lgx_unet_train_step(self.h, C.c_void_p(x12.data_ptr()), C.c_void_p(target.data_ptr()), n,
                                                h, w, C.c_float(lr), lp, stream_handle(stream)), "lgx_unet_train_step")

    def tiles(self):
        f, d = (C.c_int32 * 10)(), (C.c_int32 * 10)()
        lgx.check(lgx.lib().lgx_unet_get_tiles(self.h, f, d), "lgx_unet_get_tiles")
        return list(f), list(d)

    def set_tiles(self, fwd, dx):
        lgx.check(lgx.lib().lgx_unet_set_tiles(self.h, (C.c_int32 * 10)(*fwd), (C.c_int32 * 10)(*dx)),
                  "lgx_unet_set_tiles")

    def autotune(self, n, h, w, backward=False, stream=None):
        """Picks each convolution's register tile at this shape (overwrites activations)."""
        import json
        buf = C.create_string_buffer(1 << 14)
        lgx.check(lgx.lib().lgx_unet_autotune(self.h, n, h, w, int(backward), stream_handle(stream), buf, len(buf)),
                  "lgx_unet_autotune")
        return json.loads(buf.value.decode())

    def get(self, which=0):
        a = np.zeros(self.n_params, np.float32)
        t = C.c_int32()
        lgx.check(lgx.lib().lgx_unet_get_state(self.h, which, C.c_void_p(a.ctypes.data), self.n_params, C.byref(t)),
                  "lgx_unet_get_state")
        return a, t.value

    def set(self, which, a, adam_t=-1):
        a = np.ascontiguousarray(a, np.float32)
        lgx.check(lgx.lib().lgx_unet_set_state(self.h, which, C.c_void_p(a.ctypes.data), self.n_params, adam_t),
                  "lgx_unet_set_state")

    def denoise(self, gb, noisy, out=None, stream=None):
        if out is None:
            out = DImage(noisy.w, noisy.h, lgx.RGBA32F)
        g, i, o = gb.desc, noisy.desc, out.desc
        lgx.check(lgx.lib().lgx_unet_denoise(self.h, C.byref(g), C.byref(i), C.byref(o), stream_handle(stream)),
                  "lgx_unet_denoise")
        return out

    def save(self, path, metadata=None):
        """Checkpoint: parameters and Adam state (safetensors, no code executed
        on load); metadata: optional dict of strings stored in the header."""
        from safetensors.numpy import save_file
        p, t = self.get(0)
        m, _ = self.get(2)
        v, _ = self.get(3)
        save_file({"params": p, "adam_m": m, "adam_v": v, "adam_t": np.array([t], np.int32)}, path,
                  metadata=metadata)

    def load(self, path):
        from safetensors.numpy import load_file
        d = load_file(path)
        if d["params"].size != self.n_params:
            raise ValueError(f"checkpoint has {d['params'].size} parameters, the network {self.n_params}")
        t = int(d["adam_t"][0]) if "adam_t" in d else 0
        self.set(0, d["params"], t)
        if "adam_m" in d:
            self.set(2, d["adam_m"], t)
            self.set(3, d["adam_v"], t)

    def close(self):
        if self.h:
            lgx.lib().lgx_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GatherBatch:
    """Several views' gathers in one persistent launch (lgx_gather_batch_*).
    items: (Gather, GBufferDev, PyramidDev, out DImage, view, proj) tuples."""

    def __init__(self, items):
        n = len(items)
        ps = (lgx.GatherParams * n)()
        self._keep = items
        for i, (gp, gb, pyr, out, view, proj) in enumerate(items):
            C.memmove(C.byref(ps[i]), C.byref(gp.p), C.sizeof(lgx.GatherParams))
            ps[i].view[:] = [float(v) for v in view]
            ps[i].proj[:] = [float(v) for v in proj]
            ps[i].roi[:] = [0, 0, 0, 0]
            ps[i].step_counts = None
            ps[i].hit_counts = None
        ws = (C.c_void_p * n)(*[it[0].ws.data_ptr() for it in items])
        gs = (lgx.GBuffer * n)(*[it[1].desc for it in items])
        pyrs = (lgx.Pyramid * n)(*[it[2].desc for it in items])
        outs = (lgx.Image * n)(*[it[3].desc for it in items])
        h = C.c_void_p()
        lgx.check(lgx.lib().lgx_gather_batch_create(ps, ws, gs, pyrs, outs, n, C.byref(h)), "lgx_gather_batch_create")
        self.h = h

    def set_tile(self, tile):
        lgx.check(lgx.lib().lgx_gather_batch_set_tile(self.h, tile), "lgx_gather_batch_set_tile")

    @property
    def blocks(self):
        return int(lgx.lib().lgx_gather_batch_blocks(self.h))

    def launch(self, workgroups=0, stream=None):
        lgx.check(lgx.lib().lgx_gather_batch_launch(self.h, workgroups, stream_handle(stream)),
                  "lgx_gather_batch_launch")

    def close(self):
        if self.h:
            lgx.lib().lgx_gather_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """The denoised path for views of uneven resolution (lgx_batch_*).
    views: (GBufferDev, view, proj) tuples; net: UNet or None."""

    def __init__(self, views, dirs=4, max_steps=64, levels=6, sampler=lgx.SAMPLER_HIZ_POINT, grid=4,
                 footprint_dirs=0, tile=0, workgroups=0, net=None, chunks=0):
        n = len(views)
        vs = (lgx.BatchView * n)()
        for i, (gb, view, proj) in enumerate(views):
            vs[i].width, vs[i].height = gb.w, gb.h
            vs[i].gbuffer = gb.desc
            vs[i].view[:] = [float(v) for v in view]
            vs[i].proj[:] = [float(v) for v in proj]
        c = lgx.BatchConfig()
        c.dirs, c.max_steps, c.levels, c.sampler = dirs, max_steps, levels, sampler
        c.grid, c.grid_y = lgx.pattern(grid)
        c.footprint_dirs, c.tile, c.workgroups, c.chunks = footprint_dirs, tile, workgroups, chunks
        self._keep = (views, net)
        self.n = n
        h = C.c_void_p()
        lgx.check(lgx.lib().lgx_batch_create(C.byref(c), vs, n, net.h if net is not None else None, C.byref(h)),
                  "lgx_batch_create")
        self.h = h

    def _err(self, code, what):
        if code != lgx.OK:
            raise lgx.LgxError(code, f"{what}: {lgx.lib().lgx_batch_last_error(self.h).decode()}")

    def autotune(self, stream=None):
        import json
        buf = C.create_string_buffer(1 << 14)
        self._err(lgx.lib().lgx_batch_autotune(self.h, stream_handle(stream), buf, len(buf)), "autotune")
        return json.loads(buf.value.decode())

    def run(self, stream=None):
        self._err(lgx.lib().lgx_batch_run(self.h, stream_handle(stream)), "run")

    def capture(self, stream):
        self._err(lgx.lib().lgx_batch_capture(self.h, stream_handle(stream)), "capture")

    def replay(self, stream=None):
        self._err(lgx.lib().lgx_batch_replay(self.h, stream_handle(stream)), "replay")

    def latencies(self):
        a = (C.c_double * self.n)()
        self._err(lgx.lib().lgx_batch_latencies(self.h, a, self.n), "latencies")
        return list(a)

    def shape(self):
        t, w = C.c_int32(), C.c_int32()
        self._err(lgx.lib().lgx_batch_shape(self.h, C.byref(t), C.byref(w)), "shape")
        return t.value, w.value

    def output(self, i, which=1):
        """Copy of view i's noisy (0) or denoised (1) image as numpy (synchronises)."""
        im = lgx.Image()
        self._err(lgx.lib().lgx_batch_output(self.h, i, which, C.byref(im)), "output")
        dst = DImage(im.w, im.h, im.format)
        d = dst.desc
        lgx.check(lgx.lib().lgx_copy_image(C.byref(im), C.byref(d), stream_handle()), "lgx_copy_image")
        torch.cuda.synchronize()
        return dst.numpy()

    def close(self):
        if self.h:
            lgx.lib().lgx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mfma_probe(A, B):
    a = torch.from_numpy(np.ascontiguousarray(A, np.float32)).cuda()
    b = torch.from_numpy(np.ascontiguousarray(B, np.float32)).cuda()
    d = torch.zeros((16, 16), dtype=torch.float32, device="cuda")
    lgx.check(lgx.lib().lgx_unet_mfma_probe(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()),
                                            C.c_void_p(d.data_ptr()), stream_handle()), "lgx_unet_mfma_probe")
    torch.cuda.synchronize()
    return d.cpu().numpy()
