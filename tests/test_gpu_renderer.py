"""The SSVGI renderer on the C++ RenderGraph (lgx_renderer_*) against the
same passes called one by one through the C ABI, and against the oracle.

The renderer runs the identical kernels on identical inputs, so its images
must be bit-identical to the pass-by-pass results; captured-graph replays
must reproduce the eager frame on every replay.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from legitengine_amd import lgx  # noqa: E402
from oracle import binding as orc  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from legitengine_amd import device
    device.require_device()
    return device


def frame_inputs(dev, seed, w, h, kind, boxes=8, spheres=2):
    sc = lgx.make_scene(seed, boxes, spheres)
    view, proj = lgx.scene_matrices(sc, w, h)
    gb = lgx.render_gbuffer_cpu(sc, w, h, kind)
    return view, proj, gb, dev.GBufferDev.from_host(gb)


def raw_path(dev, g, view, proj, levels, dirs, steps, sampler, passes=2, radius=4):
    pyr = dev.PyramidDev(g, levels, sampler).build()
    sl = dev.Gather(g.w, g.h, dirs, steps, sampler, out_layout=lgx.OUT_SLICES)(view, proj, g, pyr)
    ind = dev.permute(sl, inverse=True)
    out = ind
    for _ in range(passes):
        out = dev.bilateral(out, g, radius=radius)
    torch.cuda.synchronize()
    return pyr, ind.numpy(), out.numpy()


@pytest.mark.parametrize("sampler", [lgx.SAMPLER_HIZ_POINT, lgx.SAMPLER_TRILINEAR])
@pytest.mark.parametrize("kind,w,h,levels,dirs,steps", [("f32", 256, 256, 5, 16, 32), ("mixed", 640, 360, 6, 32, 64)])
def test_renderer_matches_pass_by_pass(dev, sampler, kind, w, h, levels, dirs, steps):
    view, proj, gb, g = frame_inputs(dev, 21, w, h, kind)
    r = dev.Renderer(w, h, dirs, steps, levels, sampler, kind=kind, validate=True, profile=True)
    r.render(g, view, proj)
    torch.cuda.synchronize()
    pyr, ind, out = raw_path(dev, g, view, proj, levels, dirs, steps, sampler)
    assert r.intermediate("indirect").tobytes() == ind.tobytes()
    assert r.output().tobytes() == out.tobytes()
    for l in range(1, levels):
        assert r.intermediate("depth_pyramid", l).tobytes() == pyr.dep[l].numpy().tobytes()
        assert r.intermediate("blurred_radiance", l).tobytes() == pyr.rad_b[l].numpy().tobytes()
    names = [p["name"] for p in r.pass_times(0)]
    assert names[:2] == ["MipBuilderPass", "MipBuilderPass"]
    assert "IndirectLightPass" in names and names.count("DenoiserPass") == 2 and "InterleavePass" in names
    if kind == "f32":
        ref = orc.plain_path(gb, view, proj, dirs, steps, levels, sampler, passes=2, radius=4)
        assert orc.rel_l2(r.output(), ref["out"]) <= 1e-4


def test_capture_replay_matches_eager(dev):
    w, h = 320, 192
    view, proj, gb, g = frame_inputs(dev, 22, w, h, "mixed")
    r = dev.Renderer(w, h, 16, 32, 5, kind="mixed")
    r.render(g, view, proj)
    torch.cuda.synchronize()
    eager = r.output()
    r2 = dev.Renderer(w, h, 16, 32, 5, kind="mixed")
    s = torch.cuda.Stream()  # the legacy default stream cannot be captured
    with pytest.raises(lgx.LgxError):
        r2.capture(g, view, proj, stream=torch.cuda.default_stream())
    r2.capture(g, view, proj, stream=s)
    torch.cuda.synchronize()
    for i in range(3):
        # eager work of the program's own between replays
        junk = torch.randn(1 << 20, device="cuda") * 2.0
        torch.cuda.synchronize()
        r2.replay(stream=s)
        torch.cuda.synchronize()
        assert r2.output().tobytes() == eager.tobytes(), f"replay {i}"
        del junk
    # the failed capture left no error behind: eager frames still run
    r2.render(g, view, proj)
    torch.cuda.synchronize()
    assert r2.output().tobytes() == eager.tobytes()


def test_capture_replay_unet_frame(dev):
    """The denoised frame (4-direction gather on the 8x8 pattern, U-Net
    DenoiserPass, profiling on) captured into a graph: every replay, with the
    program's own work in between, equals the eager frame bit for bit."""
    w, h = 320, 192
    view, proj, gb, g = frame_inputs(dev, 24, w, h, "mixed")
    net = dev.UNet(1, h, w, seed=7)
    r = dev.Renderer(w, h, 4, 0, 6, lgx.SAMPLER_HIZ_POINT, 8, kind="mixed", denoiser=lgx.DENOISER_UNET,
                     profile=True, footprint_dirs=256)
    r.attach_unet(net)
    r.render(g, view, proj)
    torch.cuda.synchronize()
    eager = r.output()
    s = torch.cuda.Stream()
    r.capture(g, view, proj, stream=s)
    for i in range(3):
        junk = torch.randn(1 << 20, device="cuda") * 2.0
        torch.cuda.synchronize()
        r.replay(stream=s)
        torch.cuda.synchronize()
        assert r.output().tobytes() == eager.tobytes(), f"replay {i}"
        del junk


def test_regression_and_composite(dev):
    w, h = 256, 128
    view, proj, gb, g = frame_inputs(dev, 23, w, h, "f32")
    r = dev.Renderer(w, h, 16, 32, 5, denoiser=lgx.DENOISER_REGRESSION, kind="f32", composite=True)
    r.render(g, view, proj)
    torch.cuda.synchronize()
    ind = r.intermediate("indirect")
    den = orc.regression_denoise(ind, gb["depth"], 2)
    fin = orc.composite(orc.widen(gb["radiance"]), orc.widen(gb["albedo"]), den)
    assert np.allclose(r.output(), fin, rtol=1e-5, atol=1e-6)


def test_renderer_rejects_bad_config(dev):
    c = lgx.RendererConfig()
    c.width, c.height, c.dirs, c.max_steps, c.levels, c.grid, c.sampler = 66, 64, 16, 32, 5, 4, 1
    c.normal_format, c.albedo_format, c.radiance_format = lgx.RG16F, lgx.RGBA8, lgx.RGBA16F
    h = C.c_void_p()
    c.grid, c.grid_y = 8, 16  # 128 pattern offsets: more than the gather's tables hold
    assert lgx.lib().lgx_renderer_create(C.byref(c), C.byref(h)) == lgx.EUNSUPPORTED
    c.grid, c.grid_y, c.width, c.dirs = 4, 0, 64, 0
    assert lgx.lib().lgx_renderer_create(C.byref(c), C.byref(h)) == lgx.EUNSUPPORTED
    assert not h.value


@pytest.mark.parametrize("footprint,steps,grid,h", [(0, 64, 4, 192), (32, 64, 4, 192), (256, 0, 8, 192),
                                                    (256, 0, (4, 16), 192), (256, 0, (4, 16), 200)])
def test_unet_denoiser_pass(dev, footprint, steps, grid, h):
    """LGX_DENOISER_UNET: the renderer's DenoiserPass equals lgx_unet_denoise
    run on the renderer's own gather output, bit for bit, and the gather
    honours footprint_dirs, the step cap and the pattern size; the image
    equals the gather's own interleaved output, through slices and the
    InterleavePass when the pattern divides the frame (16 does not divide
    200: the gather writes the image directly)."""
    w = 320
    view, proj, gb, g = frame_inputs(dev, 23, w, h, "mixed")
    net = dev.UNet(1, h, w, seed=5)
    r = dev.Renderer(w, h, 4, steps, 6, lgx.SAMPLER_HIZ_POINT, grid, kind="mixed", denoiser=lgx.DENOISER_UNET,
                     validate=True, footprint_dirs=footprint)
    with pytest.raises(lgx.LgxError):
        r.render(g, view, proj)  # no network attached yet
    r.attach_unet(net)
    r.render(g, view, proj)
    torch.cuda.synchronize()
    ind = r.intermediate("indirect")
    pyr = dev.PyramidDev(g, 6, lgx.SAMPLER_HIZ_POINT).build()
    direct = dev.Gather(w, h, 4, steps, lgx.SAMPLER_HIZ_POINT, grid=grid, footprint_dirs=footprint)(view, proj, g,
                                                                                                pyr)
    assert direct.numpy().tobytes() == ind.tobytes()
    gx, gy = lgx.pattern(grid)
    if w % gx == 0 and h % gy == 0:
        sl = dev.Gather(w, h, 4, steps, lgx.SAMPLER_HIZ_POINT, grid=grid, out_layout=lgx.OUT_SLICES,
                        footprint_dirs=footprint)(view, proj, g, pyr)
        assert dev.permute(sl, inverse=True, gx=gx, gy=gy).numpy().tobytes() == ind.tobytes()
    noisy = dev.DImage.from_numpy(ind, lgx.RGBA32F)
    ref = net.denoise(g, noisy).numpy()
    assert r.output().tobytes() == ref.tobytes()
    assert np.isfinite(ref).all()
    bil = dev.Renderer(w, h, 4, 64, 6, lgx.SAMPLER_HIZ_POINT, kind="mixed")
    with pytest.raises(lgx.LgxError):
        bil.attach_unet(net)  # the bilateral renderer has no network pass
