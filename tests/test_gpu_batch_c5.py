"""The uneven batch at its configuration: 32 views of seeded scenes 3000-3031
at resolutions drawn from {512x512, 1280x720, 1920x1080, 2560x1440}, the
4-direction gather (the denoiser's input: 256-direction footprint, no step
cap, 8x8 pattern) plus the trained BF16 U-Net, autotuned launch shape, the
whole batch captured into one hipGraph and replayed.

Checks, for one view of each resolution, on three 64x64 crops (centre,
bottom-right corner, top-left edge):
  * the batch's FP32 gather within 1e-4 rel-L2 of the FP32 oracle on the
    centre crop;
  * on every crop, the FP32 gather no further from the FP64 oracle than the
    FP32 oracle is (x 1.2): near the frame's corners both FP32 evaluations
    lose ~1e-3 to cancellation in the one-pixel tangent (the shader's
    finite difference) and differ from each other by up to ~4e-4, so there the
    parity statement is about accuracy, not about equal rounding;
  * the FP64 gather of the same view (single-view path) within 1e-9 of the
    FP64 oracle on every crop (the algorithm itself, exactly).
Then the denoised images against lgx_unet_denoise on the batch's own noisy
images, bit for bit, on every replay; per-view latencies in completion
order, the median below half of the whole batch."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from legitengine_amd import lgx  # noqa: E402
from oracle import binding as orc  # noqa: E402

pytestmark = pytest.mark.gpu

SEEDS = range(3000, 3032)


@pytest.fixture(scope="module")
def dev():
    from legitengine_amd import device
    device.require_device()
    return device


def test_c5_batch(dev):
    import bench
    from legitengine_amd import denoiser as dn
    sizes = [bench.batch_resolution(s) for s in SEEDS]
    assert set(sizes) == set(bench.BATCH_SIZES), sizes  # every resolution occurs
    views = []
    for s, (w, h) in zip(SEEDS, sizes):
        sc = lgx.make_scene(s, 32, 8)
        view, proj = lgx.scene_matrices(sc, w, h)
        views.append((dev.GBufferDev.alloc(w, h, "mixed").render(sc), view, proj))
    torch.cuda.synchronize()
    net = dev.UNet(1, max(h for _, h in sizes), max(w for w, _ in sizes))
    net.load(dn.WEIGHTS)
    b = dev.Batch(views, dirs=dn.NOISY_DIRS, max_steps=dn.NOISY_STEPS, levels=dn.LEVELS, grid=dn.NOISY_GRID,
                  footprint_dirs=dn.NOISY_FOOTPRINT, net=net)
    print("autotune", b.autotune()["best"], flush=True)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b.capture(s)
    s.synchronize()
    b.replay(s)
    s.synchronize()
    first = {}
    for w_h in bench.BATCH_SIZES:
        i = sizes.index(w_h)
        w, h = w_h
        g, view, proj = views[i]
        print(f"view {i} seed {SEEDS[i]} {w}x{h}: oracle crops", flush=True)
        noisy = b.output(i, 0)
        # the oracle consumes the batch's own G-buffer bytes (rendered on the GPU)
        depth = g.depth.numpy()
        dep_b, rad_b, _, _ = orc.build_pyramids(depth, orc.widen(g.radiance.numpy()), dn.LEVELS,
                                                lgx.SAMPLER_HIZ_POINT, True)
        octn = orc.widen(g.normal.numpy())
        pyr = dev.PyramidDev(g, dn.LEVELS, lgx.SAMPLER_HIZ_POINT).build()
        k64 = dev.Gather(w, h, dn.NOISY_DIRS, dn.NOISY_STEPS, lgx.SAMPLER_HIZ_POINT, grid=dn.NOISY_GRID,
                         precision=lgx.PREC_F64, footprint_dirs=dn.NOISY_FOOTPRINT)(view, proj, g, pyr).numpy()
        for c, (x0, y0) in enumerate([(w // 2 - 32, h // 2 - 32), (w - 64, h - 64), (0, 8)]):
            kw = dict(grid=dn.NOISY_GRID, crop=(x0, y0, 64, 64), footprint_dirs=dn.NOISY_FOOTPRINT)
            o32 = orc.gather(view, proj, depth, octn, dep_b, rad_b, dn.NOISY_DIRS, dn.NOISY_STEPS,
                             lgx.SAMPLER_HIZ_POINT, **kw)
            o64 = orc.gather(view, proj, depth, octn, dep_b, rad_b, dn.NOISY_DIRS, dn.NOISY_STEPS,
                             lgx.SAMPLER_HIZ_POINT, precision="f64", **kw)
            k32 = noisy[y0:y0 + 64, x0:x0 + 64]
            e_k, e_o = orc.rel_l2(k32, o64), orc.rel_l2(o32, o64)
            print(f"  crop ({x0},{y0}): fp32 kernel vs fp32 oracle {orc.rel_l2(k32, o32):.3e}, vs fp64 oracle "
                  f"{e_k:.3e} (fp32 oracle {e_o:.3e}); fp64 kernel vs fp64 oracle "
                  f"{orc.rel_l2(k64[y0:y0 + 64, x0:x0 + 64], o64):.3e}", flush=True)
            if c == 0:
                assert orc.rel_l2(k32, o32) <= 1e-4, (i, c)
            assert e_k <= 1.2 * e_o, (i, c, e_k, e_o)
            assert orc.rel_l2(k64[y0:y0 + 64, x0:x0 + 64], o64) <= 1e-9, (i, c)
        ref_d = net.denoise(g, dev.DImage.from_numpy(noisy, lgx.RGBA32F)).numpy()
        first[i] = b.output(i, 1)
        assert first[i].tobytes() == ref_d.tobytes(), i
    lat = np.array(b.latencies())
    print("latency ms: p50 %.2f p99 %.2f max %.2f" % (np.percentile(lat, 50), np.percentile(lat, 99), lat.max()),
          flush=True)
    order = np.argsort([w * h for w, h in sizes], kind="stable")
    assert (np.diff(lat[order]) >= 0).all()  # views complete in ascending pixel count
    assert np.median(lat) < 0.5 * lat.max()
    for r in range(2):  # replays repeat the eager result
        b.replay(s)
        s.synchronize()
        for i, ref in first.items():
            assert b.output(i, 1).tobytes() == ref.tobytes(), (r, i)
    b.close()
