"""The HIP gather and bilateral on closed-form scenes, checked against the
closed forms themselves (not only against the oracle).

The sphere-around-the-camera scene makes every on-screen sample a hit of the
gather with a horizon angle known in closed form. The two-region field gives
the bilateral a known answer per column. Both are described in
tests/kat_scenes.py.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from legitengine_amd import lgx  # noqa: E402
from tests import kat_scenes as ks  # noqa: E402

pytestmark = pytest.mark.gpu

PIXELS = [(x, y) for y in (7, 15, 22, 30, 37, 45, 52) for x in (9, 20, 33, 40, 47, 58, 70)]
RGB = (0.5, 0.3, 0.15)  # sphere radiance per channel


@pytest.fixture(scope="module")
def dev():
    from legitengine_amd import device
    device.require_device()
    return device


def _sphere_gpu(dev, sc, dirs, fdirs, precision):
    w, h = sc["w"], sc["h"]
    g = dev.GBufferDev(dev.DImage.from_numpy(sc["depth"], lgx.R32F), dev.DImage.from_numpy(sc["oct"], lgx.RG32F),
                       dev.DImage.from_numpy(np.ones((h, w, 4), np.float32), lgx.RGBA32F),
                       dev.DImage.from_numpy(sc["rad"], lgx.RGBA32F))
    pyr = dev.PyramidDev(g, 1, lgx.SAMPLER_HIZ_POINT)
    ga = dev.Gather(w, h, dirs, 0, lgx.SAMPLER_HIZ_POINT, grid=1, precision=precision, footprint_dirs=fdirs)
    out = ga(sc["view"], sc["proj"], g, pyr)
    torch.cuda.synchronize()
    o = out.numpy()
    return np.array([o[y, x, :3] for x, y in PIXELS], np.float64)


@pytest.mark.parametrize("ring_ambient", [True, False])
@pytest.mark.parametrize("dirs,fdirs", [(8, 16), (16, 64)])
def test_gather_sphere_closed_form(dev, ring_ambient, dirs, fdirs):
    """FP64 kernel within 2e-6 and FP32 kernel within 2e-4 (relative) of the
    closed-form hit sums: telescoped (ambient ring) or hit by hit with the
    side weights (uniform radiance).  The three radiance channels differ, so
    a fetch that mixes up channels fails too."""
    sc = ks.sphere_scene(radiance=RGB, ring_ambient=ring_ambient)
    exp = np.array([ks.sphere_expected(sc, x, y, dirs, fdirs) for x, y in PIXELS])
    assert np.abs(exp[:, 0] - exp[:, 2]).min() > 1e-2  # the channels are told apart
    got64 = _sphere_gpu(dev, sc, dirs, fdirs, lgx.PREC_F64)
    assert np.allclose(got64, exp, rtol=0, atol=2e-6), np.abs(got64 - exp).max()
    got32 = _sphere_gpu(dev, sc, dirs, fdirs, lgx.PREC_F32)
    assert np.allclose(got32, exp, rtol=2e-4, atol=0), np.abs(got32 / exp - 1).max()


@pytest.mark.parametrize("radius", [1, 2, 4])
def test_bilateral_two_region_closed_form(dev, radius):
    """Cross-region taps weigh exp(-10.5 |dd|) exp(-20 |dn|^2) = 0.266; the
    kernel's output columns equal the closed-form ratios of tap counts."""
    depth, octn, col = ks.two_region_field()
    h, w = depth.shape
    g = dev.GBufferDev(dev.DImage.from_numpy(depth, lgx.R32F), dev.DImage.from_numpy(octn, lgx.RG32F),
                       dev.DImage.from_numpy(np.ones((h, w, 4), np.float32), lgx.RGBA32F),
                       dev.DImage.from_numpy(np.ones((h, w, 4), np.float32), lgx.RGBA32F))
    inp = dev.DImage.from_numpy(col, lgx.RGBA32F)
    out = dev.bilateral(inp, g, radius=radius).numpy()

    def decode(o):
        v = np.concatenate([o, (1 - np.abs(o).sum(-1))[..., None]], -1).astype(np.float64)
        return v / np.linalg.norm(v, axis=-1, keepdims=True)  # upper hemisphere only in this field

    exp, wc = ks.bilateral_expected(depth, octn, col, radius, split=10, decode=decode)
    assert abs(wc - math.exp(-10.5 * 0.05 - 20 * 0.04)) < 2e-6
    assert np.allclose(out[..., :3], exp, rtol=0, atol=5e-6), np.abs(out[..., :3] - exp).max()


@pytest.mark.parametrize("kind", ["f32", "mixed"])
def test_scene_gpu_matches_analytic(dev, kind):
    """The GPU G-buffer of the hand-placed scene and of three seeded config
    scenes against the float64 ray caster, with the CPU render's bars; and
    where the GPU and CPU renders differ bit for bit, the difference is a
    float32 rounding (depth within 2e-4 relative, radiance within one FP16
    step) or a pixel at a silhouette, edge or shadow boundary."""
    from tests.test_scene_pins import SCENES, check_gbuffer, scene_of
    for name, spec, w, h in SCENES:
        print(f"scene {name} {w}x{h} {kind}", flush=True)
        s = scene_of(spec)
        ref = ks.analytic_gbuffer(lgx, s, w, h)
        g = dev.GBufferDev.alloc(w, h, kind).render(s)
        torch.cuda.synchronize()
        gpu = dict(depth=g.depth.numpy(), normal=g.normal.numpy(), albedo=g.albedo.numpy(),
                   radiance=g.radiance.numpy())
        check_gbuffer(gpu, ref, kind)
        cpu = lgx.render_gbuffer_cpu(s, w, h, kind)
        dd = gpu["depth"] != cpu["depth"]
        rd = (gpu["radiance"] != cpu["radiance"]).any(-1)
        print(f"  GPU vs CPU twin: depth differs at {dd.mean():.4%}, radiance at {rd.mean():.4%} of pixels; "
              f"{(rd & ref['clear']).mean():.4%} away from boundaries", flush=True)
        rel = np.abs(gpu["depth"].astype(np.float64) / cpu["depth"] - 1)
        assert rel[ref["clear"]].max() < 2e-4  # a last-bit difference, amplified at grazing sphere hits
        r_g = gpu["radiance"][..., :3].astype(np.float64)
        r_c = cpu["radiance"][..., :3].astype(np.float64)
        step = 2e-3 if kind == "mixed" else 1e-4
        bad = (np.abs(r_g - r_c) > step * np.maximum(np.abs(r_c), 1e-3)).any(-1) & ref["clear"]
        assert not bad.any(), np.argwhere(bad)[:5]
