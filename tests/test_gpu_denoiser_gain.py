"""What the trained network gains over the method without it.

The committed weights denoise a held-out 1080p frame (seed 207; training used
seeds 1000-2999) from the 4-direction input gather. Against the 256-direction
gather accumulated in FP64 on the GPU, on four 256x256 crops of the frame,
the U-Net output must beat both the noisy input (by 5x) and two
geometry-aware bilateral passes of radius 4 over the same input, the
reference's denoiser weights (denoiser.frag:87-110), by 15 %.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from legitengine_amd import lgx  # noqa: E402

pytestmark = pytest.mark.gpu

CROPS = [(0, 0), (832, 400), (1664, 824), (256, 768)]


def test_trained_unet_beats_input_and_bilateral():
    from legitengine_amd import denoiser as dn
    from legitengine_amd import device as dev
    dev.require_device()
    fg = dn.FrameGen()
    view, proj = fg.view(207, 0)
    noisy = fg.gather(dn.NOISY_DIRS, dn.NOISY_STEPS, footprint=dn.NOISY_FOOTPRINT, grid=dn.NOISY_GRID)(
        view, proj, fg.g, fg.pyr, out=dev.DImage(dn.W, dn.H, lgx.RGBA32F))
    net = dev.UNet(1, dn.H, dn.W)
    net.load(dn.WEIGHTS)
    den = net.denoise(fg.g, noisy)
    bil = noisy
    for _ in range(2):
        bil = dev.bilateral(bil, fg.g, radius=4)
    tgt = fg.gather(dn.TARGET_DIRS, 0, lgx.PREC_F64)
    errs = {"noisy": [0.0, 0.0], "bilateral": [0.0, 0.0], "unet": [0.0, 0.0]}
    for x0, y0 in CROPS:
        tgt(view, proj, fg.g, fg.pyr, out=fg.out64, roi=(x0, y0, dn.CROP, dn.CROP))
        torch.cuda.synchronize()
        t = fg.out64.t[y0:y0 + dn.CROP, x0:x0 + dn.CROP, :3].double()
        for name, img in (("noisy", noisy), ("bilateral", bil), ("unet", den)):
            x = img.t[y0:y0 + dn.CROP, x0:x0 + dn.CROP, :3].double()
            errs[name][0] += float(((x - t) ** 2).sum())
            errs[name][1] += float((t ** 2).sum())
    rel = {k: float(np.sqrt(a / b)) for k, (a, b) in errs.items()}
    print("rel-L2 vs FP64 256-direction target on the crops:", rel, flush=True)
    assert rel["unet"] < 0.2 * rel["noisy"], rel
    assert rel["unet"] < 0.85 * rel["bilateral"], rel
