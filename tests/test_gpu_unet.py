"""BF16 MFMA U-Net kernels (lgx_unet_*) against the FP64 oracle (oracle/unet_ref.py).

Bars, stated for BF16 activations and weights with FP32 accumulation:
forward correction (output - noisy input) within 3e-2 relative L2 of the
FP64 network on the same parameters; the gradient vector within 5e-2
relative L2 overall and 0.15 per layer; the loss within 1e-2 relative; the
Adam update equal to the FP64 update from the same gradients to 1e-5.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import unet_ref as U  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from legitengine_amd import device
    device.require_device()
    return device


def bf16_round(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def test_mfma_operand_layout(dev):
    rng = np.random.default_rng(0)
    A = rng.normal(size=(16, 32)).astype(np.float32)
    B = rng.normal(size=(32, 16)).astype(np.float32)
    D = dev.mfma_probe(A, B)
    ref = bf16_round(A).astype(np.float64) @ bf16_round(B).astype(np.float64)
    assert np.allclose(D, ref, rtol=1e-5, atol=1e-5)


def case(seed, n=2, h=32, w=32):
    rng = np.random.default_rng(seed)
    layers = [(rng.normal(size=(co, ci, 3, 3)) * np.sqrt(2.0 / (9 * ci)), rng.normal(size=co) * 0.05)
              for ci, co in U.LAYERS]
    flat = U.pack(layers, rng.normal(size=(3, 32)) * 0.3, rng.normal(size=3) * 0.05).astype(np.float32)
    x12 = np.zeros((n, h, w, 12), np.float32)
    x12[..., :9] = rng.random((n, h, w, 9))
    target = np.zeros((n, h, w, 4), np.float32)
    target[..., :3] = rng.random((n, h, w, 3)) * 2.0
    return flat, x12, target


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_forward_matches_oracle(dev):
    flat, x12, target = case(1)
    net = dev.UNet(2, 32, 32)
    net.set(0, flat)
    out = net.forward(torch.from_numpy(x12).cuda())
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    ref = U.forward(flat.astype(np.float64), x12)
    assert (o[..., 3] == 1).all()
    assert rel(o[..., :3] - x12[..., :3], ref - x12[..., :3]) <= 3e-2


@pytest.mark.parametrize("kind", ["l1", "l2"])
def test_backward_matches_oracle(dev, kind):
    flat, x12, target = case(2)
    net = dev.UNet(2, 32, 32)
    net.set(0, flat)
    net.set_loss(kind)
    loss = torch.zeros(1, device="cuda")
    net.backward(torch.from_numpy(x12).cuda(), torch.from_numpy(target).cuda(), loss)
    torch.cuda.synchronize()
    g, _ = net.get(1)
    l_ref, g_ref = U.backward(flat.astype(np.float64), x12, target, kind)
    assert abs(loss.item() - l_ref) <= 1e-2 * l_ref
    assert rel(g, g_ref) <= 5e-2
    off = 0
    for ci, co in U.LAYERS:
        n = co * ci * 9 + co
        assert rel(g[off:off + n], g_ref[off:off + n]) <= 0.15, (ci, co)
        off += n
    assert rel(g[off:], g_ref[off:]) <= 0.15


def test_adam_step_matches_oracle(dev):
    flat, x12, target = case(3, n=1, h=16, w=16)
    net = dev.UNet(1, 16, 16)
    net.set(0, flat)
    net.backward(torch.from_numpy(x12).cuda(), torch.from_numpy(target).cuda())
    g, _ = net.get(1)
    net.adam_step(1e-3)
    p1, t = net.get(0)
    ref, _, _ = U.adam(flat.astype(np.float64), g.astype(np.float64), np.zeros(flat.size), np.zeros(flat.size), 1, 1e-3)
    assert t == 1
    assert np.allclose(p1, ref, rtol=1e-5, atol=1e-7)


def test_training_reduces_loss(dev):
    flat, x12, target = case(4, n=2, h=32, w=32)
    target[..., :3] = 0.5 * x12[..., :3] + 0.1  # a learnable map
    net = dev.UNet(2, 32, 32, seed=7)
    x, t = torch.from_numpy(x12).cuda(), torch.from_numpy(target).cuda()
    loss = torch.zeros(1, device="cuda")
    losses = []
    for i in range(80):
        loss.zero_()
        net.train_step(x, t, 1e-3, loss)
        losses.append(loss.item())
    # Adam's first steps at lr 1e-3 can overshoot; the tail must sit well below the start
    assert np.mean(losses[-10:]) < 0.6 * losses[0], losses


def test_checkpoint_roundtrip(dev, tmp_path):
    net = dev.UNet(1, 16, 16, seed=3)
    x = torch.rand((1, 16, 16, 12), device="cuda")
    t = torch.rand((1, 16, 16, 4), device="cuda")
    net.train_step(x, t, 1e-3)
    p = str(tmp_path / "ck.safetensors")
    net.save(p)
    other = dev.UNet(1, 16, 16, seed=99)
    other.load(p)
    assert np.array_equal(other.get(0)[0], net.get(0)[0]) and other.get(0)[1] == 1
    assert np.array_equal(other.get(2)[0], net.get(2)[0])
    a = net.forward(x)
    b = other.forward(x)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


def test_denoise_from_gbuffer_matches_records(dev):
    from legitengine_amd import lgx
    sc = lgx.make_scene(5, 8, 2)
    gb = lgx.render_gbuffer_cpu(sc, 64, 48, "mixed")
    g = dev.GBufferDev.from_host(gb)
    noisy = np.random.default_rng(0).random((48, 64, 4)).astype(np.float32)
    dn = dev.DImage.from_numpy(noisy, lgx.RGBA32F)
    net = dev.UNet(1, 48, 64, seed=1)
    out = net.denoise(g, dn)
    rec = np.zeros((1, 48, 64, 12), np.float32)
    rec[0, ..., :3] = noisy[..., :3]
    rec[0, ..., 3] = gb["depth"]
    rec[0, ..., 4:6] = gb["normal"].astype(np.float32)
    rec[0, ..., 6:9] = gb["albedo"][..., :3].astype(np.float32) * np.float32(1.0 / 255.0)
    ref = net.forward(torch.from_numpy(rec).cuda())
    torch.cuda.synchronize()
    o, r = out.numpy(), ref.cpu().numpy()[0]
    assert np.isfinite(o).all()
    assert np.allclose(o, r, atol=1e-6)
    again = net.denoise(g, dn).numpy()
    assert again.tobytes() == o.tobytes()


def test_register_tiles_bit_identical_and_autotune(dev):
    """Every convolution register tile accumulates each output in the same
    order, so forward outputs are bit-identical whatever the tiles, and so
    are the activation gradients and the convolution weight and bias
    gradients (the per-workgroup partial sums are reduced in a fixed order);
    the 1x1 head's gradients differ only by the order of their FP32 atomic
    sums.  The autotuner only picks among them."""
    from legitengine_amd import lgx
    flat, x12, target = case(5, n=2, h=48, w=80)
    net = dev.UNet(2, 48, 80)
    net.set(0, flat)
    x, t = torch.from_numpy(x12).cuda(), torch.from_numpy(target).cuda()
    couts = [co for _, co in U.LAYERS]
    cins = [32, 32, 32, 64, 64, 128, 192, 64, 96, 32]  # padded input channels
    ref_out = net.forward(x).cpu().numpy()
    net.backward(x, t)
    ref_grad = net.get(1)[0]
    wmask = np.zeros(ref_grad.size, bool)
    off = 0
    for ci, co in U.LAYERS:
        wmask[off:off + co * ci * 9 + co] = True
        off += co * ci * 9 + co
    net.backward(x, t)
    assert net.get(1)[0][wmask].tobytes() == ref_grad[wmask].tobytes()  # same call twice
    counts_f = [lgx.lib().lgx_unet_conv_tiles(c) for c in couts]
    counts_d = [lgx.lib().lgx_unet_conv_tiles(c) for c in cins]
    assert min(counts_f) >= 2 and min(counts_d) >= 2
    for k in range(1, max(counts_f + counts_d)):
        net.set_tiles([k % c for c in counts_f], [k % c for c in counts_d])
        assert net.forward(x).cpu().numpy().tobytes() == ref_out.tobytes(), k
        net.backward(x, t)
        g = net.get(1)[0]
        assert g[wmask].tobytes() == ref_grad[wmask].tobytes(), k
        assert rel(g, ref_grad) < 1e-5, k
    rep = net.autotune(2, 48, 80, backward=True)
    f, d = net.tiles()
    assert [r["best"] for r in rep["forward"]] == f
    assert [r["best"] for r in rep["input_gradient"]] == d[1:]
    assert net.forward(x).cpu().numpy().tobytes() == ref_out.tobytes()


def test_fused_head_every_tile(dev):
    """Inference fuses both ends of the network: the first two convolutions
    run in one kernel from the G-buffer (whatever tiles layers 0 and 1 have),
    and the head runs in the epilogue of the last convolution (the autotuned
    tile, or the first LDS tile when the chosen tile has no LDS form). Every
    tile gives the same bits, and the forward's output on the records."""
    from legitengine_amd import lgx
    sc = lgx.make_scene(6, 8, 2)
    gb = lgx.render_gbuffer_cpu(sc, 80, 48, "mixed")
    g = dev.GBufferDev.from_host(gb)
    noisy = np.random.default_rng(2).random((48, 80, 4)).astype(np.float32)
    dn = dev.DImage.from_numpy(noisy, lgx.RGBA32F)
    net = dev.UNet(1, 48, 80, seed=4)
    rec = np.zeros((1, 48, 80, 12), np.float32)
    rec[0, ..., :3] = noisy[..., :3]
    rec[0, ..., 3] = gb["depth"]
    rec[0, ..., 4:6] = gb["normal"].astype(np.float32)
    rec[0, ..., 6:9] = gb["albedo"][..., :3].astype(np.float32) * np.float32(1.0 / 255.0)
    ref = net.forward(torch.from_numpy(rec).cuda()).cpu().numpy()[0]
    f0, d0 = net.tiles()
    n9 = lgx.lib().lgx_unet_conv_tiles(32)
    assert n9 >= 2
    first = net.denoise(g, dn).numpy()
    assert np.allclose(first, ref, rtol=0, atol=1e-6), np.abs(first - ref).max()
    for layer in (0, 1, 9):
        for t in range(n9):
            f = list(f0)
            f[layer] = t
            net.set_tiles(f, d0)
            out = net.denoise(g, dn).numpy()
            assert out.tobytes() == first.tobytes(), (layer, t, np.abs(out - first).max())


def test_denoise_rejects_bad_images_on_the_host(dev):
    """lgx_unet_denoise checks every image it reads or writes before any
    launch: null or unpitched images, wrong formats, and extents that differ
    from the noisy image (the descriptors below are never dereferenced)."""
    import ctypes as C
    from legitengine_amd import lgx
    L = lgx.lib()
    net = dev.UNet(1, 32, 32, seed=1)
    fake = 0x100000
    w, h = 32, 32

    def gbuf(nf=lgx.RG16F, af=lgx.RGBA8, dh=h, nh=h, ah=h, df=lgx.R32F, null_normal=False):
        g = lgx.GBuffer()
        g.depth = lgx.image(fake, w, dh, df)
        g.normal = lgx.image(0 if null_normal else fake, w, nh, nf)
        g.albedo = lgx.image(fake, w, ah, af)
        g.radiance = lgx.image(fake, w, h, lgx.RGBA16F)
        return g

    noisy = lgx.image(fake, w, h, lgx.RGBA32F)
    out = lgx.image(fake, w, h, lgx.RGBA32F)
    cases = [
        (gbuf(nh=h - 4), lgx.ESIZE), (gbuf(ah=h - 4), lgx.ESIZE), (gbuf(dh=h - 4), lgx.ESIZE),
        (gbuf(df=lgx.RG32F), lgx.EFORMAT), (gbuf(nf=lgx.RG32F), lgx.EFORMAT),
        (gbuf(null_normal=True), lgx.EINVAL),
    ]
    for g, code in cases:
        assert L.lgx_unet_denoise(net.h, C.byref(g), C.byref(noisy), C.byref(out), None) == code
    bad = lgx.image(fake, w, h, lgx.RGBA32F, pitch=w - 1)
    g = gbuf()
    assert L.lgx_unet_denoise(net.h, C.byref(g), C.byref(noisy), C.byref(bad), None) == lgx.EINVAL
    small = lgx.image(fake, w, h - 4, lgx.RGBA32F)
    assert L.lgx_unet_denoise(net.h, C.byref(g), C.byref(noisy), C.byref(small), None) == lgx.ESIZE

