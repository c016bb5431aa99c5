"""Pins the FP64 U-Net restatement (oracle/unet_ref.py) independently of its
own backward formulas: FP64 central finite differences of the mean-L1 and
mean-squared losses on a 16x16 crop, per-layer directional derivatives and sampled coordinates;
closed forms for the building blocks (convolution against a direct
summation, pooling/upsampling adjointness)."""
import numpy as np
import pytest

from oracle import unet_ref as U


def random_case(seed, n=1, h=16, w=16, scale=0.3):
    rng = np.random.default_rng(seed)
    layers = [(rng.normal(size=(co, ci, 3, 3)) * np.sqrt(2.0 / (9 * ci)), rng.normal(size=co) * 0.05)
              for ci, co in U.LAYERS]
    wh = rng.normal(size=(3, 32)) * scale
    bh = rng.normal(size=3) * 0.05
    flat = U.pack(layers, wh, bh)
    x12 = np.zeros((n, h, w, 12))
    x12[..., :9] = rng.random((n, h, w, 9))
    target = rng.random((n, h, w, 4)) * 2.0
    return flat, x12, target


def test_param_count():
    assert U.param_count() == 473347
    flat, _, _ = random_case(0)
    assert flat.size == U.param_count()


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 5, 6, 3))
    w = rng.normal(size=(4, 3, 3, 3))
    b = rng.normal(size=4)
    y = U.conv(x, w, b)
    ref = np.zeros((1, 5, 6, 4))
    for yy in range(5):
        for xx in range(6):
            for co in range(4):
                s = b[co]
                for ky in range(3):
                    for kx in range(3):
                        sy, sx = yy + ky - 1, xx + kx - 1
                        if 0 <= sy < 5 and 0 <= sx < 6:
                            s += (x[0, sy, sx] * w[co, :, ky, kx]).sum()
                ref[0, yy, xx, co] = s
    assert np.allclose(y, ref, atol=1e-12)


def test_adjoints():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2, 8, 6, 3))
    g = rng.normal(size=(2, 4, 3, 3))
    assert np.isclose((U.pool(x) * g).sum(), (x * U.pool_backward(g)).sum())
    assert np.isclose((U.up(g) * x).sum(), (g * U.up_backward(x)).sum())
    w = rng.normal(size=(5, 3, 3, 3))
    dy = rng.normal(size=(2, 8, 6, 5))
    dx, dw, db = U.conv_backward(x, w, dy)
    assert np.isclose((U.conv(x, w, np.zeros(5)) * dy).sum(), (x * dx).sum())  # <conv x, dy> = <x, conv^T dy>
    eps = 1e-6
    e = rng.normal(size=w.shape)
    fd = ((U.conv(x, w + eps * e, np.zeros(5)) - U.conv(x, w - eps * e, np.zeros(5))) * dy).sum() / (2 * eps)
    assert np.isclose(fd, (dw * e).sum(), rtol=1e-7)


def layer_slices():
    out, off = [], 0
    for ci, co in U.LAYERS:
        out.append((off, off + co * ci * 9 + co))
        off += co * ci * 9 + co
    out.append((off, off + 99))
    return out


@pytest.mark.parametrize("seed,kind", [(3, "l1"), (4, "l1"), (5, "l2")])
def test_backward_matches_finite_differences(seed, kind):
    flat, x12, target = random_case(seed)
    l0, grad = U.backward(flat, x12, target, kind)
    assert np.isclose(l0, U.loss(flat, x12, target, kind))
    rng = np.random.default_rng(seed + 100)
    # small enough that no ReLU or |.| kink is crossed along the direction,
    # large enough that FP64 rounding (~1e-16 / eps) stays far below the bar
    eps = 1e-7
    # a random direction restricted to each layer in turn
    for (a, b) in layer_slices():
        v = np.zeros_like(flat)
        v[a:b] = rng.normal(size=b - a)
        fd = (U.loss(flat + eps * v, x12, target, kind) - U.loss(flat - eps * v, x12, target, kind)) / (2 * eps)
        an = grad @ v
        assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)) + 1e-9, (a, fd, an)
    # sampled coordinates across the whole vector
    for i in rng.choice(flat.size, 24, replace=False):
        e = np.zeros_like(flat)
        e[i] = 1.0
        fd = (U.loss(flat + eps * e, x12, target, kind) - U.loss(flat - eps * e, x12, target, kind)) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-4 * abs(grad[i]), (i, fd, grad[i])


def test_adam_closed_form():
    p, g = np.array([1.0, -2.0]), np.array([0.5, -0.25])
    p1, m, v = U.adam(p, g, np.zeros(2), np.zeros(2), 1, 1e-3)
    assert np.allclose(p1, p - 1e-3 * np.sign(g), atol=1e-10)  # first step moves by lr * sign(g)
