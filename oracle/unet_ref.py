"""FP64 restatement of the lgx U-Net denoiser (TEST INFRASTRUCTURE ONLY).

The network has no counterpart in the reference (it replaces the regression
DenoiserPass, SSVGIRenderer.h:265-302 / denoiser.frag:112-184); this numpy
version is the checker of the BF16 MFMA kernels: forward, analytic backward
of the mean-L1 or mean-squared loss, and Adam.  Its analytic gradient is pinned by FP64
central finite differences (tests/test_unet_oracle.py).

Parameter layout (flat, FP32 on the device): for each of the 10 convolutions
W[co][ci][3][3] then b[co]; then the head W[3][32], b[3].  Input record per
pixel: noisy irradiance rgb, radial depth, octahedral normal xy, albedo rgb,
3 unused (12 floats).
"""
import numpy as np

LAYERS = [(9, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128), (192, 64), (64, 64), (96, 32), (32, 32)]
HEAD_IN, HEAD_OUT = 32, 3


def param_count():
    return sum(co * ci * 9 + co for ci, co in LAYERS) + HEAD_IN * HEAD_OUT + HEAD_OUT


def unpack(flat):
    flat = np.asarray(flat, np.float64)
    out, off = [], 0
    for ci, co in LAYERS:
        w = flat[off:off + co * ci * 9].reshape(co, ci, 3, 3)
        off += co * ci * 9
        b = flat[off:off + co]
        off += co
        out.append((w, b))
    wh = flat[off:off + HEAD_OUT * HEAD_IN].reshape(HEAD_OUT, HEAD_IN)
    off += HEAD_OUT * HEAD_IN
    bh = flat[off:off + HEAD_OUT]
    return out, wh, bh


def pack(layers, wh, bh):
    parts = []
    for w, b in layers:
        parts += [w.ravel(), b.ravel()]
    parts += [wh.ravel(), bh.ravel()]
    return np.concatenate(parts)


def im2col(x):
    """x [n,h,w,c] -> [n,h,w,9,c] of the zero-padded 3x3 neighbourhood (tap = ky*3+kx)."""
    n, h, w, c = x.shape
    p = np.zeros((n, h + 2, w + 2, c), x.dtype)
    p[:, 1:-1, 1:-1] = x
    cols = np.empty((n, h, w, 9, c), x.dtype)
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, ky * 3 + kx] = p[:, ky:ky + h, kx:kx + w]
    return cols


def conv(x, w, b):
    co = w.shape[0]
    wm = w.transpose(2, 3, 1, 0).reshape(9 * w.shape[1], co)  # [(tap, ci), co]
    cols = im2col(x).reshape(x.shape[0], x.shape[1], x.shape[2], -1)
    return cols @ wm + b


def conv_backward(x, w, dy):
    """Gradients of y = conv(x, w, b) given dy: (dx, dw, db)."""
    n, h, wd, ci = x.shape
    co = w.shape[0]
    cols = im2col(x).reshape(-1, 9 * ci)
    dyf = dy.reshape(-1, co)
    dwm = cols.T @ dyf  # [(tap, ci), co]
    dw = dwm.reshape(3, 3, ci, co).transpose(3, 2, 0, 1)
    db = dyf.sum(0)
    wm = w.transpose(2, 3, 1, 0).reshape(9 * ci, co)
    dcols = (dyf @ wm.T).reshape(n, h, wd, 3, 3, ci)
    dp = np.zeros((n, h + 2, wd + 2, ci), dy.dtype)
    for ky in range(3):
        for kx in range(3):
            dp[:, ky:ky + h, kx:kx + wd] += dcols[:, :, :, ky, kx]
    return dp[:, 1:-1, 1:-1], dw, db


def pool(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool_backward(dy):
    return np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2) * 0.25


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_backward(dy):
    n, h, w, c = dy.shape
    return dy.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def relu(x):
    return np.maximum(x, 0)


def forward(flat, x12, keep=False):
    """x12 [n,h,w,12] -> out [n,h,w,3] (and the activations when keep)."""
    L, wh, bh = unpack(flat)
    x = np.asarray(x12, np.float64)[..., :9]
    a = {}
    a["x"] = x
    a["e1a"] = relu(conv(x, *L[0]))
    a["e1"] = relu(conv(a["e1a"], *L[1]))
    a["p1"] = pool(a["e1"])
    a["e2a"] = relu(conv(a["p1"], *L[2]))
    a["e2"] = relu(conv(a["e2a"], *L[3]))
    a["p2"] = pool(a["e2"])
    a["b1"] = relu(conv(a["p2"], *L[4]))
    a["b2"] = relu(conv(a["b1"], *L[5]))
    a["c2"] = np.concatenate([up(a["b2"]), a["e2"]], axis=-1)
    a["d2a"] = relu(conv(a["c2"], *L[6]))
    a["d2"] = relu(conv(a["d2a"], *L[7]))
    a["c1"] = np.concatenate([up(a["d2"]), a["e1"]], axis=-1)
    a["d1a"] = relu(conv(a["c1"], *L[8]))
    a["d1"] = relu(conv(a["d1a"], *L[9]))
    out = x[..., :3] + a["d1"] @ wh.T + bh
    return (out, a) if keep else out


def loss(flat, x12, target, kind="l1"):
    d = forward(flat, x12) - np.asarray(target, np.float64)[..., :3]
    return float((d * d).mean() if kind == "l2" else np.abs(d).mean())


def backward(flat, x12, target, kind="l1"):
    """Mean-L1 (kind "l1") or mean-squared ("l2") loss over rgb and its
    gradient w.r.t. the flat parameters."""
    L, wh, bh = unpack(flat)
    out, a = forward(flat, x12, keep=True)
    t = np.asarray(target, np.float64)[..., :3]
    cnt = out.size
    g = (2.0 * (out - t) if kind == "l2" else np.sign(out - t)) / cnt
    grads = [None] * 10
    dwh = g.reshape(-1, 3).T @ a["d1"].reshape(-1, 32)
    dbh = g.reshape(-1, 3).sum(0)
    dd1 = (g @ wh) * (a["d1"] > 0)

    def step(i, xin, dy, mask=None):
        dx, dw, db = conv_backward(xin, L[i][0], dy)
        grads[i] = (dw, db)
        return dx

    dx = step(9, a["d1a"], dd1)
    dd1a = dx * (a["d1a"] > 0)
    dc1 = step(8, a["c1"], dd1a)
    dd2 = up_backward(dc1[..., :64]) * (a["d2"] > 0)
    de1_c = dc1[..., 64:]
    dx = step(7, a["d2a"], dd2)
    dd2a = dx * (a["d2a"] > 0)
    dc2 = step(6, a["c2"], dd2a)
    db2 = up_backward(dc2[..., :128]) * (a["b2"] > 0)
    de2_c = dc2[..., 128:]
    dx = step(5, a["b1"], db2)
    db1 = dx * (a["b1"] > 0)
    dp2 = step(4, a["p2"], db1)
    de2 = (de2_c + pool_backward(dp2)) * (a["e2"] > 0)
    dx = step(3, a["e2a"], de2)
    de2a = dx * (a["e2a"] > 0)
    dp1 = step(2, a["p1"], de2a)
    de1 = (de1_c + pool_backward(dp1)) * (a["e1"] > 0)
    dx = step(1, a["e1a"], de1)
    de1a = dx * (a["e1a"] > 0)
    step(0, a["x"], de1a)
    d = out - t
    lval = float((d * d).mean() if kind == "l2" else np.abs(d).mean())
    return lval, pack(grads, dwh, dbh)


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (Kingma & Ba 2015), t = 1-based step index."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
