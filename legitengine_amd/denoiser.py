"""Learned denoiser: training pairs from the project's own gather, the
training loop on the BF16 MFMA U-Net, and evaluation against the
256-direction FP64 gather.

Pairs: a seeded scene (32 boxes, 8 spheres) seen from a seeded camera pose
at 1920x1080; a 256x256 crop (origin on the 16-pixel grid, so the
interleave pattern phase matches full frames) of the 4-direction gather is the input,
the 256-direction gather accumulated in FP64 on the same crop the target.
The input gather marches with the step growth and mip lod of a larger
direction count (footprint_dirs) and without a step cap: the error of a
few-direction gather against the target is dominated by the coarse
footprint and the truncated far field, not by the angular sampling the
network removes.
Run: python -m legitengine_amd.denoiser train --scenes N --steps K --out PATH
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

from . import device as dev
from . import lgx

W, H = 1920, 1080
LEVELS = 6
CROP = 256
NOISY_DIRS, TARGET_DIRS, PLAIN_DIRS, STEPS = 4, 256, 32, 64
NOISY_FOOTPRINT = 256  # the input marches with the target's step growth and lod (256 directions)
NOISY_STEPS = 0  # and the reference step rule without a cap (the far field matters most)
NOISY_GRID = 8  # interleave pattern of the input gather: 4 directions x 64 offsets = 256 angles per 8x8 block
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_1080p.safetensors")
# (logical cin, cout, resolution divisor) of the ten 3x3 convolutions; the
# 1x1 head maps 32 channels to 3
UNET_CONVS = ((9, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2), (64, 128, 4), (128, 128, 4),
              (192, 64, 2), (64, 64, 2), (96, 32, 1), (32, 32, 1))


def unet_flops_per_pixel():
    """Algorithmic forward FLOPs per full-resolution pixel (multiply-add = 2,
    logical channel counts: the first layer's padding to 32 is not counted)."""
    return sum(2 * 9 * ci * co / (d * d) for ci, co, d in UNET_CONVS) + 2 * 32 * 3


class FrameGen:
    """Device resources to render a seeded view and gather on it."""

    def __init__(self, w=W, h=H, levels=LEVELS):
        self.w, self.h = w, h
        self.g = dev.GBufferDev.alloc(w, h, "mixed")
        self.pyr = dev.PyramidDev(self.g, levels, lgx.SAMPLER_HIZ_POINT)
        self.gathers = {}
        self.out32 = dev.DImage(w, h, lgx.RGBA32F)
        self.out64 = dev.DImage(w, h, lgx.RGBA64F)

    def gather(self, dirs, steps=STEPS, precision=lgx.PREC_F32, footprint=0, grid=4):
        key = (dirs, steps, precision, footprint, grid)
        if key not in self.gathers:
            self.gathers[key] = dev.Gather(self.w, self.h, dirs, steps, lgx.SAMPLER_HIZ_POINT, grid=grid,
                                           precision=precision, footprint_dirs=footprint)
        return self.gathers[key]

    def view(self, scene_seed, pose=0, boxes=32, spheres=8):
        sc = lgx.make_scene(scene_seed, boxes, spheres)
        if pose:
            lgx.scene_pose(sc, (scene_seed << 8) | pose)
        view, proj = lgx.scene_matrices(sc, self.w, self.h)
        self.g.render(sc)
        self.pyr.build()
        return view, proj

    def records(self, noisy, y0=0, x0=0, h=None, w=None):
        """x12 records [h, w, 12] of a region from the G-buffer and a noisy image."""
        h = h or self.h
        w = w or self.w
        sl = (slice(y0, y0 + h), slice(x0, x0 + w))
        rec = torch.zeros((h, w, 12), dtype=torch.float32, device="cuda")
        rec[..., 0:3] = noisy.t[sl][..., :3]
        rec[..., 3] = self.g.depth.t[sl]
        rec[..., 4:6] = self.g.normal.t[sl].float()
        rec[..., 6:9] = self.g.albedo.t[sl][..., :3].float() * np.float32(1.0 / 255.0)
        return rec


def build_dataset(n_scenes, poses=4, seed0=1000, log=print, target_precision=lgx.PREC_F32_ACC64,
                  footprints=(NOISY_FOOTPRINT,), grid=NOISY_GRID):
    """Inputs X[f] [n, CROP, CROP, 12] for each footprint f and targets T
    [n, CROP, CROP, 4], all on the GPU.  The input gather uses the
    interleave pattern `grid` (an int: square, or (width, height)); crop
    origins are multiples of 16 pixels, so the pattern phase matches full
    frames for pattern extents dividing 16."""
    fg = FrameGen()
    n = n_scenes * poses
    X = {f: torch.empty((n, CROP, CROP, 12), dtype=torch.float32, device="cuda") for f in footprints}
    T = torch.empty((n, CROP, CROP, 4), dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(seed0)
    target_g = fg.gather(TARGET_DIRS, 0, target_precision)
    t0 = time.time()
    last = t0
    for i in range(n):
        s, p = seed0 + i // poses, i % poses
        view, proj = fg.view(s, p)
        x0 = int(rng.integers(0, (W - CROP) // 16 + 1)) * 16
        y0 = int(rng.integers(0, (H - CROP) // 16 + 1)) * 16
        roi = (x0, y0, CROP, CROP)
        for f in footprints:
            fg.gather(NOISY_DIRS, NOISY_STEPS, footprint=f, grid=grid)(view, proj, fg.g, fg.pyr, out=fg.out32,
                                                                      roi=roi)
            X[f][i] = fg.records(fg.out32, y0, x0, CROP, CROP)
        target_g(view, proj, fg.g, fg.pyr, out=fg.out64, roi=roi)
        T[i] = fg.out64.t[y0:y0 + CROP, x0:x0 + CROP].float()
        if time.time() - last > 20 or i == n - 1:
            torch.cuda.synchronize()
            last = time.time()
            log(f"dataset {i + 1}/{n} views, {last - t0:.1f} s")
    return X, T


def lr_at(it, steps, lr, schedule):
    """Learning rate of step `it`: constant, or a cosine decay from lr to lr/20."""
    if schedule == "const":
        return lr
    import math
    return lr * (0.05 + 0.95 * 0.5 * (1.0 + math.cos(math.pi * it / max(1, steps - 1))))


def crop_errors(net, X, T, batch=16):
    """Relative squared error ||out - t||^2 / ||t||^2 of every training crop
    under the current network (forward only)."""
    err = torch.empty(X.shape[0], dtype=torch.float64, device="cuda")
    for i in range(0, X.shape[0] - batch + 1, batch):
        out = net.forward(X[i:i + batch])
        d = (out[..., :3] - T[i:i + batch, ..., :3]).double()
        t = T[i:i + batch, ..., :3].double()
        err[i:i + batch] = (d * d).sum((1, 2, 3)) / ((t * t).sum((1, 2, 3)) + 1e-12)
    rest = X.shape[0] % batch
    if rest:  # the last, partial batch: padded with crops of the batch before it
        i = X.shape[0] - batch
        out = net.forward(X[i:])
        d = (out[..., :3] - T[i:, ..., :3]).double()
        t = T[i:, ..., :3].double()
        err[i:] = (d * d).sum((1, 2, 3)) / ((t * t).sum((1, 2, 3)) + 1e-12)
    return err


def train(net, X, T, steps, batch=16, lr=1e-3, log=print, seed=0, log_every=50, schedule="cosine", hard=0.0,
          refresh=5000):
    """Adam on random crops.  hard > 0: every `refresh` steps the crops'
    relative errors under the current network are measured, and a batch
    draws a crop with probability (1 - hard) / n + hard * err / sum(err):
    crops the network gets wrong relative to their own energy (dark scenes,
    small lit regions) are seen more often than their share of the
    absolute loss would give them."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    hist = []
    t0 = time.time()
    torch.cuda.synchronize()
    prob = None
    for it in range(steps):
        if hard > 0 and it % refresh == 0:
            err = crop_errors(net, X, T, batch)
            prob = ((1 - hard) / X.shape[0] + hard * err / err.sum()).float()
            log(f"step {it}: crop relative error mean {float(err.mean()):.5f}, max {float(err.max()):.5f}")
        if prob is None:
            idx = torch.randint(0, X.shape[0], (batch,), device="cuda", generator=gen)
        else:
            idx = torch.multinomial(prob, batch, replacement=True, generator=gen)
        xb = X.index_select(0, idx).contiguous()
        tb = T.index_select(0, idx).contiguous()
        loss.zero_()
        net.train_step(xb, tb, lr_at(it, steps, lr, schedule), loss)
        if it % log_every == 0 or it == steps - 1:
            hist.append((it, float(loss.item())))
            log(f"step {it} loss {hist[-1][1]:.5f} {time.time() - t0:.1f} s")
    torch.cuda.synchronize()
    return hist, time.time() - t0


def rel_l2(a, b):
    a = a[..., :3].double()
    b = b[..., :3].double()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def evaluate(net, seeds, log=print, plain_dirs=(4, 8, 16, 32), footprint=NOISY_FOOTPRINT, grid=NOISY_GRID):
    """rel-L2 of the denoised 4-direction path and of plain paths against the
    256-direction FP64 gather on full 1080p frames."""
    fg = FrameGen()
    res = []
    for s in seeds:
        view, proj = fg.view(s, 0)
        fg.gather(TARGET_DIRS, 0, lgx.PREC_F64)(view, proj, fg.g, fg.pyr, out=fg.out64)
        tgt = fg.out64.t.clone()
        row = {"seed": s}
        fg.gather(NOISY_DIRS, NOISY_STEPS, footprint=footprint, grid=grid)(view, proj, fg.g, fg.pyr, out=fg.out32)
        row["noisy4"] = rel_l2(fg.out32.t, tgt)
        den = net.denoise(fg.g, fg.out32)
        row["denoised4"] = rel_l2(den.t, tgt)
        for d in plain_dirs:
            o = fg.gather(d, STEPS if d <= PLAIN_DIRS else 0)(view, proj, fg.g, fg.pyr, out=fg.out32)
            b = o
            for _ in range(2):
                b = dev.bilateral(b, fg.g, radius=4)
            row[f"plain{d}"] = rel_l2(b.t, tgt)
            row[f"raw{d}"] = rel_l2(o.t, tgt)
        torch.cuda.synchronize()
        log(json.dumps(row))
        res.append(row)
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--scenes", type=int, default=500)
    t.add_argument("--poses", type=int, default=4)
    t.add_argument("--steps", type=int, default=2000)
    t.add_argument("--batch", type=int, default=16)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--out", default=WEIGHTS)
    t.add_argument("--resume", default=None)
    t.add_argument("--eval-seeds", default="200,201")
    t.add_argument("--report", default=None)
    t.add_argument("--footprint", type=int, default=NOISY_FOOTPRINT)
    t.add_argument("--schedule", default="cosine", choices=("cosine", "const"))
    t.add_argument("--loss", default="l1", choices=("l1", "l2"))
    t.add_argument("--grid", type=int, nargs="+", default=None, help="pattern width [height]")
    t.add_argument("--dataset-seed", type=int, default=1000, help="first scene seed of the training pairs")
    t.add_argument("--recipe", default=None, help="text stored as the checkpoint's 'training' metadata")
    t.add_argument("--hard", type=float, default=0.0,
                   help="share of each batch drawn in proportion to the crops' current relative error")
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=WEIGHTS)
    e.add_argument("--seeds", default="200,201,202,203,204,205,206,207")
    e.add_argument("--footprint", type=int, default=NOISY_FOOTPRINT)
    e.add_argument("--grid", type=int, nargs="+", default=None, help="pattern width [height]")
    args = ap.parse_args(argv)
    dev.require_device()

    def log(msg):
        print(msg, flush=True)

    if args.cmd == "train":
        args.grid = NOISY_GRID if args.grid is None else (args.grid[0] if len(args.grid) == 1 else tuple(args.grid))
        X, T = build_dataset(args.scenes, args.poses, seed0=args.dataset_seed, log=log, footprints=(args.footprint,),
                             grid=args.grid)
        X = X[args.footprint]
        net = dev.UNet(args.batch, CROP, CROP, seed=0)
        if args.resume:
            net.load(args.resume)
        net.set_loss(args.loss)
        net.autotune(args.batch, CROP, CROP, backward=True)
        hist, secs = train(net, X, T, args.steps, args.batch, args.lr, log=log, schedule=args.schedule, hard=args.hard)
        os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
        net.save(args.out, metadata={"training": args.recipe} if args.recipe else None)
        log(f"saved {args.out}: {args.steps} steps in {secs:.1f} s ({args.steps / secs:.2f} steps/s)")
        del X, T
        infer = dev.UNet(1, H, W)
        infer.load(args.out)
        infer.autotune(1, H, W)
        rows = evaluate(infer, [int(s) for s in args.eval_seeds.split(",") if s], footprint=args.footprint,
                        grid=args.grid)
        if args.report:
            with open(args.report, "w") as f:
                json.dump({"history": hist, "train_seconds": secs, "steps": args.steps, "batch": args.batch,
                           "scenes": args.scenes, "poses": args.poses, "footprint": args.footprint,
                           "schedule": args.schedule, "loss": args.loss, "lr": args.lr, "grid": args.grid,
                           "resume": args.resume, "dataset_seed": args.dataset_seed, "hard": args.hard,
                           "eval": rows}, f)
    else:
        args.grid = NOISY_GRID if args.grid is None else (args.grid[0] if len(args.grid) == 1 else tuple(args.grid))
        net = dev.UNet(1, H, W)
        net.load(args.weights)
        evaluate(net, [int(s) for s in args.seeds.split(",") if s], footprint=args.footprint, grid=args.grid)
    return 0


if __name__ == "__main__":
    sys.exit(main())
