"""Benchmark: GI ms/frame at 1920x1080 on one MI355X, plain vs denoised at
equal relative L2 against the 256-direction gather accumulated in FP64.

`value` is the plain HIP path (the configuration the metric is quoted on):
seeded scene with 32 boxes and 8 spheres, depth FP32, normals and radiance
FP16, albedo RGBA8; 32 directions, 64 march steps, 6-level Hi-Z pyramid,
4x4 interleave, 2 bilateral passes of radius 4, FP32 accumulation.  One
step = one frame of the whole plain path (mips, pre-blur, gather,
interleave, 2 bilateral passes) on a G-buffer already resident in HBM.

The line also carries the roofline of the dominant kernel (the gather;
achieved = algorithmic bytes per launch / its average launch time measured
with HIP events on the launch stream, traffic = the copy-calibrated HBM
bytes of the committed counter summary) and the CPU baseline (the FP32
oracle, timed on a bounded sample of the same frame).

"denoised" holds the other half of the metric: the renderer with a
4-direction gather (the target's step growth and lod, no step cap) and the
trained BF16 U-Net as its denoiser pass, timed the same way; the relative L2
error of its output and of the plain path at 32..256 directions against the
256-direction gather computed in FP64 on the GPU (seeded scenes 200..207);
the plain path's ms/frame at the denoised path's error (log-log
interpolation over the sweep) and the ratio.  "training" is the U-Net
training step (batch 16 of 256x256 crops) and "uneven_batch" the 32-view
batch in one graph replay.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W, H = 1920, 1080
DIRS, STEPS, LEVELS, GRID = 32, 64, 6, 4
BILAT_PASSES, BILAT_RADIUS = 2, 4
SEED = 100
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per unit (storage precision, no cache credit)
BD, BR, BC, BO = 4, 8, 8, 16
DEN_DIRS, TARGET_DIRS = 4, 256
# a second denoised configuration, for the error bar: more directions on the
# 4x4 pattern through the same network
QDEN_DIRS, QDEN_GRID = 16, 4
PLAIN_SWEEP = (32, 64, 128, 192, 256)
QUALITY_SEEDS = (200, 201, 202, 203, 204, 205, 206, 207)
BATCH_SEEDS = range(3000, 3032)
BATCH_SIZES = ((512, 512), (1280, 720), (1920, 1080), (2560, 1440))
MFMA_BF16_PEAK_TFLOPS = 2500.0


def log(msg):
    """Progress on stderr (stdout carries only the JSON line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows-stride", type=int, default=8,
                    help="CPU baseline: the gather runs on every n-th row (stratified), scaled to the frame")
    ap.add_argument("--no-denoised", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-train", action="store_true")
    ap.add_argument("--gather-tile", type=int, default=-1,
                    help="fixed gather launch shape for the plain path (default: the autotuner's choice)")
    ap.add_argument("--no-quality", action="store_true",
                    help="time the denoised path only (no error sweep against the FP64 target)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only launch a 256 MiB device copy 3x and the 1080p gather 3x (for rocprofv3 --pmc)")
    ap.add_argument("--probe-denoised", action="store_true",
                    help="with --traffic-probe: the denoised path's gather configuration")
    ap.add_argument("--unet-probe", action="store_true",
                    help="only run the 1080p U-Net denoiser pass 3x and 3 training steps (batch 16 of 256x256) "
                         "with autotuned tiles (for rocprofv3 --pmc)")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "gather_1080p_plain_traffic.md"))
    ap.add_argument("--counters-file", default=os.path.join(ROOT, "profiles", "gather_1080p_plain_counters.md"))
    ap.add_argument("--quality-seeds", default=",".join(str(s) for s in QUALITY_SEEDS))
    ap.add_argument("--weights", default=None)
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the GI path runs on a single MI355X "
              "(sharding across GPUs is described, not measured)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from legitengine_amd import device as dev
    from legitengine_amd import lgx
    dev.require_device()

    sampler = lgx.SAMPLER_HIZ_POINT
    scene = lgx.make_scene(SEED, 32, 8)
    view, proj = lgx.scene_matrices(scene, W, H)
    g = dev.GBufferDev.alloc(W, H, "mixed").render(scene)
    if args.traffic_probe:
        traffic_probe(g, view, proj, sampler, args.gather_tile, args.probe_denoised)
        return
    if args.unet_probe:
        unet_probe(args, g)
        return
    # The drop-in path: SsvgiRenderer passes declared on the C++ RenderGraph,
    # timed per pass with hipEvents on the frame's stream; the gather's launch
    # shape comes from the autotuner (all shapes give identical results).
    pyr = dev.PyramidDev(g, LEVELS, sampler).build()
    if args.gather_tile >= 0:
        plain_tile, plain_tiles = args.gather_tile, {}
    else:
        plain_tile, plain_tiles = dev.autotune_gather(g, pyr, view, proj, DIRS, STEPS, sampler)
    renderer = dev.Renderer(W, H, DIRS, STEPS, LEVELS, sampler, GRID, denoiser=lgx.DENOISER_BILATERAL,
                            denoise_passes=BILAT_PASSES, denoise_radius=BILAT_RADIUS, kind="mixed", profile=True,
                            tile=plain_tile)

    # step and hit counts (outside the timed region) for the algorithmic byte
    # model: every executed step reads one depth texel, and only the steps
    # whose sample rises above the running horizon read a radiance texel
    # (indirectLighting.frag:254-256); each pixel reads its centre depth and
    # normal and writes its output once
    counts = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    hits = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    dev.Gather(W, H, DIRS, STEPS, sampler)(view, proj, g, pyr, step_counts=counts, hit_counts=hits)
    torch.cuda.synchronize()
    total_steps = int(counts.sum().item())
    total_hits = int(hits.sum().item())
    del pyr, counts, hits
    gather_bytes = total_steps * BD + total_hits * BR + W * H * (BC + BO)
    gather_bytes_per_step_model = total_steps * (BD + BR) + W * H * (BC + BO)

    # per-pass times from eager frames (a hipEvent pair around every pass);
    # the frame time from the same frame captured once into a hipGraph and
    # replayed, the renderer's submission path (the per-frame kernel launches
    # and events of the eager frame are host overhead, reported beside it)
    ms_per_step, passes, eager_ms = time_renderer(renderer, g, view, proj, args.steps, args.warmup)
    gather_ms = sum(p["ms"] for p in passes if p["name"] == "IndirectLightPass")
    achieved = gather_bytes / (gather_ms * 1e-3) / 1e9

    log(f"plain path: {ms_per_step:.3f} ms/frame (graph replay; eager {eager_ms:.3f}), gather {gather_ms:.3f} ms")
    traffic = read_traffic(args.traffic_file)
    result = {
        "metric": "GI ms/frame @1080p: plain vs denoised at equal rel-L2 vs FP64 oracle",
        "value": round(ms_per_step, 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded analytic scene, G-buffer rendered on the GPU)",
        "config": {"workload": "plain HIP gather 1920x1080: 32 dirs, 64 steps, 6-level Hi-Z, 4x4 interleave, "
                               "2 bilateral passes r=4; depth FP32, normal/radiance FP16, albedo RGBA8",
                   "width": W, "height": H, "dirs": DIRS, "max_steps": STEPS, "levels": LEVELS, "seed": SEED},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "model": "algorithmic bytes at storage precision, no cache credit: depth 4 B per executed "
                              "step, radiance 8 B per horizon hit, centre 8 B + output 16 B per pixel",
                     "kernel": "gather_kernel (IndirectLightPass)", "kernel_ms": round(gather_ms, 4),
                     "tile": plain_tile, "tile_ms": {k: round(v, 4) for k, v in plain_tiles.items()},
                     "algorithmic_bytes": gather_bytes, "steps_executed": total_steps, "hits": total_hits,
                     "frac_radiance_every_step": round(gather_bytes_per_step_model / (gather_ms * 1e-3) / 1e9
                                                       / HBM_PEAK_GBS, 4),
                     "measured_hbm_frac": (round(traffic / (gather_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)
                                           if traffic else None),
                     **read_counters(args.counters_file)},
        "passes_ms": {f"{i}:{p['name']}": round(p["ms"], 4) for i, p in enumerate(passes)},
        "frame_submission": "frame captured once into a hipGraph (lgx_renderer_capture) and replayed per frame",
        "eager_ms_per_frame": round(eager_ms, 4),
    }

    if not args.no_denoised:
        seeds = [int(s) for s in args.quality_seeds.split(",") if s]
        result["denoised"] = denoised_leg(args, g, view, proj, sampler, ms_per_step, seeds)
        log(f"denoised path: {result['denoised']['ms_per_frame']:.3f} ms/frame, "
            f"rel-L2 {result['denoised'].get('rel_l2_vs_fp64_256')}")

    if not args.no_train:
        result["training"] = train_leg(args)
        log(f"training: {result['training']['steps_per_s']} steps/s")

    if not args.no_batch:
        result["uneven_batch"] = batch_leg(args, sampler)
        log(f"uneven batch: {result['uneven_batch']['mpixel_per_s']} Mpixel/s")

    if not args.no_cpu_baseline:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(g, view, proj, sampler, renderer.intermediate("indirect"),
                                              args.cpu_rows_stride)
    print(json.dumps(result), flush=True)


COPY_BYTES = 256 << 20


def traffic_probe(g, view, proj, sampler, tile=-1, denoised=False):
    """Kernels for the counters: a plain device copy of known size (the
    traffic calibration) and the bench's gather launch, three times each, in
    the launch shape the bench times (`tile`, or the autotuner's pick; the
    autotuner's own launches precede the copy).  `denoised`: the denoised
    path's gather (4 directions, 256-direction footprint, no step cap, 8x8
    pattern) instead of the plain one; its step and hit counts are printed."""
    import torch
    from legitengine_amd import device as dev
    from legitengine_amd import denoiser as dn
    from legitengine_amd import lgx
    pyr = dev.PyramidDev(g, LEVELS, sampler).build()
    cfg = dict(dirs=DIRS, steps=STEPS, footprint=0, grid=GRID)
    if denoised:
        cfg = dict(dirs=DEN_DIRS, steps=dn.NOISY_STEPS, footprint=dn.NOISY_FOOTPRINT, grid=dn.NOISY_GRID)
    if tile < 0:
        tile, _ = dev.autotune_gather(g, pyr, view, proj, cfg["dirs"], cfg["steps"], sampler, cfg["footprint"],
                                      grid=cfg["grid"])
    x = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    ga = dev.Gather(W, H, cfg["dirs"], cfg["steps"], sampler, grid=cfg["grid"], tile=tile,
                    footprint_dirs=cfg["footprint"])
    out = dev.DImage(W, H, lgx.RGBA32F)
    if denoised:
        counts = torch.zeros((H, W), dtype=torch.int32, device="cuda")
        hits = torch.zeros((H, W), dtype=torch.int32, device="cuda")
        dev.Gather(W, H, cfg["dirs"], cfg["steps"], sampler, grid=cfg["grid"],
                   footprint_dirs=cfg["footprint"])(view, proj, g, pyr, step_counts=counts, hit_counts=hits)
        torch.cuda.synchronize()
        print(json.dumps({"steps_executed": int(counts.sum().item()), "hits": int(hits.sum().item())}), flush=True)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    for _ in range(3):
        ga(view, proj, g, pyr, out=out)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": COPY_BYTES, "copy_bytes_written": COPY_BYTES, "gather_launches": 3,
                      "tile": tile}), flush=True)


def unet_probe(args, g):
    """Kernels for the U-Net counters: a 256 MiB device copy 3x (the traffic
    calibration), the 1080p denoiser pass (trained weights, autotuned tiles)
    3x, then 3 training steps on batch 16 of 256x256 random records."""
    import torch
    from legitengine_amd import device as dev
    from legitengine_amd import denoiser as dn
    from legitengine_amd import lgx
    x = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    net = dev.UNet(1, H, W)
    net.load(args.weights or dn.WEIGHTS)
    net.autotune(1, H, W)
    noisy = dev.DImage(W, H, lgx.RGBA32F)
    noisy.t.uniform_(0, 1)
    out = dev.DImage(W, H, lgx.RGBA32F)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    for _ in range(3):
        net.denoise(g, noisy, out=out)
    torch.cuda.synchronize()
    tnet = dev.UNet(16, dn.CROP, dn.CROP, seed=0)
    tnet.autotune(16, dn.CROP, dn.CROP, backward=True)
    xb = torch.rand((16, dn.CROP, dn.CROP, 12), device="cuda")
    tb = torch.rand((16, dn.CROP, dn.CROP, 4), device="cuda")
    torch.cuda.synchronize()
    for _ in range(3):
        tnet.train_step(xb, tb, 1e-3)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": COPY_BYTES, "denoise_launches": 3, "train_steps": 3}), flush=True)


def read_traffic(path):
    """HBM bytes per gather launch from the committed counter summary (the line
    'traffic_bytes_per_launch: N'), or None."""
    try:
        for line in open(path):
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return int(float(line.split(":", 1)[1].strip()))
    except OSError:
        return None
    return None


def read_counters(path):
    """Measured utilisation of the gather from the committed counter summary
    (lines 'counter_<key>: value': VALU issue and lane utilisation, waves
    per SIMD, L1 / L2 hit rates, texture-address busy, ...), or {}."""
    out = {}
    try:
        for line in open(path):
            line = line.strip()
            if line.startswith("counter_") and ":" in line:
                k, v = line.split(":", 1)
                try:
                    out[k.strip()[len("counter_"):]] = float(v.strip())
                except ValueError:
                    pass
    except OSError:
        return {}
    if not out:
        return {}
    res = {"measured": out}
    if "valu_issue_busy" in out:  # the unit that binds in practice: VALU issue slots used
        res["valu_frac"] = out["valu_issue_busy"]
    return res


def time_replay(r, g, view, proj, steps, warmup):
    """ms per frame of the renderer's frame captured once into a hipGraph
    (lgx_renderer_capture: the same passes, no per-pass events) and replayed
    `steps` times on a side stream, bracketed by device synchronisations."""
    import torch
    st = torch.cuda.Stream()
    r.capture(g, view, proj, st)
    for _ in range(warmup):
        r.replay(st)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.replay(st)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def time_renderer(r, g, view, proj, steps, warmup):
    """(ms per frame of the replayed frame graph, per-pass times of eager
    frames with a hipEvent pair around every pass, ms per eager frame)."""
    import torch
    for _ in range(warmup):
        r.render(g, view, proj)
    torch.cuda.synchronize()
    r.pass_times(mode=2)
    t0 = time.perf_counter()
    for _ in range(steps):
        r.render(g, view, proj)
    torch.cuda.synchronize()
    eager = (time.perf_counter() - t0) * 1e3 / steps
    passes = r.pass_times(mode=1)
    return time_replay(r, g, view, proj, steps, warmup), passes, eager


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a, np.float64)[..., :3]
    b = np.asarray(b, np.float64)[..., :3]
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def equal_error_ms(err, sweep):
    """ms/frame of a plain family at error `err`: log-log interpolation
    between its two sweep points that bracket it (None outside the sweep).
    sweep: (error, ms) points of one family in increasing direction count."""
    import math
    for (e0, m0), (e1, m1) in zip(sweep, sweep[1:]):
        if e0 >= err >= e1 and e0 > e1:
            t = (math.log(e0) - math.log(err)) / (math.log(e0) - math.log(e1))
            return math.exp(math.log(m0) + t * (math.log(m1) - math.log(m0)))
    return None


def denoised_leg(args, g, view, proj, sampler, plain32_ms, seeds):
    """Denoised 4-direction path vs the plain path at equal error against the
    FP64 256-direction gather."""
    import numpy as np
    from legitengine_amd import device as dev
    from legitengine_amd import lgx
    from legitengine_amd import denoiser as dn
    path = args.weights or dn.WEIGHTS
    if not os.path.exists(path):
        raise FileNotFoundError(f"trained U-Net weights not found at {path}")
    net = dev.UNet(1, H, W)
    net.load(path)
    unet_tune = net.autotune(1, H, W)
    pyr0 = dev.PyramidDev(g, LEVELS, sampler).build()
    den_tile, den_tiles = dev.autotune_gather(g, pyr0, view, proj, DEN_DIRS, dn.NOISY_STEPS, sampler,
                                              dn.NOISY_FOOTPRINT, grid=dn.NOISY_GRID)
    rden = dev.Renderer(W, H, DEN_DIRS, dn.NOISY_STEPS, LEVELS, sampler, dn.NOISY_GRID, denoiser=lgx.DENOISER_UNET,
                        kind="mixed", profile=True, footprint_dirs=dn.NOISY_FOOTPRINT, tile=den_tile)
    rden.attach_unet(net)
    den_ms, den_passes, den_eager = time_renderer(rden, g, view, proj, args.steps, args.warmup)
    log(f"denoised renderer timed: {den_ms:.3f} ms/frame (eager {den_eager:.3f})")
    unet_ms = sum(p["ms"] for p in den_passes if p["name"] == "DenoiserPass")
    flops = dn.unet_flops_per_pixel() * W * H
    timing = {
        "ms_per_frame": round(den_ms, 4),
        "eager_ms_per_frame": round(den_eager, 4),
        "unet_ms": round(unet_ms, 4),
        "passes_ms": {f"{i}:{p['name']}": round(p["ms"], 4) for i, p in enumerate(den_passes)},
        "unet_roofline": {"bound": "mfma", "achieved": round(flops / (unet_ms * 1e-3) / 1e12, 2),
                          "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                          "frac": round(flops / (unet_ms * 1e-3) / 1e12 / MFMA_BF16_PEAK_TFLOPS, 4),
                          "flops_per_frame": flops},
        "unet_tiles": [r["best"] for r in unet_tune["forward"]],
        "unet_layer_ms": [round(min(r["ms"]), 4) for r in unet_tune["forward"]],
        "gather_tile": den_tile,
        "gather_tile_ms": {k: round(v, 4) for k, v in den_tiles.items()},
    }
    if args.no_quality:
        return timing
    plain = {}
    for d in PLAIN_SWEEP:
        for kind, den in (("plain", lgx.DENOISER_BILATERAL), ("raw", lgx.DENOISER_NONE)):
            r = dev.Renderer(W, H, d, STEPS if d <= DIRS else 0, LEVELS, sampler, GRID, denoiser=den,
                             denoise_passes=BILAT_PASSES, denoise_radius=BILAT_RADIUS, kind="mixed", profile=True)
            ms, _, _ = time_renderer(r, g, view, proj, max(3, args.steps // 4), 1)
            plain[f"{kind}{d}"] = {"renderer": r, "ms": ms, "err": []}
            log(f"{kind}{d}: {ms:.3f} ms/frame")
    # the quality configuration: the same network on a 16-direction 4x4 input
    q_tile, _ = dev.autotune_gather(g, pyr0, view, proj, QDEN_DIRS, dn.NOISY_STEPS, sampler, dn.NOISY_FOOTPRINT,
                                    grid=QDEN_GRID)
    rq = dev.Renderer(W, H, QDEN_DIRS, dn.NOISY_STEPS, LEVELS, sampler, QDEN_GRID, denoiser=lgx.DENOISER_UNET,
                      kind="mixed", profile=True, footprint_dirs=dn.NOISY_FOOTPRINT, tile=q_tile)
    rq.attach_unet(net)
    q_ms, q_passes, _ = time_renderer(rq, g, view, proj, max(3, args.steps // 4), 1)
    log(f"quality configuration ({QDEN_DIRS} directions, {QDEN_GRID}x{QDEN_GRID}): {q_ms:.3f} ms/frame")
    den_err, noisy_err, q_err = [], [], []
    gs = dev.GBufferDev.alloc(W, H, "mixed")
    target = dev.Gather(W, H, TARGET_DIRS, 0, sampler, precision=lgx.PREC_F64)
    for s in seeds:
        sc = lgx.make_scene(s, 32, 8)
        v, p = lgx.scene_matrices(sc, W, H)
        gs.render(sc)
        pyr = dev.PyramidDev(gs, LEVELS, sampler).build()
        tgt = target(v, p, gs, pyr).numpy()
        rden.render(gs, v, p)
        den_err.append(rel_l2(rden.output(), tgt))
        noisy_err.append(rel_l2(rden.intermediate("indirect"), tgt))
        rq.render(gs, v, p)
        q_err.append(rel_l2(rq.output(), tgt))
        for e in plain.values():
            e["renderer"].render(gs, v, p)
            e["err"].append(rel_l2(e["renderer"].output(), tgt))
        log(f"seed {s}: denoised rel-L2 {den_err[-1]:.4f}")
    err = float(np.mean(den_err))
    sweep = {k: {"ms_per_frame": round(e["ms"], 4), "rel_l2": round(float(np.mean(e["err"])), 5)}
             for k, e in plain.items()}
    # each plain family (2 bilateral passes / raw) interpolated along its own
    # direction counts; the faster family at that error counts
    fam = {}
    for kind in ("plain", "raw"):
        f = equal_error_ms(err, [(float(np.mean(plain[f"{kind}{d}"]["err"])), plain[f"{kind}{d}"]["ms"])
                                 for d in PLAIN_SWEEP])
        if f is not None:
            fam[kind] = f
    eq_ms = min(fam.values()) if fam else None
    qerr = float(np.mean(q_err))
    q_fam = [equal_error_ms(qerr, [(float(np.mean(plain[f"{kind}{d}"]["err"])), plain[f"{kind}{d}"]["ms"])
                                   for d in PLAIN_SWEEP]) for kind in ("plain", "raw")]
    q_fam = [f for f in q_fam if f is not None]
    q_eq = min(q_fam) if q_fam else None
    log(f"quality configuration: rel-L2 {qerr:.5f}")
    for e in plain.values():
        e["renderer"].close()
    rq.close()
    gx, gy = lgx.pattern(dn.NOISY_GRID)
    return {
        "workload": f"renderer 1920x1080: {DEN_DIRS}-direction gather (step growth and lod of {dn.NOISY_FOOTPRINT} "
                    f"directions, reference step rule without cap), {LEVELS}-level Hi-Z, {gx}x{gy} interleave, "
                    "BF16 U-Net denoiser pass (3 levels, 32/64/128 channels, FP32 accumulation)",
        **timing,
        "rel_l2_vs_fp64_256": round(err, 5),
        "rel_l2_per_seed": [round(x, 5) for x in den_err],
        "noisy_rel_l2": round(float(np.mean(noisy_err)), 5),
        "target": f"{TARGET_DIRS}-direction gather, reference step rule (no cap), FP64 on the GPU",
        "seeds": seeds,
        "plain_sweep": sweep,
        "plain_sweep_note": "plain path = 2 bilateral passes r=4, raw = no denoiser; 64 steps at 32 directions, "
                            "the reference step rule without cap above",
        "plain_ms_at_equal_error": round(eq_ms, 4) if eq_ms else None,
        "plain_ms_at_equal_error_by_family": {k: round(v, 4) for k, v in fam.items()},
        "speedup_at_equal_error": round(eq_ms / den_ms, 3) if eq_ms else None,
        "speedup_vs_plain32": round(plain32_ms / den_ms, 3),
        "weights": os.path.relpath(path, ROOT),
        "quality_config": {
            "workload": f"the same network on a {QDEN_DIRS}-direction {QDEN_GRID}x{QDEN_GRID}-pattern gather "
                        f"(footprint {dn.NOISY_FOOTPRINT}, no cap)",
            "ms_per_frame": round(q_ms, 4),
            "passes_ms": {f"{i}:{p['name']}": round(p["ms"], 4) for i, p in enumerate(q_passes)},
            "gather_tile": q_tile,
            "rel_l2_vs_fp64_256": round(qerr, 5),
            "rel_l2_per_seed": [round(x, 5) for x in q_err],
            "plain_ms_at_equal_error": round(q_eq, 4) if q_eq else None,
            "speedup_at_equal_error": round(q_eq / q_ms, 3) if q_eq else None,
        },
    }


def train_leg(args, views=64, steps=60, warmup=5):
    """Training throughput on the training configuration: batch 16 of 256x256
    crops, forward + backward + Adam on the project's kernels (pairs from a
    small set of seeded views; the throughput does not depend on the data)."""
    import torch
    from legitengine_amd import device as dev
    from legitengine_amd import denoiser as dn
    X, T = dn.build_dataset(views // 4, 4, log=lambda m: None)
    X = X[dn.NOISY_FOOTPRINT]
    net = dev.UNet(16, dn.CROP, dn.CROP, seed=0)
    net.autotune(16, dn.CROP, dn.CROP, backward=True)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0)
    batches = [(X.index_select(0, idx).contiguous(), T.index_select(0, idx).contiguous())
               for idx in (torch.randint(0, X.shape[0], (16,), device="cuda", generator=gen) for _ in range(8))]
    for i in range(warmup):
        net.train_step(*batches[i % 8], 1e-3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        net.train_step(*batches[i % 8], 1e-3)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    flops = 3 * dn.unet_flops_per_pixel() * 16 * dn.CROP * dn.CROP  # forward + input and weight gradients
    return {"workload": "U-Net training step: batch 16 of 256x256 crops, BF16 MFMA forward + backward, FP32 Adam",
            "steps_per_s": round(steps / secs, 2), "ms_per_step": round(secs * 1e3 / steps, 3),
            "mfma_roofline": {"achieved": round(flops * steps / secs / 1e12, 1), "peak": MFMA_BF16_PEAK_TFLOPS,
                              "unit": "TFLOP/s", "frac": round(flops * steps / secs / 1e12 / MFMA_BF16_PEAK_TFLOPS, 4),
                              "flops_per_step": flops},
            "gradient_check": "FP64 central finite differences on a 16x16 crop: tests/test_unet_oracle.py"}


def batch_resolution(seed):
    """The view's resolution, drawn uniformly from the four sizes by its seed."""
    import numpy as np
    return BATCH_SIZES[int(np.random.default_rng(seed).integers(len(BATCH_SIZES)))]


def batch_leg(args, sampler):
    """32 views of uneven resolution, each its own seeded scene: one captured
    hipGraph per batch (pyramids, ONE persistent gather launch over all views,
    the U-Net per view), launch shape from the autotuner."""
    import numpy as np
    import torch
    from legitengine_amd import device as dev
    from legitengine_amd import lgx
    from legitengine_amd import denoiser as dn
    views, sizes = [], []
    for s in BATCH_SEEDS:
        w, h = batch_resolution(s)
        sc = lgx.make_scene(s, 32, 8)
        view, proj = lgx.scene_matrices(sc, w, h)
        views.append((dev.GBufferDev.alloc(w, h, "mixed").render(sc), view, proj))
        sizes.append((w, h))
    maxw, maxh = max(w for w, _ in sizes), max(h for _, h in sizes)
    net = dev.UNet(1, maxh, maxw)
    path = args.weights or dn.WEIGHTS
    net.load(path)
    net.autotune(1, 1080, 1920)  # register tiles at the most common extent
    b = dev.Batch(views, dirs=DEN_DIRS, max_steps=dn.NOISY_STEPS, levels=LEVELS, sampler=sampler, grid=dn.NOISY_GRID,
                  footprint_dirs=dn.NOISY_FOOTPRINT, net=net)
    log(f"batch of {len(views)} views created")
    tune = b.autotune()
    log(f"autotune best {tune['best']}")
    stream = torch.cuda.Stream()
    b.capture(stream)
    for _ in range(max(1, args.warmup)):
        b.replay(stream)
    stream.synchronize()
    # throughput: back-to-back replays
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.replay(stream)
    stream.synchronize()
    batch_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # latency: per view, from the batch's start stamp to the view's stamp
    lat, stamped = [], []
    for _ in range(args.steps):
        b.replay(stream)
        one = b.latencies()  # synchronises; the stamps come from inside the replay
        lat.extend(one)
        stamped.append(max(one))
    pixels = sum(w * h for w, h in sizes)
    counts = {f"{w}x{h}": sizes.count((w, h)) for (w, h) in BATCH_SIZES}
    tile, wg = b.shape()
    b.close()
    # the same views one at a time through the single-view path (pyramids,
    # lgx_gather with the batch's launch shape, lgx_unet_denoise), smallest
    # first, eager on one stream: what the batch is measured against
    seq = []
    for i in sorted(range(len(views)), key=lambda i: sizes[i][0] * sizes[i][1]):
        (g, view, proj), (w, h) = views[i], sizes[i]
        ga = dev.Gather(w, h, DEN_DIRS, dn.NOISY_STEPS, sampler, grid=dn.NOISY_GRID, tile=tile,
                        footprint_dirs=dn.NOISY_FOOTPRINT)
        seq.append((g, view, proj, dev.PyramidDev(g, LEVELS, sampler), ga, dev.DImage(w, h, lgx.RGBA32F),
                    dev.DImage(w, h, lgx.RGBA32F)))

    def sequential():
        for g, view, proj, pyr, ga, noisy, den in seq:
            pyr.build(stream)
            ga(view, proj, g, pyr, out=noisy, stream=stream)
            net.denoise(g, noisy, out=den, stream=stream)

    sequential()
    stream.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        sequential()
    e1.record(stream)
    e1.synchronize()
    seq_ms = e0.elapsed_time(e1) / args.steps
    del seq
    return {
        "workload": f"{len(views)} views (seeds {BATCH_SEEDS[0]}-{BATCH_SEEDS[-1]}, own scenes), resolutions drawn "
                    "from {512x512, 1280x720, 1920x1080, 2560x1440}: mips, Hi-Z and pre-blur per view, one "
                    f"persistent {DEN_DIRS}-direction gather launch over all views, BF16 U-Net per view (smallest "
                    "first); the batch is one hipGraph replay",
        "views": counts,
        "megapixels": round(pixels / 1e6, 3),
        "batch_ms": round(batch_ms, 4),
        "mpixel_per_s": round(pixels / 1e6 / (batch_ms * 1e-3), 1),
        "sequential_ms": round(seq_ms, 4),
        "sequential_mpixel_per_s": round(pixels / 1e6 / (seq_ms * 1e-3), 1),
        "sequential_note": "the same views one at a time (pyramids, lgx_gather, lgx_unet_denoise), smallest first, "
                           "eager on one stream",
        "stamped_batch_ms": round(float(np.median(stamped)), 4),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4),
                       "min": round(float(min(lat)), 4)},
        "autotune": {"best": tune["best"], "cus": tune["cus"],
                     "table": [(r["tile"], r["workgroups"], round(r["ms"], 4)) for r in tune["table"]]},
        "shape": {"tile": tile, "workgroups": wg},
        "replays": args.steps,
    }


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(g, view, proj, sampler, gpu_gather, rows_stride):
    """The FP32 oracle on the GPU's own G-buffer bytes, the whole plain path:
    full-frame pyramids and pre-blur, the gather on a stratified sample of
    whole rows (every `rows_stride`-th row, the offset cycling through the
    stride so all pattern rows and both borders are covered; its time scaled
    by H / rows), and the full-frame bilateral passes. One warm-up run, then
    the median of three."""
    import numpy as np
    from oracle import binding as orc
    threads = orc.lib().orc_max_threads()
    depth = g.depth.numpy()
    octn = orc.widen(g.normal.numpy())
    radiance = orc.widen(g.radiance.numpy())
    n_rows = (H + rows_stride - 1) // rows_stride
    rows = sorted({min(H - 1, i * rows_stride + i % rows_stride) for i in range(n_rows)})
    runs = []
    for rep in range(4):
        t0 = time.perf_counter()
        dep_b, rad_b, _, _ = orc.build_pyramids(depth, radiance, LEVELS, sampler, True)
        t1 = time.perf_counter()
        ref, _ = orc.gather_rows(view, proj, depth, octn, dep_b, rad_b, DIRS, STEPS, sampler, rows)
        t2 = time.perf_counter()
        img = np.zeros((H, W, 4), np.float32)
        for _ in range(BILAT_PASSES):
            img = orc.bilateral(img, depth, octn, BILAT_RADIUS)
        t3 = time.perf_counter()
        frame_s = (t1 - t0) + (t2 - t1) * H / len(rows) + (t3 - t2)
        log(f"cpu baseline run {rep}: {frame_s * 1e3:.0f} ms/frame (gather rows {(t2 - t1) * 1e3:.0f} ms)")
        if rep > 0:
            runs.append((frame_s, t1 - t0, (t2 - t1) * H / len(rows), t3 - t2))
    runs.sort()
    med = runs[1]
    return {"value": round(med[0] * 1e3, 1), "unit": "ms/frame", "cores": int(threads), "kind": "port",
            "cpu_model": cpu_model(), "omp_max_threads": int(threads),
            "sample": f"FP32 oracle on the GPU's G-buffer, whole plain path: full-frame pyramids and pre-blur, "
                      f"{BILAT_PASSES} full-frame bilateral passes, gather on {len(rows)} stratified whole rows "
                      f"(every {rows_stride}th, offset cycling) scaled by {H / len(rows):.2f}; 1 warm-up, "
                      f"median of 3",
            "runs_ms": [round(r[0] * 1e3, 1) for r in runs],
            "split_ms": {"pyramids": round(med[1] * 1e3, 1), "gather": round(med[2] * 1e3, 1),
                         "bilateral": round(med[3] * 1e3, 1)},
            "gather_rel_l2_on_rows": orc.rel_l2(gpu_gather[rows], ref)}


if __name__ == "__main__":
    main()
