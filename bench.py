"""Benchmark: Mpaths/s of the path-tracing sample pass on one MI355X.

Workload: 1920x1080 view (resolution (1080, 1920) in the reference's (H, W)
order) of a 1,003,520-triangle displaced icosphere (PBR metal, roughness 0.4)
on a Lambertian quad under a procedural 2048x1024 fp32 HDRI with a 1e4-radiance
0.5 degree sun; depth 12, Russian roulette from bounce 3, seed 0, MIS on
(fast mode: next-event estimation to the HDRI, power heuristic); synthetic,
seeded scene data.  A step is one sample pass over all 2,073,600 pixels
(one path per pixel); K steps render passes 0..K-1, the 64-spp image is K = 64.
The same passes with the reference's estimator (passes 1..K, bit-exact with
the oracle) are reported beside it.

Output: one JSON line with the contract fields plus
  roofline          traversal (k_extend) bytes / traversal time against HBM
                    peak, with the algorithmic byte model counted by the oracle
  parity_estimator  Mpaths/s and kernel times of the reference's estimator
  time_to_rmse      seconds to RMSE 0.01 on this scene, plain and U-Net denoised;
                    time_to_rmse_end_to_end the same on the 500-sphere scene
  cpu_baseline      the CPU oracle (the reference's estimator) on this host's
                    cores, on full 1080p passes for ~12 s
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpaths/s @1080p; secs to RMSE≤0.01 vs 4096spp ref (plain vs denoised)"
W, H = 1920, 1080
PATHS_PER_STEP = W * H
# Traversal work per path of this workload, counted by oracle/traversal_work.py
# on every 61st pixel over sample passes 1..4 (135,976 paths): 1.7945 segments,
# 23.873 interior nodes (64 B), 1.8217 triangle tests (48 B), 1.9960 top-level
# objects (128 B), 40 B ray/hit per segment.
TRAVERSAL_BYTES_PER_PATH = 1942.58
# The same model per traced segment (ray): the path's 24 accumulate bytes aside,
# (1942.58 - 24) / 1.7945 = 1069.14 B.  k_extend traces the segments of the
# streaming iterations; once few paths are live, tail mode (k_tail) traces the
# rest, so k_extend's algorithmic bytes are its traced rays (nr_last_timings[7])
# times this.
TRAVERSAL_BYTES_PER_SEGMENT = (TRAVERSAL_BYTES_PER_PATH - 24.0) / 1.7945
# The scattered rays of the MIS estimator follow the same distribution (next-event
# estimation adds shadow rays, traced by k_shadow, and changes no scattering
# decision's probability), so the same model applies to its k_extend.
# HBM bytes k_extend moves per traced ray on this workload, from rocprofv3 FETCH_SIZE (scaled by the
# factor a 1 GiB device-to-device copy gives) + WRITE_SIZE over the 64 passes under the tuning the
# autotuner picks (64 slots, 2 groups, 8 shading workgroups per CU, tail below 1,048,576 paths):
# profiles/k_extend_c2_1080p.md.
TRAVERSAL_HBM_BYTES_PER_SEGMENT = {0: 483.0, 1: 461.3}  # keyed by estimator (abi.MODE_PARITY, MODE_FAST)
HBM_PEAK_GBS = 8000.0


def cpu_baseline(budget_s: float = 12.0, max_passes: int = 63):
    """The oracle's render pass over the same workload, timed on the host cores."""
    from nectarrender_amd import abi, scenes
    from oracle import pyoracle as po
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    s, cam, rp = scenes.config2(spp=64)
    os_ = po.OracleScene(s)
    oc = po.OracleCamera(cam, rp["seed"])
    d = abi.render_desc(1, 64, rp["depth"], rp["seed"])
    t0 = time.perf_counter()
    n = 0
    while n < max_passes and (time.perf_counter() - t0) < budget_s:
        po.render_pass(os_, oc, d, 1 + n, threads=threads)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": round(n * PATHS_PER_STEP / dt / 1e6, 3), "unit": "Mpaths/s", "cores": threads,
            "kind": "port",
            "sample": f"{n} full 1920x1080 sample passes ({n * PATHS_PER_STEP} paths) of the same scene, "
                      f"{dt:.1f} s, OpenMP over 16x16 tiles"}


def quality_section(log, ctx=None, scene=None, end_to_end=False):
    """Seconds to RMSE <= 0.01 (display-referred) vs a 4096-sample float64 reference, plain and denoised:
    on the benchmark scene with MIS (ctx holds it), or (end_to_end) on the 1080p three-sphere scene with
    500 media/dielectric spheres and the reference's estimator (4032 samples)."""
    import torch

    from nectarrender_amd import abi, native, quality, scenes
    from nectarrender_amd.denoise import WEIGHTS, UNet
    own = ctx is None
    if end_to_end:
        s, cam, rp = scenes.config5(spp=64)
        workload = ("1920x1080 (H=1080, W=1920) three-sphere scene + 500 constant-medium/dielectric spheres, "
                    "depth 8, reference estimator, white balance 6500 K tint 0, ACES, sqrt gamma")
        ref_desc = "64 renders x 63 passes (1..63 of 64), float64 mean, disjoint seeds"
    else:
        s, cam, rp = scene
        rp = dict(rp, mode=abi.MODE_FAST)
        workload = ("the benchmark scene (1080p, 1,003,520-triangle icosphere, 2048x1024 HDRI with 1e4 sun, "
                    "depth 12, MIS on), white balance 6500 K tint 0, ACES, sqrt gamma")
        ref_desc = "64 renders x 64 passes (0..63 of 64), float64 mean, disjoint seeds"
    if own:
        ctx = native.Context(0)
        ctx.upload(s)
    ref_disp, _ = quality.reference(ctx, abi, torch, cam, rp, log=log)
    unet = UNet(0).load() if os.path.exists(WEIGHTS) else None
    plain, den = quality.curves(ctx, abi, torch, cam, rp, ref_disp, unet)
    out = {"workload": workload, "target_rmse": quality.TARGET, "reference": ref_desc,
           "plain_s": quality.time_to_target(plain),
           "plain": [[k, round(t, 5), round(e, 5)] for k, t, e in plain]}
    if unet is not None:
        pts = [(k, tr + td, e) for k, tr, td, e in den]
        out["denoised_s"] = quality.time_to_target(pts)
        out["denoised"] = [[k, round(tr, 5), round(td, 5), round(e, 5)] for k, tr, td, e in den]
        # denoiser throughput at this size (HIP events around the U-Net launch sequence)
        if end_to_end:
            # 4 samples + U-Net with the per-tile fallback (64^2 tiles whose error estimate exceeds 0.02
            # restart with 64 samples and are denoised again), timed end to end
            from nectarrender_amd.adaptive import render_denoised_adaptive
            H, W = int(cam.resolution[0]), int(cam.resolution[1])
            frame = torch.empty((3, H, W), dtype=torch.float32, device="cuda")
            render_denoised_adaptive(ctx, unet, 0, cam, rp, quality.display_desc(abi), frame, seed=quality.TEST_SEED)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            info = render_denoised_adaptive(ctx, unet, 0, cam, rp, quality.display_desc(abi), frame,
                                            seed=quality.TEST_SEED)
            torch.cuda.synchronize()
            ta = time.perf_counter() - t0
            d = quality._display(ctx, abi, torch, frame, H, W)
            out["denoised_fallback"] = {
                "seconds": round(ta, 5), "rmse": round(quality.rmse(d, ref_disp), 5),
                "fallback_tiles": info["fallback_tiles"], "tiles": info["tiles"],
                "max_tile_rmse_estimate": round(float(info["tile_rmse"].max()), 5),
                "note": "4 samples + U-Net; two 4-sample plain renders judge each 64x64 tile; tiles above 0.02 "
                        "re-render with 64 samples and the U-Net runs again"}
        ms, flops = unet.last_stats()
        out["denoiser"] = {"ms": round(ms, 3), "tflops": round(flops / ms / 1e9, 2), "peak_tflops": 2500.0,
                           "frac": round(flops / ms / 1e9 / 2500.0, 4), "flops_per_frame": flops,
                           "dtype": "bf16 (fp32 accumulation)"}
        unet.close()
    else:
        out["denoised_s"] = None
        out["denoised_note"] = "no trained weights at " + WEIGHTS
    if out["plain_s"] and out["denoised_s"]:
        out["speedup"] = round(out["plain_s"] / out["denoised_s"], 2)
    four = [(k, tr + td, e) for k, tr, td, e in den if k == 4] if unet is not None else []
    if four and out["plain_s"]:
        # the 4-sample + U-Net frame on its own: its time, its error, and the speedup over the
        # plain path tracer's time to the same error
        k, t4, e4 = four[0]
        same = quality.time_to_target(plain, target=e4)
        out["denoised_4spp"] = {"seconds": round(t4, 5), "rmse": round(e4, 5),
                                "plain_s_to_same_rmse": same, "speedup": round(same / t4, 2) if same else None}
    if own:
        ctx.close()
    return out


def views_section(log):
    """The 16-view configuration (10,000 spheres of every material, views of 256x256 .. 1920x1080,
    10,142,016 pixels, 32 samples: passes 1..31, depth 16, reference estimator): all views in one
    nr_render_views call against one nr_render call per view, on the same passes."""
    from nectarrender_amd import abi, native, scenes
    s, cams, rp = scenes.config3()
    ctx = native.Context(0)
    ctx.upload(s)
    views = list(range(len(cams)))
    for v, cam in enumerate(cams):
        ctx.set_camera(v, cam, rp["seed"] + v)
    px = sum(int(c.resolution[0]) * int(c.resolution[1]) for c in cams)
    d = abi.render_desc(1, 32, rp["depth"], rp["seed"])
    ctx.render_views(views, abi.render_desc(1, 2, rp["depth"], rp["seed"]))   # warmup pass
    out = {"workload": f"16 views, {px} pixels, 10,000 spheres (PBR, dielectric, Lambertian, emissive, "
                       "constant media), passes 1..31, depth 16, reference estimator"}
    for name in ("one_call", "per_view"):
        for v in views:
            ctx.reset(v)
        ctx.sync()
        t0 = time.perf_counter()
        if name == "one_call":
            ctx.render_views(views, d)
        else:
            for v in views:
                ctx.render(v, d)
        ctx.sync()
        dt = time.perf_counter() - t0
        out[name] = {"seconds": round(dt, 4), "Mpaths_s": round(px * 31 / dt / 1e6, 1)}
        log(f"views: {name} {dt:.3f} s")
    out["speedup"] = round(out["per_view"]["seconds"] / out["one_call"]["seconds"], 2)
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-quality", action="store_true", help="skip the seconds-to-RMSE measurement")
    ap.add_argument("--streams", type=int, default=0,
                    help="stream groups per render (1..4; default: the library's default)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the library's default tuning")
    ap.add_argument("--tuning", default="",
                    help="slots,groups,shade,tail[,graph iterations]: set this tuning instead of autotuning")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one 1 GiB device-to-device copy (a kernel of known traffic for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards by "
              f"tiles or samples, with one final sum of the per-GPU AOVs as the only exchange)", file=sys.stderr)
        sys.exit(2)
    K, Wm = args.steps, args.warmup

    if not args.no_quality or args.calibrate_copy:
        import torch                      # its HIP runtime initialises before the library's
        torch.cuda.init()
    from nectarrender_amd import abi, native, scenes
    from nectarrender_amd.quality import with_spp

    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    scene, cam, rp = scenes.config2(spp=64)
    ctx = native.Context(0)
    # scene upload (outside the Mpaths/s region): mesh tree on the host, then on the GPU (same tree)
    build = {}
    for name, mode in (("host_tree_s", abi.BUILD_HOST), ("device_tree_s", abi.BUILD_DEVICE)):
        ctx.set_build_mode(mode)
        tb = time.perf_counter()
        ctx.upload(scene)
        ctx.sync()
        build[name] = round(time.perf_counter() - tb, 4)
    tuning = None
    if args.tuning:
        ctx.set_tuning(*[int(v) for v in args.tuning.split(",")])
    elif args.streams:
        ctx.set_render_streams(args.streams)
    elif not args.no_autotune:
        # the library's autotuner on the K MIS passes the timed region renders (path slots per pixel
        # above the pass count cannot show their effect): path slots, stream groups, shading grid,
        # tail threshold, graph capture (the result does not depend on them)
        ctx.set_camera(0, with_spp(cam, K), rp["seed"])
        tuning = ctx.autotune(0, abi.render_desc(0, K, rp["depth"], rp["seed"], mode=abi.MODE_FAST))

    def measure(mode):
        """K timed passes (fast mode: 0..K-1 of a K-sample pattern; the reference's estimator: 1..K of
        K + 1), then the same passes again with per-kernel HIP events on one stream group."""
        p0 = 0 if mode == abi.MODE_FAST else 1
        ctx.set_camera(0, with_spp(cam, K + p0), rp["seed"])
        if Wm > 0:
            ctx.render(0, abi.render_desc(p0, p0 + min(Wm, K), rp["depth"], rp["seed"], mode=mode))
        ctx.sync()
        ctx.reset(0)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.render(0, abi.render_desc(p0, p0 + K, rp["depth"], rp["seed"], mode=mode))
        ctx.sync()
        t1 = time.perf_counter()
        img = ctx.read(0)
        ctx.reset(0)
        ctx.set_timing(True)
        ctx.render(0, abi.render_desc(p0, p0 + K, rp["depth"], rp["seed"], mode=mode))
        ctx.sync()
        # [total, k_extend, shading, generate+accumulate, iterations, k_shadow, k_tail, rays k_extend traced]
        tm = ctx.timings()
        ctx.set_timing(False)
        if not np.array_equal(img, ctx.read(0)):
            raise RuntimeError("the timed and the event-timed renders differ")
        paths = K * PATHS_PER_STEP
        n_extend = int(tm[4])    # one k_extend launch per streaming iteration
        traced = float(tm[7])
        alg_bytes = TRAVERSAL_BYTES_PER_SEGMENT * traced
        achieved = alg_bytes / (tm[1] / 1e3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": round(TRAVERSAL_HBM_BYTES_PER_SEGMENT[mode] * traced / n_extend),
                "kernel": "k_extend", "launches": n_extend, "avg_launch_ms": round(tm[1] / n_extend, 4),
                "bytes_per_launch": round(alg_bytes / n_extend), "rays_traced": int(traced),
                "segments_per_path_in_k_extend": round(traced / paths, 4)}
        gpu = {"total": round(tm[0], 3), "traversal": round(tm[1], 3), "shading": round(tm[2], 3),
               "generate_accumulate": round(tm[3], 3)}
        if mode == abi.MODE_FAST:
            gpu["shadow_rays"] = round(tm[5], 3)
        if len(tm) > 6:
            gpu["tail"] = round(tm[6], 3)
        gpu["note"] = ("event-timed rerun of the same passes on one stream group; the Mpaths/s figure is the "
                       "default two stream groups")
        return paths / (t1 - t0) / 1e6, (t1 - t0), gpu, roof, img

    value, dt, gpu_ms, roof, img = measure(abi.MODE_FAST)
    par_value, par_dt, par_gpu, par_roof, _ = measure(abi.MODE_PARITY)
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpaths/s", "n_gpus": 1, "steps": K,
        "warmup": Wm, "ms_per_step": round(dt / K * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "1920x1080 (H=1080, W=1920), 1,003,520-triangle displaced icosphere PBR metal "
                               "rough 0.4 on Lambertian quad, 2048x1024 fp32 HDRI with 1e4 sun, depth 12, "
                               "MIS on (next-event estimation to the HDRI, power heuristic), "
                               "one path per pixel per step",
                   "spp_passes": K, "ray_depth": rp["depth"], "paths_per_step": PATHS_PER_STEP},
        "gpu_ms": gpu_ms,
        "image_mean": float(img.mean()),
        "scene_upload": dict(build, triangles=1_003_520),
        "roofline": roof,
        "tuning": tuning or ctx.get_tuning(),
        "parity_estimator": {"value": round(par_value, 3), "unit": "Mpaths/s",
                             "ms_per_step": round(par_dt / K * 1e3, 4),
                             "note": "the reference's estimator (passes 1..K, bit-exact with the oracle)",
                             "gpu_ms": par_gpu, "roofline": par_roof},
    }
    if not args.no_quality:
        log = lambda m: print(m, file=sys.stderr, flush=True)   # noqa: E731
        out["time_to_rmse"] = quality_section(log, ctx, (scene, cam, rp))
        out["time_to_rmse_end_to_end"] = quality_section(log, end_to_end=True)
        out["batched_views"] = views_section(log)
    ctx.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
