"""Train the U-Net denoiser on images rendered by this library.

Data: `scenes.training_scene(i)` (randomised ground / sphere / medium / mesh
scenes with simple or HDRI skies, in the style of config1 / config5) and, for a
fraction `--hdri-mesh` of them, `scenes.training_scene_hdri_mesh(i)` (displaced
icospheres under an HDRI with a small bright sun, in the style of config2),
rendered at `--res` (256 x 256): a target of `--target-passes` samples with the
MIS estimator (the same expectation as the reference's, far less sun noise) and
noisy inputs of 1, 2, 4 and 8 samples with an independent seed, with the
reference's estimator for half of the scenes and the MIS estimator for the
other half.  Inputs are beauty, DIFFUSE as albedo and WORLD_NORMAL as normal
(what the reference hands to its denoiser, csrc/engine/src/engine.cpp:25-28)
and the DEPTH AOV.  Every render uses a pattern of its own sample count.
`--workers` render contexts render different scenes concurrently (one host
thread each; the library releases the GIL in every call).

Training: `--crop` x `--crop` crops (the whole 256 x 256 frame by default) with
random flips, batch `--batch`, L1 on log1p radiance, Adam (lr 1e-3, decayed to
1e-4 over the last 30% of the steps or of `--train-budget`); everything runs in libnectar_hip's kernels,
torch only holds device buffers.

    python -m nectarrender_amd.train_denoiser --scenes 2000 --target-passes 4096 --batch 16 --steps 8000
"""
from __future__ import annotations

import argparse
import sys
import threading
import time

import numpy as np

INPUT_PASSES = (1, 2, 4, 8)


def render_set(ctx, abi, scenes, i, res, target_passes, torch, hdri_mesh_frac=0.0):
    """Target (3,H,W) and inputs (K,10,H,W) = [beauty, albedo, normal, depth] of training scene i."""
    from .quality import with_spp
    mesh_scene = hdri_mesh_frac > 0.0 and (i * 0.6180339887) % 1.0 < hdri_mesh_frac
    gen = scenes.training_scene_hdri_mesh if mesh_scene else scenes.training_scene
    s, cam, rp = gen(i, res, spp=target_passes)
    ctx.upload(s)
    ctx.set_camera(0, with_spp(cam, target_passes), rp["seed"])
    ctx.reset(0)
    ctx.render(0, abi.render_desc(0, target_passes, rp["depth"], rp["seed"], mode=abi.MODE_FAST))
    target = torch.from_numpy(ctx.read(0)).cuda()
    seed2 = (rp["seed"] * 7 + 12345) & 0x7FFFFFFF
    mode = abi.MODE_FAST if i % 2 else abi.MODE_PARITY
    p0 = 0 if mode == abi.MODE_FAST else 1
    noisy = []
    for k in INPUT_PASSES:
        ctx.set_camera(0, with_spp(cam, k + p0), seed2)
        ctx.reset(0)
        ctx.render(0, abi.render_desc(p0, k + p0, rp["depth"], seed2, mode=mode))
        noisy.append(torch.cat([torch.from_numpy(ctx.read(0, L)).cuda()
                                for L in (abi.LAYER_BEAUTY, abi.LAYER_DIFFUSE, abi.LAYER_WORLD_NORMAL,
                                          abi.LAYER_DEPTH)]))
    return target, torch.stack(noisy)


def render_dataset(args, torch):
    """Render scenes 0..args.scenes-1 with `args.workers` concurrent contexts (scene i on
    worker i mod workers), stopping at the render budget; returns the scenes in index order
    up to the first one missing."""
    from . import abi, native, scenes
    results = {}
    lock = threading.Lock()
    t0 = time.time()
    errors = []

    def work(w):
        try:
            ctx = native.Context(0)
            # small images: more samples in flight
            ctx.set_tuning(slots=8, stream_groups=args.stream_groups, shade_wg_per_cu=8)
            for i in range(w, args.scenes, args.workers):
                if args.render_budget > 0 and time.time() - t0 > args.render_budget:
                    break
                r = render_set(ctx, abi, scenes, i, args.res, args.target_passes, torch, args.hdri_mesh)
                with lock:
                    results[i] = r
                    n = len(results)
                if n % 32 == 0:
                    print(f"data: {n}/{args.scenes} scenes rendered, {time.time() - t0:.1f} s", flush=True)
            ctx.close()
        except Exception as e:                    # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=work, args=(w,)) for w in range(args.workers)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    n = 0
    while n in results:
        n += 1
    if n < args.scenes:
        print(f"data: render budget reached, {n} scenes in order", flush=True)
    targets = torch.stack([results[i][0] for i in range(n)])     # (S,3,H,W)
    noisy = torch.stack([results[i][1] for i in range(n)])       # (S,K,10,H,W)
    return targets, noisy, time.time() - t0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=2000)
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--target-passes", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--crop", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--hdri-mesh", type=float, default=0.3, help="fraction of config2-style scenes")
    ap.add_argument("--workers", type=int, default=4, help="render contexts running concurrently")
    ap.add_argument("--stream-groups", type=int, default=2, help="stream groups of each render context")
    ap.add_argument("--render-budget", type=float, default=0.0,
                    help="stop rendering scenes after this many seconds (0: render all)")
    ap.add_argument("--train-budget", type=float, default=0.0,
                    help="stop training after this many seconds (0: run every step)")
    args = ap.parse_args(argv)

    import torch

    from .denoise import WEIGHTS, UNet

    targets, noisy, t_data = render_dataset(args, torch)
    S, K, R, c = targets.shape[0], noisy.shape[1], args.res, args.crop
    print(f"data ready: {S} scenes x {K} noise levels, {t_data:.1f} s", flush=True)

    net = UNet(0, seed=args.seed)
    g = torch.Generator(device="cpu").manual_seed(args.seed)
    t1 = time.time()
    hist = []
    steps = args.steps
    for step in range(args.steps):
        si = torch.randint(0, S, (args.batch,), generator=g)
        ki = torch.randint(0, K, (args.batch,), generator=g)
        ys = torch.randint(0, R - c + 1, (args.batch,), generator=g)
        xs = torch.randint(0, R - c + 1, (args.batch,), generator=g)
        fl = torch.randint(0, 4, (args.batch,), generator=g)
        xb, tg = [], []
        for b in range(args.batch):
            y0, x0 = int(ys[b]), int(xs[b])
            crop = noisy[si[b], ki[b], :, y0:y0 + c, x0:x0 + c]
            tc = targets[si[b], :, y0:y0 + c, x0:x0 + c]
            if fl[b] & 1:
                crop, tc = crop.flip(-1), tc.flip(-1)
            if fl[b] & 2:
                crop, tc = crop.flip(-2), tc.flip(-2)
            xb.append(crop)
            tg.append(tc)
        xb = torch.stack(xb)
        frac = step / max(1, args.steps - 1)
        if args.train_budget > 0:                 # the schedule follows whichever runs out first
            frac = max(frac, (time.time() - t1) / args.train_budget)
        lr = args.lr if frac < 0.7 else args.lr * (0.1 ** ((frac - 0.7) / 0.3))
        loss = net.train_step(xb[:, 0:3].contiguous(), xb[:, 3:6].contiguous(), xb[:, 6:9].contiguous(),
                              xb[:, 9:10].contiguous(), torch.stack(tg).contiguous(), lr)
        hist.append(loss)
        if (step + 1) % 200 == 0:
            print(f"step {step + 1}/{args.steps}: loss {np.mean(hist[-200:]):.4f}, lr {lr:.2e}, "
                  f"{(time.time() - t1) / (step + 1) * 1e3:.1f} ms/step", flush=True)
        if args.train_budget > 0 and time.time() - t1 > args.train_budget:
            steps = step + 1
            print(f"train budget reached after {steps} steps", flush=True)
            break
    out = args.out or WEIGHTS
    net.save(out)
    print(f"saved {out} ({net.n_params} parameters): {S} scenes, {args.target_passes}-sample targets, "
          f"{steps} steps of {args.batch} crops of {c}x{c}, training {time.time() - t1:.1f} s", flush=True)


if __name__ == "__main__":
    sys.exit(main())
