"""Train the U-Net denoiser on images rendered by this library.

Data: `scenes.training_scene(i)` (randomised ground / sphere / medium / mesh
scenes with simple or HDRI skies, in the style of config1 / config5) and, for a
fraction `--hdri-mesh` of them, `scenes.training_scene_hdri_mesh(i)` (displaced
icospheres under an HDRI with a small bright sun, in the style of config2),
rendered at 256x256: a target of `--target-passes` samples with the MIS
estimator (the same expectation as the reference's, far less sun noise) and
noisy inputs of 1, 2, 4 and 8 samples with an independent seed, with the
reference's estimator for half of the scenes and the MIS estimator for the
other half (beauty, DIFFUSE as albedo, WORLD_NORMAL as normal, the inputs the
reference hands to its denoiser, csrc/engine/src/engine.cpp:25-28).  Every
render uses a pattern of its own sample count.  Training: random 128x128 crops
with flips, batch `--batch`, L1 on log1p radiance, Adam (lr 1e-3, decayed to
1e-4 over the last 30% of steps); everything runs in libnectar_hip's kernels,
torch only holds device buffers.

    python -m nectarrender_amd.train_denoiser --scenes 2000 --target-passes 2048 --batch 16 --steps 8000
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np


def render_set(ctx, abi, scenes, i, res, target_passes, input_passes, torch, hdri_mesh_frac=0.0):
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
    for k in input_passes:
        ctx.set_camera(0, with_spp(cam, k + p0), seed2)
        ctx.reset(0)
        ctx.render(0, abi.render_desc(p0, k + p0, rp["depth"], seed2, mode=mode))
        noisy.append(torch.stack([torch.from_numpy(ctx.read(0, L)).cuda()
                                  for L in (abi.LAYER_BEAUTY, abi.LAYER_DIFFUSE, abi.LAYER_WORLD_NORMAL)]))
    return target, torch.stack(noisy)          # (3,H,W), (K,3,3,H,W)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=192)
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--target-passes", type=int, default=1023)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--hdri-mesh", type=float, default=0.3, help="fraction of config2-style scenes")
    ap.add_argument("--render-budget", type=float, default=0.0,
                    help="stop rendering scenes after this many seconds (0: render all)")
    args = ap.parse_args(argv)

    import torch

    from . import abi, native, scenes
    from .denoise import WEIGHTS, UNet

    input_passes = (1, 2, 4, 8)
    ctx = native.Context(0)
    ctx.set_tuning(slots=8, stream_groups=2, shade_wg_per_cu=8)    # small images: more samples in flight
    t0 = time.time()
    targets, noisy = [], []
    for i in range(args.scenes):
        t, n = render_set(ctx, abi, scenes, i, args.res, args.target_passes, input_passes, torch, args.hdri_mesh)
        targets.append(t)
        noisy.append(n)
        if (i + 1) % 16 == 0:
            print(f"data: {i + 1}/{args.scenes} scenes rendered, {time.time() - t0:.1f} s", flush=True)
        if args.render_budget > 0 and time.time() - t0 > args.render_budget:
            print(f"data: render budget reached after {i + 1} scenes", flush=True)
            break
    ctx.close()
    args.scenes = len(targets)
    targets = torch.stack(targets)             # (S,3,H,W)
    noisy = torch.stack(noisy)                 # (S,K,3,3,H,W)
    print(f"data ready: {args.scenes} scenes x {len(input_passes)} noise levels, {time.time() - t0:.1f} s", flush=True)

    net = UNet(0, seed=args.seed)
    g = torch.Generator(device="cpu").manual_seed(args.seed)
    S, K, R, c = args.scenes, len(input_passes), args.res, args.crop
    t1 = time.time()
    hist = []
    for step in range(args.steps):
        si = torch.randint(0, S, (args.batch,), generator=g)
        ki = torch.randint(0, K, (args.batch,), generator=g)
        ys = torch.randint(0, R - c + 1, (args.batch,), generator=g)
        xs = torch.randint(0, R - c + 1, (args.batch,), generator=g)
        fl = torch.randint(0, 4, (args.batch,), generator=g)
        bs, al, nm, tg = [], [], [], []
        for b in range(args.batch):
            y0, x0 = int(ys[b]), int(xs[b])
            crop = noisy[si[b], ki[b], :, :, y0:y0 + c, x0:x0 + c]
            tc = targets[si[b], :, y0:y0 + c, x0:x0 + c]
            if fl[b] & 1:
                crop, tc = crop.flip(-1), tc.flip(-1)
            if fl[b] & 2:
                crop, tc = crop.flip(-2), tc.flip(-2)
            bs.append(crop[0])
            al.append(crop[1])
            nm.append(crop[2])
            tg.append(tc)
        frac = step / max(1, args.steps - 1)
        lr = args.lr if frac < 0.7 else args.lr * (0.1 ** ((frac - 0.7) / 0.3))
        loss = net.train_step(torch.stack(bs).contiguous(), torch.stack(al).contiguous(),
                              torch.stack(nm).contiguous(), torch.stack(tg).contiguous(), lr)
        hist.append(loss)
        if (step + 1) % 200 == 0:
            print(f"step {step + 1}/{args.steps}: loss {np.mean(hist[-200:]):.4f}, lr {lr:.2e}, "
                  f"{(time.time() - t1) / (step + 1) * 1e3:.1f} ms/step", flush=True)
    out = args.out or WEIGHTS
    net.save(out)
    print(f"saved {out} ({net.n_params} parameters), training {time.time() - t1:.1f} s", flush=True)


if __name__ == "__main__":
    sys.exit(main())
