"""A denoised frame with per-tile fallback to more samples.

1. Render the view with `k` samples and run the U-Net on it (D).
2. Judge it with the plain path tracer: two more independent `k`-sample renders
   of the same camera (J1, J2, each with its own seed, in judge views).  Per
   `tile`^2 tile, `nr_tile_error` estimates D's mean squared error in display
   space from (D - J1)^2 - (J1 - J2)^2 / 2 (high by the judges' own display-space
   bias: about 2x the MSE of a 4-sample U-Net frame on the 500-sphere scene).
3. Tiles whose estimated RMSE exceeds `threshold` restart their accumulation
   with `fallback` samples (`nr_render_tiles`; the other tiles keep their `k`
   samples), and the U-Net runs again on the frame.

The display transform is the one the error target is stated in (white balance,
tonemap, gamma; `nr_post_device`).  Images are torch device tensors (device
memory only); every computation runs in the library's kernels.
"""
from __future__ import annotations

import numpy as np

from . import abi
from .quality import with_spp


def first_pass(mode: int) -> int:
    return 0 if mode == abi.MODE_FAST else 1


def render_denoised_adaptive(ctx, unet, view: int, cam, rp: dict, post: abi.NrPostDesc, out, *, k: int = 4,
                             threshold: float = 0.02, tile: int = 64, fallback: int = 64, seed: int = 1,
                             judges=(14, 15)) -> dict:
    """Renders view `view` (camera `cam`, depth and mode from `rp`) and writes the denoised frame into the
    3 x H x W float32 device tensor `out`.  Returns the tile error estimates (RMSE) and the tiles that fell
    back.  The judge views are (re)configured with the same camera."""
    import torch

    H, W = int(cam.resolution[0]), int(cam.resolution[1])
    mode = rp.get("mode", abi.MODE_PARITY)
    p0 = first_pass(mode)
    ctx.set_camera(view, with_spp(cam, k + p0), seed)
    ctx.reset(view)
    ctx.render(view, abi.render_desc(p0, k + p0, rp["depth"], seed, mode=mode))
    for j, js in zip(judges, (seed + 7919, seed + 104729)):
        ctx.set_camera(j, with_spp(cam, k + p0), js)
        ctx.reset(j)
        ctx.render(j, abi.render_desc(p0, k + p0, rp["depth"], js, mode=mode))
    unet.denoise_view(ctx, view, out)
    torch.cuda.synchronize()
    disp = [torch.empty((3, H, W), dtype=torch.float32, device="cuda") for _ in range(3)]
    ctx.post_device(out.data_ptr(), H, W, post, 0, disp[0].data_ptr())
    for j, dst in zip(judges, disp[1:]):
        ctx.post_device(ctx.aov_device_ptr(j, abi.LAYER_BEAUTY), H, W, post, 0, dst.data_ptr())
    rows, cols = (H + tile - 1) // tile, (W + tile - 1) // tile
    err = torch.empty(rows * cols, dtype=torch.float32, device="cuda")
    ctx.tile_error(disp[0].data_ptr(), disp[1].data_ptr(), disp[2].data_ptr(), H, W, tile, err.data_ptr())
    rmse = np.sqrt(np.maximum(err.cpu().numpy(), 0.0)).reshape(rows, cols)
    bad = rmse > threshold
    if bad.any():
        ctx.set_camera(view, with_spp(cam, fallback + p0), seed)
        ends = np.where(bad, fallback + p0, 0).astype(np.uint32)
        ctx.render_tiles(view, abi.render_desc(p0, fallback + p0, rp["depth"], seed, mode=mode), ends, tile)
        unet.denoise_view(ctx, view, out)
        torch.cuda.synchronize()
    return {"tile_rmse": rmse, "fallback_tiles": int(bad.sum()), "tiles": int(bad.size)}
