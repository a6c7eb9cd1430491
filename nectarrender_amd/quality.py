"""Time to a target image error, plain and denoised.

The error is the RMSE over all pixels and RGB of the display-referred float
image (white balance at 6500 K with zero tint, ACES, sqrt gamma, clamp to
[0, 1], i.e. the float image process_stream_kernel produces before the x255,
csrc/data/src/stream.cu:146-159) against a reference of 64 independent
64-sample renders (seeds disjoint from the measured renders) averaged in
float64: 4032 sample passes per pixel.

Plain: render K passes for K = 1, 2, 4, ... and time the render; the time to
reach the target is interpolated log-log between the bracketing K.  Every
render is what the engine produces for a K-sample image: with the reference's
estimator samples_per_pixel = K + 1 and passes 1..K (engine.cpp:43), in fast
mode samples_per_pixel = K and passes 0..K-1.  Pass s takes x-stratum s of the
N-rooks pattern (camera.h:45-61), so K passes of a larger pattern would cover
only part of each pixel and the error would stall at the missing antialiasing
instead of falling with K.
Denoised: render K passes, run the U-Net (nr_denoise), time both.
rp["mode"] selects the estimator of every render (abi.MODE_PARITY when absent).
"""
from __future__ import annotations

import time

import numpy as np

REF_SEED0 = 1_000_000
TEST_SEED = 77
TARGET = 0.01


def display_desc(abi):
    return abi.post_desc(linear_to_gamma=True, apply_white_balance=True, wb_temperature=6500.0, wb_tint=0.0,
                         apply_tonemapping=True, tm_method=2)


def _display(ctx, abi, torch, img_chw, H, W):
    src = torch.as_tensor(np.ascontiguousarray(img_chw, np.float32)).cuda() if isinstance(img_chw, np.ndarray) \
        else img_chw.contiguous()
    out = torch.empty((3, H, W), dtype=torch.float32, device="cuda")
    ctx.post_device(src.data_ptr(), H, W, display_desc(abi), 0, out.data_ptr())
    return out


def with_spp(cam, n: int):
    """Copy of a camera descriptor with an n-sample pattern."""
    c = type(cam).from_buffer_copy(cam)
    c.samples_per_pixel = int(n)
    return c


def first_pass(abi, rp) -> int:
    """0 in fast mode, 1 with the reference's estimator (engine.cpp:43)."""
    return 0 if rp.get("mode", abi.MODE_PARITY) == abi.MODE_FAST else 1


def reference(ctx, abi, torch, cam, rp, n_renders=64, spp=64, log=print):
    """float64 mean of n_renders independent spp-sample renders (spp - 1 passes with the reference's
    estimator); returns the display image on the device and the linear mean."""
    H, W = int(cam.resolution[0]), int(cam.resolution[1])
    acc = np.zeros((3, H, W), np.float64)
    t0 = time.time()
    p0 = first_pass(abi, rp)
    for i in range(n_renders):
        ctx.set_camera(0, with_spp(cam, spp), REF_SEED0 + i)
        ctx.reset(0)
        ctx.render(0, abi.render_desc(p0, spp, rp["depth"], REF_SEED0 + i, mode=rp.get("mode", abi.MODE_PARITY)))
        acc += ctx.read(0)
        if (i + 1) % 16 == 0:
            log(f"reference: {i + 1}/{n_renders} renders of {spp - p0} passes, {time.time() - t0:.1f} s")
    mean = (acc / n_renders).astype(np.float32)
    return _display(ctx, abi, torch, mean, H, W), mean


def rmse(a, b):
    return float(((a.double() - b.double()) ** 2).mean().sqrt())


def curves(ctx, abi, torch, cam, rp, ref_disp, unet=None, plain_ks=(1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024),
           denoise_ks=(1, 2, 4, 8, 16, 32)):
    H, W = int(cam.resolution[0]), int(cam.resolution[1])
    plain, den = [], []
    out = torch.empty((3, H, W), dtype=torch.float32, device="cuda") if unet is not None else None
    p0 = first_pass(abi, rp)
    for K in sorted(set(plain_ks) | set(denoise_ks)):
        ctx.set_camera(0, with_spp(cam, K + p0), TEST_SEED)
        ctx.reset(0)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.render(0, abi.render_desc(p0, K + p0, rp["depth"], TEST_SEED, mode=rp.get("mode", abi.MODE_PARITY)))
        ctx.sync()
        t_render = time.perf_counter() - t0
        if K in plain_ks:
            d = _display(ctx, abi, torch, ctx.read(0), H, W)
            plain.append((K, t_render, rmse(d, ref_disp)))
        if unet is not None and K in denoise_ks:
            if not den:                            # first call sizes the U-Net's buffers: not timed
                unet.denoise_view(ctx, 0, out)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            unet.denoise_view(ctx, 0, out)
            torch.cuda.synchronize()
            t_den = time.perf_counter() - t1
            d = _display(ctx, abi, torch, out, H, W)
            den.append((K, t_render, t_den, rmse(d, ref_disp)))
    return plain, den


def time_to_target(points, target=TARGET):
    """points: (K, seconds, rmse) sorted by K; log-log interpolation of seconds at rmse == target."""
    for (k0, t0, e0), (k1, t1, e1) in zip(points, points[1:]):
        if e0 > target >= e1:
            a = (np.log(e0) - np.log(target)) / (np.log(e0) - np.log(e1))
            return float(np.exp(np.log(max(t0, 1e-9)) + a * (np.log(max(t1, 1e-9)) - np.log(max(t0, 1e-9)))))
    if points and points[0][2] <= target:
        return float(points[0][1])
    return None
