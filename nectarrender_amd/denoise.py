"""The learned denoiser of libnectar_hip.so (include/nectar_denoise.h) from Python.

Device buffers are torch tensors on the GPU (plain memory and streams); every
computation runs in the library's HIP kernels.  There is no CPU path: without
the library or a GPU every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import native

_HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(_HERE, "weights", "unet.npy")


def _lib():
    L = native.lib()
    if not getattr(L, "_unet_bound", False):
        vp, fp = C.c_void_p, C.c_void_p
        sig = {
            "nr_unet_param_count": (C.c_size_t, []),
            "nr_unet_create": (vp, [C.c_int, C.c_uint32]),
            "nr_unet_destroy": (None, [vp]),
            "nr_unet_set_params": (C.c_int, [vp, C.POINTER(C.c_float), C.c_size_t]),
            "nr_unet_get_params": (C.c_int, [vp, C.POINTER(C.c_float), C.c_size_t]),
            "nr_unet_forward": (C.c_int, [vp, fp, fp, fp, fp, C.c_int32, C.c_int32, C.c_int32, fp]),
            "nr_unet_train_step": (C.c_int, [vp, fp, fp, fp, fp, fp, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                             C.POINTER(C.c_float)]),
            "nr_unet_gradients": (C.c_int, [vp, fp, fp, fp, fp, fp, C.c_int32, C.c_int32, C.c_int32,
                                            C.POINTER(C.c_float), C.POINTER(C.c_float)]),
            "nr_denoise": (C.c_int, [vp, C.c_int, vp, fp]),
            "nr_unet_last_stats": (C.c_int, [vp, C.POINTER(C.c_float), C.POINTER(C.c_double)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L._unet_bound = True
    return L


def _check(rc):
    if rc != 0:
        raise native.NectarError(f"nectar_hip error {rc}: {native.lib().nr_last_error().decode()}")


def _dev(t, shape=None):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise ValueError("expected a contiguous float32 CUDA tensor")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"expected shape {shape}, got {tuple(t.shape)}")
    return C.c_void_p(t.data_ptr())


class UNet:
    """U-Net denoiser on HIP device `device` (He-normal init from `seed`)."""

    def __init__(self, device: int = 0, seed: int = 0):
        import torch                      # device buffers are torch tensors; its runtime starts first
        torch.cuda.init()
        L = _lib()
        self._h = L.nr_unet_create(device, seed)
        if not self._h:
            raise native.NectarError(f"nr_unet_create failed: {L.nr_last_error().decode()}")
        self.n_params = int(L.nr_unet_param_count())

    def close(self):
        if self._h:
            _lib().nr_unet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters ------------------------------------------------------------
    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        _check(_lib().nr_unet_get_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        _check(_lib().nr_unet_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size))

    def load(self, path: str = WEIGHTS):
        self.set_params(np.load(path, allow_pickle=False))
        return self

    def save(self, path: str = WEIGHTS):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.save(path, self.get_params())

    # compute ---------------------------------------------------------------
    def forward(self, beauty, albedo, normal, depth, out=None):
        """(N,3,H,W) beauty, albedo, normal and (N,1,H,W) depth, float32 CUDA tensors ->
        denoised radiance (N,3,H,W)."""
        import torch
        N, _, H, W = beauty.shape
        if out is None:
            out = torch.empty_like(beauty)
        _check(_lib().nr_unet_forward(self._h, *self._inputs(beauty, albedo, normal, depth), N, H, W,
                                      _dev(out, beauty.shape)))
        return out

    def train_step(self, beauty, albedo, normal, depth, target, lr=1e-3) -> float:
        N, _, H, W = beauty.shape
        loss = C.c_float()
        _check(_lib().nr_unet_train_step(self._h, *self._inputs(beauty, albedo, normal, depth),
                                         _dev(target, beauty.shape), N, H, W, float(lr), C.byref(loss)))
        return loss.value

    def gradients(self, beauty, albedo, normal, depth, target):
        N, _, H, W = beauty.shape
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        _check(_lib().nr_unet_gradients(self._h, *self._inputs(beauty, albedo, normal, depth),
                                        _dev(target, beauty.shape), N, H, W, g.ctypes.data_as(C.POINTER(C.c_float)),
                                        C.byref(loss)))
        return loss.value, g

    @staticmethod
    def _inputs(beauty, albedo, normal, depth):
        N, _, H, W = beauty.shape
        return (_dev(beauty, (N, 3, H, W)), _dev(albedo, (N, 3, H, W)), _dev(normal, (N, 3, H, W)),
                _dev(depth, (N, 1, H, W)))

    def denoise_view(self, ctx: native.Context, view: int, out):
        """Denoise a render context view's beauty (uses its DIFFUSE, WORLD_NORMAL and DEPTH layers) into
        `out` (3,H,W)."""
        H, W = ctx.views[view]
        _check(_lib().nr_denoise(ctx._h, view, self._h, _dev(out, (3, H, W))))
        return out

    def last_stats(self):
        ms, fl = C.c_float(), C.c_double()
        _check(_lib().nr_unet_last_stats(self._h, C.byref(ms), C.byref(fl)))
        return ms.value, fl.value


def _time_forward(H: int, W: int, reps: int) -> dict:
    """Median wall time (HIP events inside the library) of one U-Net forward over an H x W
    frame of synthetic inputs (beauty 0..2, albedo 0..1, normals -1..1, depth 0..10), shipped
    weights."""
    import torch
    net = UNet(0).load()
    g = torch.Generator(device="cuda").manual_seed(0)
    b = torch.rand((1, 3, H, W), device="cuda", generator=g) * 2
    a = torch.rand((1, 3, H, W), device="cuda", generator=g)
    n = torch.rand((1, 3, H, W), device="cuda", generator=g) * 2 - 1
    d = torch.rand((1, 1, H, W), device="cuda", generator=g) * 10
    out = torch.empty_like(b)
    for _ in range(3):
        net.forward(b, a, n, d, out)
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        net.forward(b, a, n, d, out)
        ms.append(net.last_stats()[0])
    flops = net.last_stats()[1]
    med = sorted(ms)[len(ms) // 2]
    return {"H": H, "W": W, "forward_ms": round(med, 4), "flops": flops,
            "tflops": round(flops / med / 1e9, 2)}


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="time the U-Net forward pass on HIP device 0")
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args()
    print(json.dumps(_time_forward(args.height, args.width, args.reps)), flush=True)
