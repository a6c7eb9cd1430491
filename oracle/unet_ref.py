"""CPU float64 restatement of the denoiser network (test infrastructure only).

Mirrors include/nectar_denoise.h: inputs log1p(beauty), log1p(albedo),
clamped normal, log1p(depth) (10 channels); five levels of 32/48/64/96/128 channels with two
3x3 convolutions + ReLU per level and side, 2x2 average pooling, nearest 2x
upsampling, concatenation [encoder, upsampled], a 3x3 output convolution
added to log1p(beauty); loss = mean |y - log1p(target)| over pixels and RGB.

The OIDN network the reference calls (denoise.h:25-83) is not available, so
there is no reference output to match ("parity unpinned"); this file is the
checker for the GPU implementation's arithmetic.  With ``bf16=True`` weights
and every stored activation are rounded to bfloat16 where the GPU stores them,
so forward values differ only by accumulation order.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

CH = (32, 48, 64, 96, 128)
IN_CH = 10


def layer_shapes():
    """(cin, cout) of every convolution in parameter order."""
    out = []
    prev = IN_CH
    for c in CH:
        out += [(prev, c), (c, c)]
        prev = c
    for l in range(len(CH) - 2, -1, -1):
        out += [(CH[l] + CH[l + 1], CH[l]), (CH[l], CH[l])]
    out.append((CH[0], 3))
    return out


def param_count():
    return sum(co * 9 * ci + co for ci, co in layer_shapes())


def unpack(flat):
    """Flat parameters -> list of (weight [cout, cin, 3, 3], bias [cout]) tensors (float64)."""
    t = torch.as_tensor(np.asarray(flat), dtype=torch.float64)
    out, o = [], 0
    for ci, co in layer_shapes():
        w = t[o:o + co * 9 * ci].reshape(co, 9, ci).permute(0, 2, 1).reshape(co, ci, 3, 3)
        o += co * 9 * ci
        b = t[o:o + co]
        o += co
        out.append((w, b))
    return out


def pack_grads(layers):
    """Inverse of unpack for gradient tensors."""
    parts = []
    for w, b in layers:
        co, ci = w.shape[:2]
        parts.append(w.reshape(co, ci, 9).permute(0, 2, 1).reshape(-1))
        parts.append(b.reshape(-1))
    return torch.cat(parts)


def _rb(x, on):
    return x.to(torch.bfloat16).to(torch.float64) if on else x


def inputs(beauty, albedo, normal, depth, bf16=False):
    """(N,3,H,W) beauty, albedo, normal and (N,1,H,W) depth -> network input (N,10,H,W) and
    log1p(beauty)."""
    b = torch.as_tensor(beauty, dtype=torch.float64)
    a = torch.as_tensor(albedo, dtype=torch.float64)
    n = torch.as_tensor(normal, dtype=torch.float64)
    z = torch.as_tensor(depth, dtype=torch.float64)
    xl = torch.log1p(torch.nan_to_num(b, nan=0.0).clamp(min=0))
    x = torch.cat([xl, torch.log1p(torch.nan_to_num(a, nan=0.0).clamp(min=0)),
                   torch.nan_to_num(n, nan=0.0).clamp(-1, 1),
                   torch.log1p(torch.nan_to_num(z, nan=0.0).clamp(min=0))], 1)
    # the GPU keeps log1p(beauty) in fp32 for the residual and rounds the network input to bf16
    return _rb(x, bf16), xl.to(torch.float32).to(torch.float64) if bf16 else xl


def forward(layers, x, xl, bf16=False):
    """Log-space output y (N,3,H,W)."""
    L = [(_rb(w, bf16), b) for w, b in layers]
    conv = lambda t, i, relu=True: (lambda r: F.relu(r) if relu else r)(F.conv2d(t, L[i][0], L[i][1], padding=1))
    enc, h = [], x
    for l in range(len(CH)):
        h = _rb(conv(h, 2 * l), bf16)
        h = _rb(conv(h, 2 * l + 1), bf16)
        enc.append(h)
        if l < len(CH) - 1:
            h = _rb(F.avg_pool2d(h, 2), bf16)
    d = enc[-1]
    k = 10
    for l in range(len(CH) - 2, -1, -1):
        u = F.interpolate(d, scale_factor=2, mode="nearest")
        h = torch.cat([enc[l], u], 1)
        h = _rb(conv(h, k), bf16)
        d = _rb(conv(h, k + 1), bf16)
        k += 2
    return conv(d, k, relu=False) + xl


def loss_and_grads(flat, beauty, albedo, normal, depth, target, bf16=False):
    layers = [(w.clone().requires_grad_(True), b.clone().requires_grad_(True)) for w, b in unpack(flat)]
    x, xl = inputs(beauty, albedo, normal, depth, bf16)
    y = forward(layers, x, xl, bf16)
    t = torch.log1p(torch.as_tensor(target, dtype=torch.float64).clamp(min=0))
    loss = (y - t).abs().mean()
    loss.backward()
    g = pack_grads([(w.grad, b.grad) for w, b in layers])
    return float(loss.detach()), g.detach().numpy(), y.detach().numpy()


def denoise(flat, beauty, albedo, normal, depth, bf16=False):
    with torch.no_grad():
        x, xl = inputs(beauty, albedo, normal, depth, bf16)
        y = forward(unpack(flat), x, xl, bf16)
    return torch.expm1(y.clamp(min=0)).numpy(), y.numpy()

