"""The denoiser's CPU float64 checker (oracle/unet_ref.py), pinned without a GPU.

* finite differences: the gradients it returns equal central differences of
  its own loss (float64, a sample of parameters in every layer);
* closed form: all-zero weights make the network the identity on radiance;
* the depth input: channel 9 of the first convolution is log1p(depth);
* layout: the flat parameter order round-trips through unpack/pack_grads.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import unet_ref as U


def _batch(rng, N=1, H=16, W=16):
    b = np.exp(rng.normal(size=(N, 3, H, W))).astype(np.float32)
    a = rng.random((N, 3, H, W)).astype(np.float32)
    n = rng.normal(size=(N, 3, H, W)).astype(np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    d = (rng.random((N, 1, H, W)) * 5).astype(np.float32)
    t = (b * 0.5 + 0.2).astype(np.float32)
    return b, a, n, d, t


def test_param_layout_roundtrip():
    rng = np.random.default_rng(0)
    p = rng.normal(size=U.param_count())
    assert np.array_equal(U.pack_grads(U.unpack(p)).numpy(), p)
    assert U.param_count() == 1016419


def test_zero_weights_identity():
    rng = np.random.default_rng(1)
    b, a, n, d, _ = _batch(rng)
    out, _ = U.denoise(np.zeros(U.param_count()), b, a, n, d)
    np.testing.assert_allclose(out, b, rtol=1e-12)


def test_depth_channel_is_log1p_depth():
    """The first convolution sees log1p(depth) as input channel 9: with only that channel's
    weights non-zero (0.1 on every tap) and depth constant at z, its pre-activation is
    bias + 0.9 log1p(z) inside the image and bias + 0.4 log1p(z) at a corner (4 of 9 taps
    inside).  A swapped channel, a missing log1p or a wrong padding fails it."""
    rng = np.random.default_rng(3)
    b, a, n, _, _ = _batch(rng, H=16, W=16)
    z = 3.5
    d = np.full((1, 1, 16, 16), z, np.float32)
    ci, co = U.layer_shapes()[0]
    p = np.zeros(U.param_count())
    w0 = np.zeros((co, 9, ci))
    w0[0, :, 9] = 0.1
    p[:co * 9 * ci] = w0.reshape(-1)
    p[co * 9 * ci] = 1.0                     # bias of output channel 0
    x, _ = U.inputs(b, a, n, d)
    np.testing.assert_allclose(x[:, 9].numpy(), np.log1p(z), rtol=1e-15)
    w, bias = U.unpack(p)[0]
    h = F.conv2d(x, w, bias, padding=1)[0, 0].numpy()
    np.testing.assert_allclose(h[1:-1, 1:-1], 1.0 + 0.9 * np.log1p(z), rtol=1e-12)
    np.testing.assert_allclose(h[0, 0], 1.0 + 0.4 * np.log1p(z), rtol=1e-12)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    b, a, n, d, t = _batch(rng)
    p = rng.normal(size=U.param_count()) * 0.05
    loss, g, _ = U.loss_and_grads(p, b, a, n, d, t)
    # one weight and the bias of a few layers, at positions where the L1 kink is not crossed
    checks, off = [], 0
    for li, (ci, co) in enumerate(U.layer_shapes()):
        if li in (0, 5, 9, 12, 17, 18):
            checks += [off + 7 % (co * 9 * ci), off + co * 9 * ci]
            if li == 0:
                checks.append(off + 9)       # a depth weight (output 0, tap 0, channel 9)
        off += co * 9 * ci + co
    h = 1e-6
    for k in checks:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        lp = U.loss_and_grads(pp, b, a, n, d, t)[0]
        lm = U.loss_and_grads(pm, b, a, n, d, t)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(g[k]), (k, fd, g[k])
