"""The denoiser's CPU float64 checker (oracle/unet_ref.py), pinned without a GPU.

* finite differences: the gradients it returns equal central differences of
  its own loss (float64, a sample of parameters in every layer);
* closed form: all-zero weights make the network the identity on radiance;
* layout: the flat parameter order round-trips through unpack/pack_grads.
"""
import numpy as np
import pytest
import torch

from oracle import unet_ref as U


def _batch(rng, N=1, H=16, W=16):
    b = np.exp(rng.normal(size=(N, 3, H, W))).astype(np.float32)
    a = rng.random((N, 3, H, W)).astype(np.float32)
    n = rng.normal(size=(N, 3, H, W)).astype(np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    t = (b * 0.5 + 0.2).astype(np.float32)
    return b, a, n, t


def test_param_layout_roundtrip():
    rng = np.random.default_rng(0)
    p = rng.normal(size=U.param_count())
    assert np.array_equal(U.pack_grads(U.unpack(p)).numpy(), p)
    assert U.param_count() == 1016131


def test_zero_weights_identity():
    rng = np.random.default_rng(1)
    b, a, n, _ = _batch(rng)
    out, _ = U.denoise(np.zeros(U.param_count()), b, a, n)
    np.testing.assert_allclose(out, b, rtol=1e-12)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    b, a, n, t = _batch(rng)
    p = rng.normal(size=U.param_count()) * 0.05
    loss, g, _ = U.loss_and_grads(p, b, a, n, t)
    # one weight and the bias of a few layers, at positions where the L1 kink is not crossed
    checks, off = [], 0
    for li, (ci, co) in enumerate(U.layer_shapes()):
        if li in (0, 5, 9, 12, 17, 18):
            checks += [off + 7 % (co * 9 * ci), off + co * 9 * ci]
        off += co * 9 * ci + co
    h = 1e-6
    for k in checks:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        lp = U.loss_and_grads(pp, b, a, n, t)[0]
        lm = U.loss_and_grads(pm, b, a, n, t)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(g[k]), (k, fd, g[k])
