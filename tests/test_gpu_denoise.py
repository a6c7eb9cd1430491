"""The denoiser's HIP kernels (BF16 MFMA, FP32 accumulation) against the float64 checker.

Tolerances (bf16 storage of weights and activations): forward log-radiance
within 2e-2 absolute of the checker run with the same bf16 rounding points;
parameter gradients with relative L2 error below 5e-2 and cosine above 0.995
per layer; Adam updates to 1e-6 relative.  Shapes: small crops, odd sizes, and
one step at the training configuration's batch (16 crops of 256 x 256).
"""
import numpy as np
import pytest

from oracle import unet_ref as U

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.fixture(scope="module")
def net():
    from nectarrender_amd import denoise
    n = denoise.UNet(0, seed=3)
    n.p0 = n.get_params()                      # every test perturbs these, whatever ran before it
    yield n
    n.close()


def _batch(rng, N, H, W):
    b = np.exp(rng.normal(size=(N, 3, H, W)) * 0.8).astype(np.float32)
    a = rng.random((N, 3, H, W)).astype(np.float32)
    n = rng.normal(size=(N, 3, H, W)).astype(np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    d = (rng.random((N, 1, H, W)) * 8).astype(np.float32)
    d[:, :, : H // 4] = 0.0                      # camera rays that miss
    t = (0.5 * b + 0.25 * np.roll(b, 1, axis=3) + 0.25 * np.roll(b, 1, axis=2)).astype(np.float32)
    return b, a, n, d, t


def _cuda(torch, *xs):
    return [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in xs]


def _params(net, rng):
    # larger than the initial output layer so every path of the network matters
    p = (net.p0 + rng.normal(size=net.p0.shape).astype(np.float32) * 0.01).astype(np.float32)
    net.set_params(p)
    return p


@pytest.mark.parametrize("shape", [(2, 32, 48), (1, 37, 53)])
def test_forward_matches_checker(torch_cuda, net, shape):
    rng = np.random.default_rng(10)
    N, H, W = shape
    p = _params(net, rng)
    b, a, n, d, _ = _batch(rng, N, H, W)
    out = net.forward(*_cuda(torch_cuda, b, a, n, d)).cpu().numpy()
    Hp, Wp = -(-H // 16) * 16, -(-W // 16) * 16
    pad = lambda x: np.pad(x, ((0, 0), (0, 0), (0, Hp - H), (0, Wp - W)))
    ref, y = U.denoise(p, pad(b), pad(a), pad(n), pad(d), bf16=True)
    got_y = np.log1p(out)
    err = np.abs(got_y - np.maximum(y[:, :, :H, :W], 0))
    assert err.max() < 2e-2, err.max()
    assert np.isfinite(out).all()


def test_forward_at_1080p_matches_checker(torch_cuda, net):
    """The benchmark frame's shape (1080 x 1920, padded to 1088 x 1920): the full-resolution
    layers run 8-row workgroups and the first and last convolutions read the render layers and
    write the planar image themselves, a configuration only frames of about a megapixel reach."""
    rng = np.random.default_rng(14)
    p = _params(net, rng)
    H, W = 1080, 1920
    b, a, n, d, _ = _batch(rng, 1, H, W)
    out = net.forward(*_cuda(torch_cuda, b, a, n, d)).cpu().numpy()
    pad = lambda x: np.pad(x, ((0, 0), (0, 0), (0, 8), (0, 0)))
    _, y = U.denoise(p, pad(b), pad(a), pad(n), pad(d), bf16=True)
    err = np.abs(np.log1p(out) - np.maximum(y[:, :, :H, :W], 0))
    assert np.isfinite(out).all()
    assert err.max() < 2e-2, err.max()


def _check_gradients(g, ref_g):
    off = 0
    for li, (ci, co) in enumerate(U.layer_shapes()):
        sz = co * 9 * ci + co
        gg, rr = g[off:off + sz].astype(np.float64), ref_g[off:off + sz]
        off += sz
        rel = np.linalg.norm(gg - rr) / max(np.linalg.norm(rr), 1e-30)
        cos = gg @ rr / max(np.linalg.norm(gg) * np.linalg.norm(rr), 1e-30)
        assert rel < 5e-2 and cos > 0.995, (li, rel, cos)


def test_gradients_match_checker(torch_cuda, net):
    rng = np.random.default_rng(11)
    p = _params(net, rng)
    b, a, n, d, t = _batch(rng, 2, 32, 32)
    loss, g = net.gradients(*_cuda(torch_cuda, b, a, n, d, t))
    ref_loss, ref_g, _ = U.loss_and_grads(p, b, a, n, d, t, bf16=True)
    assert abs(loss - ref_loss) < 1e-3 * max(1.0, ref_loss), (loss, ref_loss)
    _check_gradients(g, ref_g)


def test_train_step_at_training_batch(torch_cuda, net):
    """One step at the training configuration's shape, 16 crops of 256 x 256: loss and every
    layer's gradient against the checker, and the Adam update the step applies."""
    rng = np.random.default_rng(13)
    p = _params(net, rng)
    batch = _batch(rng, 16, 256, 256)
    dev = _cuda(torch_cuda, *batch)
    loss, g = net.gradients(*dev)
    ref_loss, ref_g, _ = U.loss_and_grads(p, *batch, bf16=True)
    assert abs(loss - ref_loss) < 1e-3 * max(1.0, ref_loss), (loss, ref_loss)
    _check_gradients(g, ref_g)
    net.set_params(p)                            # resets Adam state
    step_loss = net.train_step(*dev, lr=1e-3)
    assert step_loss == pytest.approx(loss, rel=1e-6)
    m = 0.1 * g.astype(np.float64)
    v = 0.001 * g.astype(np.float64) ** 2
    expect = p - 1e-3 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    np.testing.assert_allclose(net.get_params(), expect, rtol=1e-5, atol=1e-7)


def test_adam_step_and_training_reduces_loss(torch_cuda, net):
    rng = np.random.default_rng(12)
    net.set_params(net.get_params())            # resets Adam state
    p0 = net.get_params()
    b, a, n, d, t = _cuda(torch_cuda, *_batch(rng, 2, 32, 32))
    _, g = net.gradients(b, a, n, d, t)
    net.train_step(b, a, n, d, t, lr=1e-3)
    p1 = net.get_params()
    m = 0.1 * g.astype(np.float64)
    v = 0.001 * g.astype(np.float64) ** 2
    expect = p0 - 1e-3 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    np.testing.assert_allclose(p1, expect, rtol=1e-5, atol=1e-7)
    losses = [net.train_step(b, a, n, d, t, lr=1e-3) for _ in range(60)]
    assert losses[-1] < 0.7 * losses[0], losses[::10]


def test_denoise_view(torch_cuda, net):
    from nectarrender_amd import abi, native, scenes
    s, cam, rp = scenes.cornell(spp=4, res=64)
    ctx = native.Context(0)
    ctx.upload(s)
    ctx.set_camera(0, cam, rp["seed"])
    ctx.render(0, abi.render_desc(1, 4, rp["depth"], rp["seed"]))
    out = torch_cuda.empty((3, 64, 64), dtype=torch_cuda.float32, device="cuda")
    net.denoise_view(ctx, 0, out)
    beauty = ctx.read(0)
    ref, _ = U.denoise(net.get_params(), beauty[None], ctx.read(0, abi.LAYER_DIFFUSE)[None],
                       ctx.read(0, abi.LAYER_WORLD_NORMAL)[None], ctx.read(0, abi.LAYER_DEPTH)[None], bf16=True)
    got = out.cpu().numpy()
    assert np.abs(np.log1p(got) - np.log1p(ref[0])).max() < 2e-2
    ms, flops = net.last_stats()
    assert ms > 0 and flops > 0
    ctx.close()
