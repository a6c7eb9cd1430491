"""Fast mode (NR_MODE_FAST): next-event estimation to the HDRI sky with power-heuristic
MIS against the scattering technique, and a 64-bit per-sample subsequence.

The reference has no such estimator, so there is nothing to match bit for bit; what
fixes it instead:
- a closed form: a convex Lambertian object of albedo a under a constant sky of
  radiance 1 has radiance a (white furnace);
- unbiasedness: on scenes with every material type, an object light and a medium, the
  fast estimator's expectation equals the reference estimator's (z-test over seeds);
- the point of the mode: at equal samples its error against a converged image is lower.
"""
import numpy as np
import pytest

from nectarrender_amd import abi, scenes

pytestmark = pytest.mark.gpu


def blob_hdri(W=256, H=128, radiance=40.0, width_deg=12.0, elev_deg=50.0, azim_deg=-60.0, background=0.4):
    """CHW equirect: smooth background gradient + a bright disc (same texel -> direction map as the lookup)."""
    u = (np.arange(W) + 0.5) / W
    v = (np.arange(H) + 0.5) / H
    phi = (u - 0.5) * 2 * np.pi
    th = v * np.pi
    d = np.stack([np.sin(th)[:, None] * np.cos(phi)[None, :], np.cos(th)[:, None] * np.ones((1, W)),
                  np.sin(th)[:, None] * np.sin(phi)[None, :]])
    el, az = np.radians(elev_deg), np.radians(azim_deg)
    sd = np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)])
    cosang = np.tensordot(sd, d, axes=(0, 0))
    img = np.empty((3, H, W), np.float32)
    for c, tint in enumerate((1.0, 0.9, 0.8)):
        img[c] = background * tint * (0.6 + 0.4 * np.cos(th))[:, None]
    img[:, cosang >= np.cos(np.radians(width_deg / 2))] = radiance
    return img


def _render(ctx, s, cam, spp, mode, seed=3, depth=6, streams=None, upload=True):
    if upload:
        ctx.upload(s)
    ctx.set_camera(0, cam, seed)
    if streams is not None:
        ctx.set_render_streams(streams)
    ctx.reset(0)
    # fast mode renders passes 0..spp-1, the reference's estimator passes 1..spp-1
    ctx.render(0, abi.render_desc(0 if mode == abi.MODE_FAST else 1, spp, depth=depth, seed=seed, mode=mode))
    img = ctx.read(0, abi.LAYER_BEAUTY)
    if streams is not None:
        ctx.set_render_streams(2)
    return img


def _furnace_scene(albedo=0.5, spp=256):
    s = abi.SceneDesc()
    s.sphere((0, 0, 0), 1.0, abi.Material.lambertian((albedo, albedo, albedo)))
    s.sky_hdri(np.ones((3, 32, 64), np.float32))
    cam = abi.camera_desc((48, 48), (0, 0, 4), (0, 0, 0), spp, 2.0, 4.0, 0.0, 2.0, 1.0)
    return s, cam


def test_fast_white_furnace(ctx):
    s, cam = _furnace_scene(0.5, spp=256)
    par = _render(ctx, s, cam, 256, abi.MODE_PARITY)
    inside = np.all(np.abs(par - 0.5) < 1e-5, axis=0)        # pixels whose every sample hits the sphere
    assert inside.sum() > 300
    fast = _render(ctx, s, cam, 256, abi.MODE_FAST)
    v = fast[0, inside]                                       # grey albedo: the channels carry the same samples
    m, se = float(v.mean()), float(v.std(ddof=1) / np.sqrt(v.size))
    print(f"furnace: {v.size} pixels, fast-mode mean {m:.5f} +- {se:.5f} (exact 0.5)")
    assert abs(m - 0.5) < max(5 * se, 1e-4)
    assert abs(m - 0.5) < 3e-3
    # sky pixels see the constant radiance 1 directly
    sky = np.all(np.abs(par - 1.0) < 1e-6, axis=0)
    np.testing.assert_allclose(fast[:, sky], 1.0, rtol=1e-6)


def _mixed_hdri_scene(res=40):
    s = abi.SceneDesc()
    s.quad((0, 0, 0), (0, 0, 0), (10, 1, 10), abi.Material.lambertian((0.5, 0.5, 0.5)))
    s.sphere((-1.1, 0.5, 0), 0.5, abi.Material.pbr(albedo=(1.0, 0.6, 0.15), roughness=0.35, metallic=1.0))
    s.sphere((0, 0.5, 0), 0.5, abi.Material.dielectric((1, 1, 1), 1.5))
    s.sphere((1.1, 0.5, 0), 0.5, abi.Material.lambertian((0.2, 0.4, 0.8)))
    s.sphere((0.0, 0.35, 1.0), 0.35, abi.Material.pbr(albedo=(0.8, 0.8, 0.8), roughness=0.8, metallic=0.0))
    s.constant_medium(abi.SceneDesc.bound(abi.HIT_SPHERE, (-0.75, 0.32, 1.2), radius=0.3), 1.5, (0.9, 0.9, 0.9))
    s.object_light(abi.SceneDesc.bound(abi.HIT_QUAD, (1.5, 2.5, 1.0), (0, 0, 180), (0.6, 1, 0.6)), 6.0)
    s.sky_hdri(blob_hdri())
    cam = abi.camera_desc((res, res), (0, 1, 4), (-10, 0, 0), 4096, 2.0, 2.0, 0.0, 2.0, 1.0)
    return s, cam


def test_fast_is_unbiased_against_the_reference_estimator(ctx):
    """Block means of fast and reference renders over independent seeds agree to within 5 sigma."""
    s, cam = _mixed_hdri_scene()
    ctx.upload(s)
    seeds, spp, B = range(10, 18), 256, 8
    blocks = {}
    for mode in (abi.MODE_PARITY, abi.MODE_FAST):
        per_seed = []
        for sd in seeds:
            img = _render(ctx, s, cam, spp, mode, seed=sd, upload=False)
            C, H, W = img.shape
            per_seed.append(img.reshape(C, H // B, B, W // B, B).mean(axis=(2, 4)))
        blocks[mode] = np.stack(per_seed)                    # seeds x C x bh x bw
    mp, mf = blocks[abi.MODE_PARITY].mean(0), blocks[abi.MODE_FAST].mean(0)
    n = len(seeds)
    se = np.sqrt(blocks[abi.MODE_PARITY].var(0, ddof=1) / n + blocks[abi.MODE_FAST].var(0, ddof=1) / n)
    z = np.abs(mp - mf) / np.maximum(se, 1e-6)
    print(f"max |z| over {z.size} block means: {z.max():.2f}; image mean parity {mp.mean():.5f} fast {mf.mean():.5f}")
    assert z.max() < 5.0
    assert abs(mp.mean() - mf.mean()) < 0.01 * mp.mean()
    # fast mode's spread over seeds is smaller
    assert blocks[abi.MODE_FAST].std(0).mean() < blocks[abi.MODE_PARITY].std(0).mean()


def test_fast_lowers_error_at_equal_samples(ctx):
    s, cam = _mixed_hdri_scene(res=48)
    s.sky_hdri(blob_hdri(radiance=2000.0, width_deg=2.0))     # a small, bright sun
    ctx.upload(s)
    ref = _render(ctx, s, cam, 4096, abi.MODE_FAST, seed=99, upload=False)
    err = {}
    for mode in (abi.MODE_PARITY, abi.MODE_FAST):
        img = _render(ctx, s, cam, 16, mode, seed=5, upload=False)
        err[mode] = float(np.sqrt(np.mean((img - ref) ** 2)))
    print(f"RMSE at 16 spp vs 4096-spp image: reference estimator {err[abi.MODE_PARITY]:.4f}, "
          f"fast {err[abi.MODE_FAST]:.4f}")
    assert err[abi.MODE_FAST] < 0.5 * err[abi.MODE_PARITY]


def test_fast_is_deterministic_and_independent_of_stream_groups(ctx):
    s, cam = _mixed_hdri_scene(res=64)
    a = _render(ctx, s, cam, 8, abi.MODE_FAST, streams=1)
    b = _render(ctx, s, cam, 8, abi.MODE_FAST, streams=3, upload=False)
    c = _render(ctx, s, cam, 8, abi.MODE_FAST, upload=False)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a, c)


@pytest.mark.parametrize("tail", [0, 1 << 30])
def test_fast_tail_mode_gives_the_same_image(ctx, tail):
    """The tail kernel runs the fast estimator's per-path steps (next-event estimation and shadow rays
    included) in the same order: the image is the wavefront's bit for bit."""
    s, cam = _mixed_hdri_scene(res=56)
    ctx.set_tuning(4, 2, 8, 262144)
    ref = _render(ctx, s, cam, 6, abi.MODE_FAST)
    ctx.set_tuning(3, 2, 8, tail)
    try:
        got = _render(ctx, s, cam, 6, abi.MODE_FAST, upload=False)
    finally:
        ctx.set_tuning(4, 2, 8, 262144)
    np.testing.assert_array_equal(got, ref)


def test_fast_with_a_simple_sky_runs_without_sky_sampling(ctx):
    """No HDRI tables: fast mode is the reference estimator with the 64-bit subsequence."""
    s, cam, _ = scenes.config1(spp=16)
    cam = abi.camera_desc((64, 64), (0, 1, 4), (-10, 0, 0), 64, 2.0, 2.0, 0.0, 2.0, 1.0)
    par = _render(ctx, s, cam, 64, abi.MODE_PARITY)
    fast = _render(ctx, s, cam, 64, abi.MODE_FAST, upload=False)
    assert np.isfinite(fast).all()
    assert abs(par.mean() - fast.mean()) < 0.02 * par.mean()
    assert not np.array_equal(par, fast)


def test_fast_config2_scene_small(ctx):
    """The 1M-triangle HDRI scene at reduced resolution renders in fast mode with finite, positive radiance."""
    s, _, _ = scenes.config2(spp=4, freq=40, hdri_w=512, hdri_h=256)
    cam = abi.camera_desc((108, 192), (0.0, 0.5, 3.8), (-30.0, 20.15, 0.0), 4, 2.0, 3.8, 0.0, 2.0, 1.0)
    img = _render(ctx, s, cam, 4, abi.MODE_FAST, depth=12)
    assert np.isfinite(img).all() and (img >= 0).all() and img.mean() > 0


def test_fast_pass_numbering(ctx):
    """Fast mode folds pass s as sample s + 1 (passes 0..n-1 make an n-sample image); the reference's
    estimator starts at pass 1 and refuses 0."""
    s, cam = _furnace_scene(0.5, spp=8)
    ctx.upload(s)
    ctx.set_camera(0, cam, 3)
    with pytest.raises(RuntimeError):
        ctx.render(0, abi.render_desc(0, 4, depth=4, seed=3, mode=abi.MODE_PARITY))
    ctx.reset(0)
    ctx.render(0, abi.render_desc(0, 8, depth=4, seed=3, mode=abi.MODE_FAST))
    whole = ctx.read(0)
    ctx.reset(0)
    for p in range(8):                                          # one call per pass: same running mean
        ctx.render(0, abi.render_desc(p, p + 1, depth=4, seed=3, mode=abi.MODE_FAST))
    np.testing.assert_array_equal(ctx.read(0), whole)
    # pass 0 alone is the first sample itself (weight 1)
    ctx.reset(0)
    ctx.render(0, abi.render_desc(0, 1, depth=4, seed=3, mode=abi.MODE_FAST))
    first = ctx.read(0)
    sky = np.all(np.abs(whole - 1.0) < 1e-6, axis=0)
    np.testing.assert_array_equal(first[:, sky], 1.0)
