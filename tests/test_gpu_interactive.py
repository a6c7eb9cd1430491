"""GPU parity of the interactive path through the C ABI: viewport preview, picking,
selection outline, overlay and the readback chain, against the CPU oracle.

Integer and byte results (hits, masks, uint8 images) must be identical; the
viewport layers are fp32 computed in the same operation order, so they must be
bit-exact too.
"""
import threading

import numpy as np
import pytest

from nectarrender_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _mesh_scene():
    """A 23k-triangle displaced icosphere (BVH deeper than the 8 LDS stack levels) on a quad, simple sky."""
    v, t = scenes.displaced_icosphere(freq=34)
    s = abi.SceneDesc()
    m = s.add_mesh(v, t)
    s.mesh(m, (0, 0.2, 0), (10, 20, 0), (1, 1, 1), abi.Material.pbr(albedo=(0.7, 0.8, 0.9), roughness=0.3))
    s.quad((0, -1.0, 0), (0, 0, 0), (6, 1, 6), abi.Material.lambertian((0.4, 0.5, 0.3)))
    s.sky_simple((0.8, 0.9, 1.0), (0.2, 0.3, 0.6))
    cam = abi.camera_desc((72, 96), (0, 0.6, 3.2), (-12, 4, 0), 4, 2.0, 3.2, 0.0, 2.0, 1.0)
    return s, cam


def _scene(name):
    if name == "config1":
        s, _, _ = scenes.config1(spp=4)
        return s, abi.camera_desc((64, 96), (0, 1, 4), (-10, 0, 0), 4, 2.0, 2.0, 0.0, 2.0, 1.0)
    if name == "mixed":
        s, cam, _ = scenes.mixed_scene(spp=4, res=80)
        return s, cam
    return _mesh_scene()


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("name", ["config1", "mixed", "mesh"])
@pytest.mark.parametrize("grid", [False, True])
def test_viewport_matches_oracle(ctx, oracle, name, grid):
    s, cam = _scene(name)
    ctx.upload(s)
    ctx.set_camera(0, cam, 3)
    ctx.viewport(0, grid)
    got = [ctx.read(0, L) for L in (abi.LAYER_BEAUTY, abi.LAYER_DIFFUSE, abi.LAYER_WORLD_NORMAL)]
    ref = oracle.viewport(oracle.OracleScene(s), oracle.OracleCamera(cam, 3), grid)
    for g, r, lname in zip(got, ref, ("beauty", "diffuse", "normal")):
        diff = _bits(g) != _bits(r)
        assert not diff.any(), f"{lname}: {np.count_nonzero(diff)} values differ, max |d| {np.abs(g - r).max()}"
    assert (got[1] != 0).any() and (got[1] == 0).all(axis=0).any()   # hits and misses present


@pytest.mark.parametrize("name", ["config1", "mixed", "mesh"])
def test_pick_matches_oracle(ctx, oracle, name):
    s, cam = _scene(name)
    ctx.upload(s)
    ctx.set_camera(0, cam, 0)
    sc, oc = oracle.OracleScene(s), oracle.OracleCamera(cam, 0)
    rng = np.random.default_rng(11)
    n_hit = 0
    for su, sv in rng.random((48, 2)).astype(np.float32):
        g, o = ctx.pick(0, float(su), float(sv)), oracle.pick(sc, oc, float(su), float(sv))
        assert bytes(g) == bytes(o), (su, sv, g.object_id, o.object_id)
        n_hit += g.object_id >= 0
    assert n_hit > 10


@pytest.mark.parametrize("name,objects", [("config1", (0, 1, 2, 3)), ("mixed", (0, 3, 5)), ("mesh", (0, 1))])
@pytest.mark.parametrize("radius", [0, 3])
def test_selection_mask_matches_oracle(ctx, oracle, name, objects, radius):
    import torch
    s, cam = _scene(name)
    ctx.upload(s)
    ctx.set_camera(0, cam, 0)
    H, W = int(cam.resolution[0]), int(cam.resolution[1])
    sc, oc = oracle.OracleScene(s), oracle.OracleCamera(cam, 0)
    d_mask = torch.full((H, W), 7, dtype=torch.uint8, device="cuda")
    for obj in objects:
        ctx.selection_mask_device(0, obj, radius, d_mask.data_ptr())
        ref = oracle.selection_mask(sc, oc, obj, radius)
        np.testing.assert_array_equal(d_mask.cpu().numpy(), ref)


def test_overlay_device_matches_oracle(ctx, oracle):
    import torch
    rng = np.random.default_rng(5)
    H, W = 37, 53
    img = rng.integers(0, 256, (3, H, W), dtype=np.uint8)
    mask = rng.integers(0, 256, (H, W), dtype=np.uint8)
    mask[rng.random((H, W)) < 0.4] = 0
    col = (0.98, 0.75, 0.17)
    d_img, d_mask = torch.from_numpy(img).cuda(), torch.from_numpy(mask).cuda()
    ctx.overlay_device(d_img.data_ptr(), H, W, d_mask.data_ptr(), col)
    np.testing.assert_array_equal(d_img.cpu().numpy(), oracle.overlay(img, mask, col))


def _render(ctx, s, cam, depth=4, spp=4):
    ctx.upload(s)
    ctx.set_camera(0, cam, 0)
    ctx.reset(0)
    ctx.render(0, abi.render_desc(1, spp, depth, 0))
    ctx.sync()


def test_readback_matches_post_and_overlay(ctx, oracle):
    import torch
    s, cam = _scene("config1")
    _render(ctx, s, cam)
    H, W = int(cam.resolution[0]), int(cam.resolution[1])
    post = abi.post_desc(tm_method=2, wb_temperature=5600.0)
    ref_u8, _ = oracle.post(ctx.read(0, abi.LAYER_BEAUTY), post)
    np.testing.assert_array_equal(ctx.readback(0, post), ref_u8)
    # relinked layer + the selection overlay of object 3
    post_n = abi.post_desc(layer=abi.LAYER_WORLD_NORMAL)
    d_mask = torch.zeros((H, W), dtype=torch.uint8, device="cuda")
    ctx.selection_mask_device(0, 3, 2, d_mask.data_ptr())
    got = ctx.readback(0, post_n, d_overlay=d_mask.data_ptr(), overlay_color=(0.98, 0.75, 0.17))
    ref_n, _ = oracle.post(ctx.read(0, abi.LAYER_WORLD_NORMAL), post_n)
    ref = oracle.overlay(ref_n, d_mask.cpu().numpy(), (0.98, 0.75, 0.17))
    assert (d_mask.cpu().numpy() > 0).any()
    np.testing.assert_array_equal(got, ref)


def test_readback_denoised_equals_denoise_then_post(ctx):
    import torch
    from nectarrender_amd import denoise
    net = denoise.UNet(0).load()
    s, cam = _scene("config1")
    _render(ctx, s, cam)
    H, W = int(cam.resolution[0]), int(cam.resolution[1])
    post = abi.post_desc(tm_method=2)
    got = ctx.readback(0, post, denoiser=net._h)
    den = torch.empty((3, H, W), device="cuda")
    net.denoise_view(ctx, 0, den)
    u8 = torch.empty((3, H, W), dtype=torch.uint8, device="cuda")
    ctx.post_device(den.data_ptr(), H, W, post, d_u8=u8.data_ptr())
    np.testing.assert_array_equal(got, u8.cpu().numpy())
    assert not np.array_equal(got, ctx.readback(0, post))       # the denoiser changed the image


def test_readback_concurrent_with_render(ctx, oracle):
    """Readbacks on another thread while passes render: each returns a whole image, and the last one
    equals the post-processed final layer."""
    s, cam = _scene("config1")
    ctx.upload(s)
    ctx.set_camera(0, abi.camera_desc((64, 96), (0, 1, 4), (-10, 0, 0), 24, 2.0, 2.0, 0.0, 2.0, 1.0), 0)
    ctx.reset(0)
    post = abi.post_desc()
    stop = threading.Event()
    seen, errors = [], []

    def reader():
        try:
            while not stop.is_set():
                seen.append(ctx.readback(0, post))
        except Exception as e:          # surfaced below
            errors.append(e)

    t = threading.Thread(target=reader)
    t.start()
    try:
        for k in range(1, 24):
            ctx.render(0, abi.render_desc(k, k + 1, 4, 0))
        ctx.sync()
    finally:
        stop.set()
        t.join(timeout=60)
    assert not t.is_alive() and not errors, errors
    assert len(seen) > 0 and all(a.shape == (3, 64, 96) for a in seen)
    ref, _ = oracle.post(ctx.read(0, abi.LAYER_BEAUTY), post)
    np.testing.assert_array_equal(ctx.readback(0, post), ref)
