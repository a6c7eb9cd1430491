"""Live scene edits as partial updates: SceneInterface.add_object (csrc/scene/include/interface.h:112-116,
Scene::add_hittable, csrc/scene/src/scene.cpp:101-104) and swap_skylight (interface.h:145-151).  After
each edit the hits and the renders equal those of a fresh upload of the edited scene and the oracle's, and
adding an object to the 1,003,520-triangle scene rebuilds no mesh tree."""
import ctypes as C
import time

import numpy as np
import pytest

from nectarrender_amd import abi, native, scenes
from nectarrender_amd.abi import Material, SceneDesc, Texture

pytestmark = pytest.mark.gpu


def _hits_equal(got, ref):
    np.testing.assert_array_equal(got["object_id"], ref["object_id"])
    np.testing.assert_array_equal(got["prim"], ref["prim"])
    hit = got["object_id"] >= 0
    for f in ("t", "p", "n"):
        np.testing.assert_array_equal(got[f][hit], ref[f][hit], err_msg=f)
    np.testing.assert_array_equal(got["material_id"][hit], ref["material_id"][hit])


def _render(ctx, cam, rp, mode, passes=2):
    p0 = 0 if mode == abi.MODE_FAST else 1
    ctx.set_camera(0, cam, rp["seed"])
    ctx.reset(0)
    ctx.render(0, abi.render_desc(p0, p0 + passes, rp["depth"], rp["seed"], mode=mode))
    return [ctx.read(0, layer) for layer in (abi.LAYER_BEAUTY, abi.LAYER_DIFFUSE, abi.LAYER_WORLD_NORMAL)]


def _same(a, b):
    for x, y in zip(a, b):
        assert ((x == y) | (np.isnan(x) & np.isnan(y))).all()


def _check_against_fresh(ctx, oracle, s, cam, rp, rays):
    """Hits and renders (both estimators) of the edited context, then of a fresh upload, then the oracle's."""
    got = ctx.trace_closest(rays)
    edited = {m: _render(ctx, cam, rp, m) for m in (abi.MODE_PARITY, abi.MODE_FAST)}
    ctx.upload(s)
    _hits_equal(got, ctx.trace_closest(rays))
    for m in (abi.MODE_PARITY, abi.MODE_FAST):
        _same(edited[m], _render(ctx, cam, rp, m))
    _hits_equal(got, oracle.OracleScene(s).trace_closest(rays))
    ref = oracle.render(oracle.OracleScene(s), oracle.OracleCamera(cam, rp["seed"]),
                        abi.render_desc(1, 3, rp["depth"], rp["seed"]))
    _same(edited[abi.MODE_PARITY], ref)


def test_add_objects_equal_a_fresh_upload(ctx, oracle):
    s, cam, rp = scenes.mixed_scene(res=48)
    rays = scenes.random_rays(40_000, 21, center=(0, 0.3, 0), spread=2.5)
    n0 = len(s.objects)
    ctx.upload(s)
    # a sphere with a material of its own
    s.sphere((0.0, 0.75, 0.0), 0.3, Material.pbr(albedo=(0.2, 0.5, 0.9), roughness=0.3, metallic=0.5))
    ctx.add_objects(s)
    _check_against_fresh(ctx, oracle, s, cam, rp, rays)
    # the uploaded mesh again, with a texture on a new image; then a new mesh large enough for
    # the device tree builder, a medium in a cube and an object light, registered after the lights
    rng = np.random.default_rng(5)
    img = s.add_image((rng.random((3, 8, 8)) * 255).astype(np.uint8))
    s.mesh(0, (0.9, 0.5, -0.2), (0, 45, 10), (0.8, 0.8, 0.8), Material.lambertian(Texture(image=img, scale=2.0)))
    P, T = scenes.geodesic_sphere(24)
    mi = s.add_mesh(scenes.mesh_vertices((P * 0.25).astype(np.float32), T), T)
    assert T.shape[0] >= 4096
    s.mesh(mi, (-1.0, 0.7, -0.3), (0, 0, 0), (1.0, 1.3, 1.0), Material.dielectric((1.0, 0.9, 0.9), 1.3))
    s.constant_medium(SceneDesc.bound(abi.HIT_CUBE, (0.3, -0.3, 0.5), (0, 10, 0), (0.3, 0.3, 0.3)), 3.0,
                      (0.7, 0.9, 0.7))
    s.object_light(SceneDesc.bound(abi.HIT_SPHERE, (0.8, 1.4, -0.6), radius=0.15), 12.0, (1.0, 0.8, 0.6))
    ctx.add_objects(s)
    got = ctx.trace_closest(rays)
    assert set(np.unique(got["object_id"])) >= {n0 + 1, n0 + 2}, "the added objects are hit"
    _check_against_fresh(ctx, oracle, s, cam, rp, rays)


def test_set_sky_equals_a_fresh_upload(ctx, oracle):
    s, cam, rp = scenes.mixed_scene(res=48)
    rays = scenes.random_rays(10_000, 22, center=(0, 0.3, 0), spread=2.5)
    ctx.upload(s)
    s.sky_hdri(scenes.procedural_hdri(96, 48, seed=3, sun_diameter_deg=6.0, sun_radiance=9.0), rotation=-70.0,
               intensity=1.3)
    ctx.set_sky(s)
    _check_against_fresh(ctx, oracle, s, cam, rp, rays)
    s.sky_simple((1.0, 0.9, 0.8), (0.3, 0.5, 1.0))
    ctx.set_sky(s)
    _check_against_fresh(ctx, oracle, s, cam, rp, rays)


def test_add_object_to_the_1m_triangle_scene_rebuilds_only_the_top_level(ctx, oracle):
    s, cam, rp = scenes.config2()
    ctx.upload(s)
    rays = scenes.random_rays(20_000, 23, center=(0, 0, 0), spread=3.0)
    times = []
    for k, pos in enumerate([(1.3, -0.6, 0.4), (-1.4, -0.7, -0.2), (0.2, 1.2, 0.9)]):
        s.sphere(pos, 0.35, Material.lambertian((0.2 + 0.2 * k, 0.6, 0.3)))
        ctx.sync()
        t0 = time.perf_counter()
        ctx.add_objects(s)
        times.append(time.perf_counter() - t0)
    print(f"add_object on the 1,003,520-triangle scene: {[round(t * 1e3, 3) for t in times]} ms")
    assert max(times) < 5e-3
    got = ctx.trace_closest(rays)
    assert (got["object_id"] >= 2).any()
    ctx.upload(s)
    _hits_equal(got, ctx.trace_closest(rays))


def test_invalid_addition_leaves_the_scene(ctx):
    s, cam, rp = scenes.mixed_scene(res=32)
    ctx.upload(s)
    rays = scenes.random_rays(5_000, 24, center=(0, 0.3, 0), spread=2.5)
    before = ctx.trace_closest(rays)
    bad = abi.NrObjectDesc(abi.HIT_SPHERE, 0, 999, -1, abi.F3(0, 0, 0), abi.F3(0, 0, 0), abi.F3(1, 1, 1), 0.5, 0.0)
    L = native.lib()
    assert L.nr_scene_add_object(ctx._h, C.byref(bad), None) == 1          # NR_ERR_INVALID
    assert b"material index" in L.nr_last_error()
    _hits_equal(ctx.trace_closest(rays), before)
