"""The `_pathtracer` module's Python surface against the reference's bindings.

Names, submodules, defaults and value semantics follow
csrc/bindings/bind_{core,material,hittable,lights,data,scene,engine,host}.h;
scene objects built through this API must lower to exactly the C ABI
descriptors the parity tests upload through `abi.SceneDesc`.  No GPU needed.
"""
import numpy as np
import pytest

from nectarrender_amd import abi, scenes

P = pytest.importorskip("nectarrender_amd._pathtracer")
V3 = P.core.vector.Vector3
Col = P.core.vector.Color
Mat = P.material.Material
Hit = P.hittable.Hittable
Sky = P.light.Skylight


def test_surface():
    names = {
        "core": ["Transform"], "core.vector": ["Vector2", "Vector3", "Color", "dot", "cross", "reflect", "refract"],
        "core.matrix": ["Matrix3", "rotation_from_euler"], "core.random": ["Generator"],
        "cuda": ["destroy_cublas_handle", "cudaDeviceSynchronize"],
        "material": ["Material", "MaterialType", "MaterialCore", "Lambertian", "PBR", "Dielectric", "Emissive",
                     "Isotropic"],
        "material.texture": ["Texture", "TextureType"],
        "hittable": ["Hittable", "HitRecord", "Quad", "Sphere", "Cube", "Mesh", "ObjectLight", "ConstantMedium"],
        "light": ["Skylight", "SkylightType", "SimpleSkylightConfig", "HDRISkylightConfig", "Simple", "HDRI"],
        "data": ["DataObject", "StreamState", "StreamConfig", "TonemapMethod", "DenoiseFilterQuality", "TransferStream",
                 "RenderLayersConfig", "RenderLayers", "LayerType"],
        "scene": ["Scene"],
        "engine": ["RenderEngine", "EngineType", "EngineState", "EnginePollResponse", "EngineRequests"],
        "engine.camera": ["Camera", "CameraParams"],
        "host.memory": ["CUDAMemInfo", "get_cuda_meminfo", "get_cuda_memory_used", "get_cuda_memory_free",
                        "get_cuda_memory_total"],
    }
    for mod, attrs in names.items():
        m = P
        for part in mod.split("."):
            m = getattr(m, part)
        for a in attrs:
            assert hasattr(m, a), f"_pathtracer.{mod}.{a} missing"
    eng = P.engine.RenderEngine
    for meth in ("set_scene", "idle", "render", "camera", "layers", "scene", "stream", "requests",
                 "update_streaming_layer", "get_state", "is_idle", "is_rendering", "get_engine_type",
                 "set_engine_type", "n_samples", "set_n_samples", "max_depth", "set_max_depth",
                 "set_axis_grid_visible", "on_render_started", "on_frame_finished", "on_render_finished",
                 "on_stopped", "on_restarted", "on_reset", "on_shutdown", "poll_updates"):
        assert hasattr(eng, meth), meth


def test_vector_semantics():
    v = V3(1.0, 2.0, 4.0)
    assert [(v / 2.0)[i] for i in range(3)] == [np.float32(1 / 2.0) * x for x in (1.0, 2.0, 4.0)]
    w = V3(1.0, 2.0, 4.0)
    w /= 2.0                                                                   # (t + FMIN) on /=
    assert w[0] == np.float32(np.float32(1.0) * np.float32(1.0 / np.float32(2.0 + 1e-8)))
    d = V3(1.0) / V3(0.0)                                                      # element division biased by FMIN
    assert d[0] == np.float32(1.0) / np.float32(1e-8)
    assert (V3(2.0) * 3.0)[1] == 6.0 and (V3(2.0) + 1.0)[2] == 3.0
    assert P.core.vector.dot(V3(1, 2, 3), V3(4, 5, 6)) == 32.0
    c = P.core.vector.cross(V3(1, 0, 0), V3(0, 1, 0))
    assert [c[i] for i in range(3)] == [0, 0, 1]
    r = P.core.vector.reflect(V3(1, -1, 0), V3(0, 1, 0))
    assert [r[i] for i in range(3)] == [1, 1, 0]
    t = P.core.vector.refract(V3(0, -1, 0), V3(0, 1, 0), 1.0)                 # straight through
    assert [t[i] for i in range(3)] == [0, -1, 0]
    assert Col(0.8)[0] == np.float32(0.8) and Col(0.8)[1] == 0.0              # Color(r, g=0, b=0)
    assert V3(0.2)[2] == np.float32(0.2)                                      # Vector3(e) fills


def test_matrix_camera_params():
    R = P.core.matrix.rotation_from_euler(V3(0.1, 0.2, 0.3)).numpy()
    cx, sx, cy, sy, cz, sz = (np.cos(0.1), np.sin(0.1), np.cos(0.2), np.sin(0.2), np.cos(0.3), np.sin(0.3))
    ref = (np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]]) @ np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
           @ np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]]))
    np.testing.assert_allclose(R, ref, atol=1e-6)
    cp = P.engine.camera.CameraParams(P.core.vector.Vector2(512, 256), V3(0, 1, 4), V3(-10, 0, 0), 16, 2.0, 2.0, 0.0,
                                      2.0, 1.0)
    assert cp.shutter_time == np.float32(1.0) / np.float32(1.0 + 1e-8)
    assert cp.aspect_ratio == 0.5
    cam = P.engine.camera.Camera(cp)
    # the camera position projects behind itself: None; the look-at point projects to the centre
    assert cam.project_to_screen(V3(0, 1, 5)) is None
    fwd = cp.R.forward()
    c = cam.project_to_screen(V3(0 - fwd[0], 1 - fwd[1], 4 - fwd[2]))
    assert abs(c.x() - (511.0 / 2) / 512) < 1e-5 and abs(c.y() - (255.0 / 2) / 256) < 1e-5


def test_defaults():
    q = Hit.quad()
    lw = P.scene.Scene([q], [], None).lowered()
    o = lw["objects"][0]
    assert o["scale"] == [0.0, 0.0, 0.0]                                      # bind_hittable.h default scale 0
    m = lw["materials"][o["material"]]
    assert m["type"] == abi.MAT_LAMBERTIAN and m["tex"][0][1] == [1.0, 0.0, 1.0]   # make_default_material
    pbr = Mat.pbr()
    cols = [tuple(pbr.get_tracked_texture(i).constant_color[k] for k in range(3)) for i in range(5)]
    assert cols == [(1, 1, 1), (np.float32(0.8),) * 3, (0, 0, 0), (0, 0, 0), (0.5, 0.5, 1.0)]
    sk = Sky.simple()
    assert [sk.config_simple().end[k] for k in range(3)] == [np.float32(0.5), np.float32(0.7), 1.0]
    sc = P.data.StreamConfig()
    assert (sc.apply_denoising, sc.linear_to_gamma, sc.apply_white_balance, sc.wb_temperature, sc.wb_tint,
            sc.apply_tonemapping, sc.tm_method, sc.tm_white_point, sc.tm_alpha) == \
        (True, True, True, 6800.0, 1.0, True, P.data.TonemapMethod.REINHARD, 1.0, 1.0)
    rc = P.data.RenderLayersConfig()
    assert (rc.beauty, rc.diffuse, rc.world_normal, rc.depth) == (True, True, True, False)
    assert Mat.emissive().__repr__() == "Material(type = Emissive)"


def _cornell_api():
    """The reference's config.py scene written against the module API (config.py:19-74)."""
    white = Mat.lambertian(Col.white())
    light = Hit.object_light(Hit.quad(V3(0.0, 0.499, 0.0), V3(0.0, 180.0, 0.0), V3(0.2)), 35.0, Col.white())
    hittables = [
        Hit.quad(V3(0.0, -0.5, 0.0), V3(0.0, 0.0, 0.0), V3(1.0), white),
        Hit.quad(V3(0.0, 0.5, 0.0), V3(0.0, 180.0, 0.0), V3(1.0), Mat.lambertian(Col.white())),
        Hit.quad(V3(0.5, 0.0, 0.0), V3(0.0, 0.0, 90.0), V3(1.0), Mat.lambertian(Col.red())),
        Hit.quad(V3(-0.5, 0.0, 0.0), V3(0.0, 0.0, -90.0), V3(1.0), Mat.lambertian(Col.green())),
        Hit.quad(V3(0.0, 0.0, -0.5), V3(-90.0, 0.0, 0.0), V3(1.0), Mat.lambertian(Col.white())),
        Hit.cube(V3(0.2, -0.35, 0.2), V3(0.0, 35.0, 0.0), V3(0.3), Mat.lambertian(Col.white())),
        Hit.cube(V3(-0.2, -0.2, -0.1), V3(0.0, 35.0, 0.0), V3(0.3, 0.6, 0.3), Mat.lambertian(Col.white())),
    ]
    return P.scene.Scene(hittables, [light], P.light.Skylight())


def _lowered_from_desc(s: abi.SceneDesc):
    objs = [dict(type=o["type"], bound_type=o["bound_type"], material=o["material"], mesh=o["mesh"],
                 position=[np.float32(x) for x in o["position"]], rotation=[np.float32(x) for x in o["rotation"]],
                 scale=[np.float32(x) for x in o["scale"]], radius=np.float32(o["radius"]),
                 density=np.float32(o["density"])) for o in s.objects]
    mats = []
    for m in s.materials:
        tex = []
        for k in range(5):
            t = m.tex[k] if k < len(m.tex) else abi.Texture()
            tex.append((t.image, [np.float32(x) for x in t.color], np.float32(t.scale if t.image >= 0 else 1.0)))
        mats.append(dict(type=m.type, tex=tex, tint=[np.float32(x) for x in m.tint], ior=np.float32(m.ior),
                         brightness=np.float32(m.brightness)))
    return objs, mats


def test_cornell_lowers_like_scenedesc():
    lw = _cornell_api().lowered()
    s, _, _ = scenes.cornell()
    objs, mats = _lowered_from_desc(s)
    assert len(lw["objects"]) == len(objs) == 8
    for a, b in zip(lw["objects"], objs):
        for k in b:
            assert a[k] == b[k], (k, a, b)
    assert len(lw["materials"]) == len(mats)
    for a, b in zip(lw["materials"], mats):
        assert a["type"] == b["type"]
        if a["type"] != abi.MAT_DIELECTRIC:
            assert a["tex"][0][:2] == tuple(b["tex"][0][:2]) or list(a["tex"][0][:2]) == list(b["tex"][0][:2])
        assert a["brightness"] == b["brightness"]
    assert lw["sky"]["type"] == abi.SKY_NULL


def test_mixed_types_lower(tmp_path):
    img = (np.random.default_rng(1).random((3, 8, 8)) * 255).astype(np.uint8)
    tex = P.material.texture.Texture.from_array(img, 0.5)
    objp = tmp_path / "t.obj"
    objp.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    mesh = Hit.mesh(str(objp), V3(0, 0, 0), V3(0, 0, 0), V3(1), Mat.pbr(albedo=tex))
    sph = Hit.sphere(V3(1, 0, 0), 0.5, Mat.dielectric(Col(0.9, 1, 0.9), 1.45))
    fog = Hit.constant_medium(Hit.cube(V3(0, 1, 0), V3(0, 20, 0), V3(0.5)), 2.0, Col(0.8, 0.8, 0.9))
    hdr = Sky.hdri_from_array(np.ones((3, 4, 8), np.float32), 30.0, 0.8)
    lw = P.scene.Scene([mesh, sph, fog], [], hdr).lowered()
    o = lw["objects"]
    assert [x["type"] for x in o] == [abi.HIT_MESH, abi.HIT_SPHERE, abi.HIT_CONSTANT_MEDIUM]
    assert o[0]["mesh"] == 0 and o[2]["bound_type"] == abi.HIT_CUBE and o[2]["density"] == 2.0
    assert o[1]["radius"] == 0.5 and o[1]["scale"] == [1, 1, 1]
    assert lw["materials"][o[0]["material"]]["tex"][0][0] == 0                # image texture slot
    assert lw["materials"][o[2]["material"]]["type"] == abi.MAT_ISOTROPIC
    assert lw["n_images"] == 1 and lw["n_meshes"] == 1
    assert lw["sky"]["type"] == abi.SKY_HDRI and lw["sky"]["shape"] == [3, 4, 8]
    assert lw["sky"]["rotation"] == 30.0 and lw["sky"]["intensity"] == np.float32(0.8)
    with pytest.raises(RuntimeError):
        Hit.object_light(fog, 1.0, Col.white())


def test_engine_needs_gpu_and_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    cp = P.engine.camera.CameraParams(P.core.vector.Vector2(8, 8), V3(0), V3(0), 4, 2.0, 2.0, 0.0, 2.0, 1.0)
    with pytest.raises(RuntimeError):
        P.engine.RenderEngine(P.engine.camera.Camera(cp), 4, 0)
