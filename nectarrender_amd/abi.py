"""ctypes mirror of include/nectar_hip.h plus a builder for nr_scene_desc.

The structs here are the plain-data boundary of the HIP library.  A
``SceneDesc`` collects objects, materials, images, meshes and a sky in numpy
form (registration order = object id, materials[0] = the registry's default
purple Lambertian, csrc/material/include/material/registry.h:23-26) and
produces a ``nr_scene_desc`` whose pointers stay alive as long as the
``SceneDesc`` does.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

NR_OK = 0

HIT_QUAD, HIT_SPHERE, HIT_CUBE, HIT_MESH, HIT_CONSTANT_MEDIUM, HIT_OBJECT_LIGHT = 1, 2, 3, 4, 5, 6
MAT_LAMBERTIAN, MAT_PBR, MAT_DIELECTRIC, MAT_EMISSIVE, MAT_ISOTROPIC = 1, 2, 3, 4, 5
SKY_NULL, SKY_SIMPLE, SKY_HDRI = 0, 1, 2
LAYER_BEAUTY, LAYER_DIFFUSE, LAYER_WORLD_NORMAL, LAYER_DEPTH = 0, 1, 4, 6
MODE_PARITY, MODE_FAST = 0, 1

F3 = C.c_float * 3


class NrObjectDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("bound_type", C.c_int32), ("material", C.c_int32), ("mesh", C.c_int32),
                ("position", F3), ("rotation", F3), ("scale", F3), ("radius", C.c_float), ("density", C.c_float)]


class NrTextureDesc(C.Structure):
    _fields_ = [("image", C.c_int32), ("color", F3), ("scale", C.c_float)]


class NrImageDesc(C.Structure):
    _fields_ = [("C", C.c_int32), ("H", C.c_int32), ("W", C.c_int32), ("data", C.POINTER(C.c_uint8))]


class NrMaterialDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("tex", NrTextureDesc * 5), ("tint", F3), ("ior", C.c_float),
                ("brightness", C.c_float)]


class NrMeshDesc(C.Structure):
    _fields_ = [("n_vertices", C.c_int32), ("n_triangles", C.c_int32), ("vertices", C.POINTER(C.c_float)),
                ("triangles", C.POINTER(C.c_uint32))]


class NrSkyDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("start", F3), ("end", F3), ("rotation", C.c_float), ("intensity", C.c_float),
                ("C", C.c_int32), ("H", C.c_int32), ("W", C.c_int32), ("data", C.POINTER(C.c_float))]


class NrSceneDesc(C.Structure):
    _fields_ = [("n_objects", C.c_int32), ("objects", C.POINTER(NrObjectDesc)),
                ("n_materials", C.c_int32), ("materials", C.POINTER(NrMaterialDesc)),
                ("n_images", C.c_int32), ("images", C.POINTER(NrImageDesc)),
                ("n_meshes", C.c_int32), ("meshes", C.POINTER(NrMeshDesc)),
                ("sky", NrSkyDesc)]


class NrCameraDesc(C.Structure):
    _fields_ = [("resolution", C.c_float * 2), ("position", F3), ("rotation", F3),
                ("samples_per_pixel", C.c_uint32), ("focal_length", C.c_float), ("focus_distance", C.c_float),
                ("aperture", C.c_float), ("sensor_width", C.c_float), ("shutter_speed", C.c_float)]


class NrTuning(C.Structure):
    """nr_tuning: launch tunings of nr_render (results do not depend on them)."""
    _fields_ = [("slots", C.c_int32), ("stream_groups", C.c_int32), ("shade_wg_per_cu", C.c_int32),
                ("tail_paths", C.c_int32), ("graph_iterations", C.c_int32)]


class NrRenderDesc(C.Structure):
    _fields_ = [("sample_begin", C.c_uint32), ("sample_end", C.c_uint32), ("ray_depth", C.c_uint32),
                ("seed", C.c_uint32), ("do_rr", C.c_int32), ("rr_start_bounce", C.c_uint32),
                ("rr_survival", C.c_float), ("mode", C.c_int32)]


class NrHit(C.Structure):
    _fields_ = [("object_id", C.c_int32), ("material_id", C.c_int32), ("prim", C.c_int32),
                ("front_face", C.c_int32), ("t", C.c_float), ("p", F3), ("n", F3), ("tangent", F3),
                ("uv", C.c_float * 2)]


# numpy view of nr_hit (64 bytes)
HIT_DTYPE = np.dtype([("object_id", "<i4"), ("material_id", "<i4"), ("prim", "<i4"), ("front_face", "<i4"),
                      ("t", "<f4"), ("p", "<f4", 3), ("n", "<f4", 3), ("tangent", "<f4", 3), ("uv", "<f4", 2)])
assert HIT_DTYPE.itemsize == C.sizeof(NrHit) == 64


# child-pair tree node (nr_mesh_tree): both child boxes, child refs (>= 0 node, < 0 ~leaf slot), pads
NODE_DTYPE = np.dtype([("lmin", "<f4", 3), ("lmax", "<f4", 3), ("rmin", "<f4", 3), ("rmax", "<f4", 3),
                       ("left", "<i4"), ("right", "<i4"), ("pad", "<i4", 2)])
assert NODE_DTYPE.itemsize == 64
BUILD_HOST, BUILD_DEVICE, BUILD_AUTO = 0, 1, 2


class NrPostDesc(C.Structure):
    _fields_ = [("linear_to_gamma", C.c_int32), ("apply_white_balance", C.c_int32), ("wb_temperature", C.c_float),
                ("wb_tint", C.c_float), ("apply_tonemapping", C.c_int32), ("tm_method", C.c_int32),
                ("tm_white_point", C.c_float), ("tm_alpha", C.c_float), ("fix_reinhard", C.c_int32),
                ("layer", C.c_int32)]


class NrReadbackDesc(C.Structure):
    _fields_ = [("post", NrPostDesc), ("denoiser", C.c_void_p), ("d_overlay", C.c_void_p),
                ("overlay_color", C.c_float * 3)]


def post_desc(linear_to_gamma=True, apply_white_balance=True, wb_temperature=6800.0, wb_tint=1.0,
              apply_tonemapping=True, tm_method=0, tm_white_point=1.0, tm_alpha=1.0, fix_reinhard=False,
              layer=LAYER_BEAUTY):
    """StreamConfig defaults of csrc/data/include/data/stream.h:13-30."""
    return NrPostDesc(int(linear_to_gamma), int(apply_white_balance), wb_temperature, wb_tint,
                      int(apply_tonemapping), int(tm_method), tm_white_point, tm_alpha, int(fix_reinhard),
                      int(layer))


@dataclass
class Texture:
    color: tuple = (0.0, 0.0, 0.0)
    image: int = -1
    scale: float = 1.0


@dataclass
class Material:
    type: int
    tex: list = field(default_factory=list)
    tint: tuple = (1.0, 1.0, 1.0)
    ior: float = 1.5
    brightness: float = 35.0

    @staticmethod
    def lambertian(albedo=(1.0, 1.0, 1.0)):
        return Material(MAT_LAMBERTIAN, [albedo if isinstance(albedo, Texture) else Texture(tuple(albedo))])

    @staticmethod
    def pbr(albedo=(1.0, 1.0, 1.0), roughness=0.8, metallic=0.0, emission=(0.0, 0.0, 0.0), normal=(0.5, 0.5, 1.0)):
        def tx(v):
            if isinstance(v, Texture):
                return v
            if np.isscalar(v):
                return Texture((float(v),) * 3)
            return Texture(tuple(float(x) for x in v))
        return Material(MAT_PBR, [tx(albedo), tx(roughness), tx(metallic), tx(emission), tx(normal)])

    @staticmethod
    def dielectric(tint=(1.0, 1.0, 1.0), ior=1.5):
        return Material(MAT_DIELECTRIC, [], tint=tuple(tint), ior=float(ior))

    @staticmethod
    def emissive(albedo=(1.0, 1.0, 1.0), brightness=35.0):
        return Material(MAT_EMISSIVE, [albedo if isinstance(albedo, Texture) else Texture(tuple(albedo))],
                        brightness=float(brightness))

    @staticmethod
    def isotropic(albedo=(1.0, 1.0, 1.0)):
        return Material(MAT_ISOTROPIC, [albedo if isinstance(albedo, Texture) else Texture(tuple(albedo))])


DEFAULT_MATERIAL = Material.lambertian((1.0, 0.0, 1.0))   # make_default_material, material.h:650-652


class SceneDesc:
    """Registration-ordered scene description."""

    def __init__(self):
        self.objects: list[dict] = []
        self.materials: list[Material] = [DEFAULT_MATERIAL]
        self.images: list[np.ndarray] = []
        self.meshes: list[tuple[np.ndarray, np.ndarray]] = []
        self.sky = dict(type=SKY_NULL)
        self._keep = []

    # -- registration -----------------------------------------------------
    def add_material(self, m: Material) -> int:
        self.materials.append(m)
        return len(self.materials) - 1

    def add_image(self, chw_u8: np.ndarray) -> int:
        a = np.ascontiguousarray(chw_u8, dtype=np.uint8)
        assert a.ndim == 3 and a.shape[0] >= 3
        self.images.append(a)
        return len(self.images) - 1

    def add_mesh(self, vertices: np.ndarray, triangles: np.ndarray) -> int:
        v = np.ascontiguousarray(vertices, dtype=np.float32)
        t = np.ascontiguousarray(triangles, dtype=np.uint32)
        assert v.ndim == 2 and v.shape[1] == 11 and t.ndim == 2 and t.shape[1] == 3
        self.meshes.append((v, t))
        return len(self.meshes) - 1

    def _obj(self, type_, material, position=(0, 0, 0), rotation=(0, 0, 0), scale=(1, 1, 1), radius=0.0,
             density=0.0, bound_type=0, mesh=-1):
        mid = self.add_material(material) if material is not None else 0
        self.objects.append(dict(type=type_, bound_type=bound_type, material=mid, mesh=mesh,
                                 position=tuple(map(float, position)), rotation=tuple(map(float, rotation)),
                                 scale=tuple(map(float, scale)), radius=float(radius), density=float(density)))
        return len(self.objects) - 1

    def quad(self, position, rotation, scale, material):
        return self._obj(HIT_QUAD, material, position, rotation, scale)

    def cube(self, position, rotation, scale, material):
        return self._obj(HIT_CUBE, material, position, rotation, scale)

    def sphere(self, position, radius, material):
        return self._obj(HIT_SPHERE, material, position, radius=radius)

    def mesh(self, mesh_index, position, rotation, scale, material):
        return self._obj(HIT_MESH, material, position, rotation, scale, mesh=mesh_index)

    @staticmethod
    def bound(type_, position=(0, 0, 0), rotation=(0, 0, 0), scale=(1, 1, 1), radius=0.0, mesh=-1):
        """Geometry of an object wrapped by a medium or light (not registered itself)."""
        return dict(type=type_, position=position, rotation=rotation, scale=scale, radius=radius, mesh=mesh)

    def constant_medium(self, bound: dict, density: float, albedo=(1.0, 1.0, 1.0)):
        # Hittable::constant_medium (hittable.h:221-237): material = isotropic(albedo)
        return self._obj(HIT_CONSTANT_MEDIUM, Material.isotropic(albedo), bound["position"], bound["rotation"],
                         bound["scale"], bound["radius"], density, bound_type=bound["type"], mesh=bound["mesh"])

    def object_light(self, bound: dict, brightness: float, albedo=(1.0, 1.0, 1.0)):
        # Hittable::object_light (hittable.h:251-267): material = emissive(albedo, brightness)
        return self._obj(HIT_OBJECT_LIGHT, Material.emissive(albedo, brightness), bound["position"],
                         bound["rotation"], bound["scale"], bound["radius"], bound_type=bound["type"],
                         mesh=bound["mesh"])

    def sky_simple(self, start=(1.0, 1.0, 1.0), end=(0.5, 0.7, 1.0)):
        self.sky = dict(type=SKY_SIMPLE, start=tuple(start), end=tuple(end))

    def sky_hdri(self, chw: np.ndarray, rotation=0.0, intensity=1.0):
        a = np.ascontiguousarray(chw, dtype=np.float32)
        assert a.ndim == 3 and a.shape[0] >= 3
        self.sky = dict(type=SKY_HDRI, data=a, rotation=float(rotation), intensity=float(intensity))

    # -- lowering ---------------------------------------------------------
    def counts(self):
        """(objects, images, materials, meshes) registered so far."""
        return len(self.objects), len(self.images), len(self.materials), len(self.meshes)

    def object_c(self, i: int) -> NrObjectDesc:
        o = self.objects[i]
        return NrObjectDesc(o["type"], o["bound_type"], o["material"], o["mesh"], F3(*o["position"]),
                            F3(*o["rotation"]), F3(*o["scale"]), o["radius"], o["density"])

    @staticmethod
    def _material_c(m) -> NrMaterialDesc:
        md = NrMaterialDesc()
        md.type = m.type
        for k in range(5):
            t = m.tex[k] if k < len(m.tex) else Texture()
            md.tex[k] = NrTextureDesc(int(t.image), F3(*t.color), float(t.scale))
        md.tint = F3(*m.tint)
        md.ior = m.ior
        md.brightness = m.brightness
        return md

    def sky_c(self) -> NrSkyDesc:
        sky = NrSkyDesc()
        sky.type = self.sky["type"]
        sky.intensity = 1.0
        if sky.type == SKY_SIMPLE:
            sky.start = F3(*self.sky["start"])
            sky.end = F3(*self.sky["end"])
        elif sky.type == SKY_HDRI:
            a = self.sky["data"]
            sky.rotation = self.sky["rotation"]
            sky.intensity = self.sky["intensity"]
            sky.C, sky.H, sky.W = a.shape
            sky.data = a.ctypes.data_as(C.POINTER(C.c_float))
        return sky

    def tail_c(self, img0: int, mat0: int, mesh0: int, mat1: int = None, mesh1: int = None) -> NrSceneDesc:
        """The images, materials [mat0, mat1) and meshes [mesh0, mesh1) registered after an upload,
        as the additions of nr_scene_add_object (no objects); images from img0 on, as far as
        the materials up to mat1 use them."""
        mat1 = len(self.materials) if mat1 is None else mat1
        mesh1 = len(self.meshes) if mesh1 is None else mesh1
        img1 = img0
        for m in self.materials[mat0:mat1]:
            for t in m.tex:
                img1 = max(img1, int(t.image) + 1)
        return self._lower([], self.materials[mat0:mat1], self.images[img0:img1], self.meshes[mesh0:mesh1],
                           None)

    def _lower(self, objects, materials, images, meshes, sky) -> NrSceneDesc:
        keep = []
        n = len(objects)
        objs = (NrObjectDesc * max(n, 1))()
        for i in range(n):
            objs[i] = self.object_c(objects[i])
        mats = (NrMaterialDesc * max(len(materials), 1))()
        for i, m in enumerate(materials):
            mats[i] = self._material_c(m)
        imgs = (NrImageDesc * max(len(images), 1))()
        for i, a in enumerate(images):
            imgs[i] = NrImageDesc(a.shape[0], a.shape[1], a.shape[2], a.ctypes.data_as(C.POINTER(C.c_uint8)))
        mesh_arr = (NrMeshDesc * max(len(meshes), 1))()
        for i, (v, t) in enumerate(meshes):
            mesh_arr[i] = NrMeshDesc(v.shape[0], t.shape[0], v.ctypes.data_as(C.POINTER(C.c_float)),
                                     t.ctypes.data_as(C.POINTER(C.c_uint32)))
        d = NrSceneDesc(n, objs, len(materials), mats, len(images), imgs, len(meshes), mesh_arr,
                        sky if sky is not None else NrSkyDesc())
        keep += [objs, mats, imgs, mesh_arr]
        self._keep.append(keep)
        return d

    def to_c(self) -> NrSceneDesc:
        self._keep = []
        return self._lower(list(range(len(self.objects))), self.materials, self.images, self.meshes,
                           self.sky_c())


def camera_desc(resolution=(512, 512), position=(0, 0, 0), rotation=(0, 0, 0), spp=16, focal_length=2.0,
                focus_distance=2.0, aperture=0.0, sensor_width=2.0, shutter_speed=1.0) -> NrCameraDesc:
    """CameraParams (csrc/engine/include/engine/camera.h:68-122); resolution is (H, W)."""
    return NrCameraDesc((C.c_float * 2)(*map(float, resolution)), F3(*map(float, position)),
                        F3(*map(float, rotation)), int(spp), float(focal_length), float(focus_distance),
                        float(aperture), float(sensor_width), float(shutter_speed))


def render_desc(sample_begin=1, sample_end=16, depth=8, seed=54321, do_rr=True, rr_start=3, rr_survival=0.95,
                mode=MODE_PARITY) -> NrRenderDesc:
    """TraceConfig as RenderEngine::render fills it (csrc/engine/src/engine.cpp:52-58)."""
    return NrRenderDesc(int(sample_begin), int(sample_end), int(depth), int(seed), int(do_rr), int(rr_start),
                        float(rr_survival), int(mode))
