"""ctypes wrapper of the CPU oracle (oracle/_build/libnr_oracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker / the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from nectarrender_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libnr_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    fp = C.POINTER(C.c_float)
    sig = {
        "nro_scene_build": (vp, [C.POINTER(abi.NrSceneDesc)]),
        "nro_scene_free": (None, [vp]),
        "nro_scene_first_light": (C.c_int, [vp]),
        "nro_scene_mesh_depth": (C.c_int, [vp, C.c_int]),
        "nro_trace_closest": (C.c_int, [vp, fp, C.c_int, C.POINTER(abi.NrHit), C.c_int]),
        "nro_trace_counts": (C.c_int, [vp, fp, C.c_int, C.POINTER(C.c_int64)]),
        "nro_camera_build": (vp, [C.POINTER(abi.NrCameraDesc), C.c_uint32]),
        "nro_camera_free": (None, [vp]),
        "nro_camera_ray": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, fp]),
        "nro_nrooks": (C.c_int, [C.c_uint32, C.c_uint32, fp]),
        "nro_render_pass": (C.c_int, [vp, vp, C.POINTER(abi.NrRenderDesc), C.c_uint32, C.POINTER(C.c_int32),
                                      C.c_int, fp, fp, fp, fp, C.POINTER(C.c_int64), C.c_int]),
        "nro_render": (C.c_int, [vp, vp, C.POINTER(abi.NrRenderDesc), fp, fp, fp, fp, C.c_int]),
        "nro_accumulate": (None, [fp, fp, C.c_size_t, C.c_uint32]),
        "nro_post": (C.c_int, [fp, C.c_int, C.c_int, C.POINTER(abi.NrPostDesc), fp, C.POINTER(C.c_uint8)]),
        "nro_philox_block": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "nro_philox_uniform": (C.c_float, [C.c_uint32, C.c_uint32, C.c_uint32]),
        "nro_pcg_hash_int": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
        "nro_pcg_float": (C.c_float, [C.c_float, C.c_uint32, C.c_uint32]),
        "nro_pcg_float3": (C.c_float, [fp, C.c_uint32, C.c_uint32]),
        "nro_pcg_float_ray": (C.c_float, [fp, fp]),
        "nro_pcg_vector2": (None, [C.c_float, C.c_uint32, fp]),
        "nro_bvh_build": (C.c_int, [fp, C.c_int, C.POINTER(C.c_int32), fp, C.POINTER(C.c_int32)]),
        "nro_sky_sample": (C.c_int, [vp, fp, fp]),
        "nro_ggx_distribution": (C.c_float, [C.c_float, C.c_float]),
        "nro_pdf_pbr_value": (C.c_float, [fp, fp, C.c_float, C.c_float, fp]),
        "nro_pdf_pbr_generate": (C.c_int, [fp, fp, C.c_float, C.c_float, C.c_uint32, C.c_uint32, C.c_uint32, fp]),
        "nro_pdf_cosine_value": (C.c_float, [fp, fp]),
        "nro_pdf_cosine_generate": (C.c_int, [fp, C.c_uint32, C.c_uint32, C.c_uint32, fp]),
        "nro_mathfn": (None, [C.c_int, fp, fp, fp, C.c_int]),
        "nro_material_evaluate": (C.c_int, [vp, C.c_int, C.POINTER(abi.NrHit), fp, fp, fp]),
        "nro_material_sample": (C.c_int, [vp, C.c_int, C.POINTER(abi.NrHit), fp, C.c_uint32, C.c_uint32, C.c_uint32,
                                          fp, fp, fp, C.POINTER(C.c_int)]),
        "nro_light_pdf": (C.c_float, [vp, C.c_int, fp, fp]),
        "nro_light_random": (C.c_int, [vp, C.c_int, fp, C.c_uint32, C.c_uint32, C.c_uint32, fp]),
        "nro_texture_sample": (C.c_int, [vp, C.c_int, C.c_int, C.c_float, C.c_float, fp]),
        "nro_hash_constants": (None, [C.POINTER(C.c_uint32)]),
        "nro_viewport": (C.c_int, [vp, vp, C.c_int, fp, fp, fp, C.c_int]),
        "nro_triplanar_grid": (C.c_int, [fp, fp]),
        "nro_center_ray": (C.c_int, [vp, C.c_uint32, C.c_uint32, fp]),
        "nro_screen_ray": (C.c_int, [vp, C.c_float, C.c_float, fp]),
        "nro_pick": (C.c_int, [vp, vp, C.c_float, C.c_float, C.POINTER(abi.NrHit)]),
        "nro_selection_mask": (C.c_int, [vp, vp, C.c_int32, C.c_int, C.POINTER(C.c_uint8)]),
        "nro_overlay": (C.c_int, [C.POINTER(C.c_uint8), C.POINTER(C.c_uint8), fp, C.c_int, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def f3(v):
    return np.ascontiguousarray(v, np.float32)


class OracleScene:
    def __init__(self, scene: abi.SceneDesc):
        self._keep = scene
        d = scene.to_c()
        self._free = lib().nro_scene_free
        self.h = lib().nro_scene_build(C.byref(d))
        if not self.h:
            raise RuntimeError("oracle scene build failed")

    def __del__(self):
        # the free function is bound at construction: at interpreter exit the
        # module's globals may already be gone
        if getattr(self, "h", None):
            self._free(self.h)
            self.h = None

    def trace_closest(self, rays: np.ndarray, threads: int = 0) -> np.ndarray:
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        hits = np.empty(r.shape[0], abi.HIT_DTYPE)
        rc = lib().nro_trace_closest(self.h, _fp(r), r.shape[0], hits.ctypes.data_as(C.POINTER(abi.NrHit)), threads)
        if rc:
            raise RuntimeError("oracle traversal failed")
        return hits

    def trace_counts(self, rays: np.ndarray):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        c = np.zeros(3, np.int64)
        lib().nro_trace_counts(self.h, _fp(r), r.shape[0], c.ctypes.data_as(C.POINTER(C.c_int64)))
        return c

    def first_light(self):
        return lib().nro_scene_first_light(self.h)

    def sky(self, d):
        out = np.zeros(3, np.float32)
        lib().nro_sky_sample(self.h, _fp(f3(d)), _fp(out))
        return out

    # single-vertex probes
    def material_evaluate(self, material: int, rec, view, light):
        out = np.zeros(3, np.float32)
        if lib().nro_material_evaluate(self.h, material, C.byref(rec), _fp(f3(view)), _fp(f3(light)), _fp(out)):
            raise RuntimeError("bad material")
        return out

    def material_sample(self, material: int, rec, ray6, seed: int, subseq: int, offset: int = 0):
        """(kind, direction, pdf, attenuation, next draw): Material::scatter + sample_cosine_brdf."""
        d, a = np.zeros(3, np.float32), np.zeros(3, np.float32)
        pdf, kind = C.c_float(), C.c_int()
        nxt = lib().nro_material_sample(self.h, material, C.byref(rec), _fp(f3(ray6)), seed, subseq, offset, _fp(d),
                                        C.byref(pdf), _fp(a), C.byref(kind))
        if nxt < 0:
            raise RuntimeError("bad material")
        return kind.value, d, pdf.value, a, nxt

    def light_pdf(self, object_id: int, origin, direction) -> float:
        return lib().nro_light_pdf(self.h, object_id, _fp(f3(origin)), _fp(f3(direction)))

    def light_random(self, object_id: int, origin, seed: int, subseq: int, offset: int = 0):
        out = np.zeros(3, np.float32)
        if lib().nro_light_random(self.h, object_id, _fp(f3(origin)), seed, subseq, offset, _fp(out)):
            raise RuntimeError("no such object")
        return out

    def texture_sample(self, material: int, slot: int, u: float, v: float):
        out = np.zeros(3, np.float32)
        if lib().nro_texture_sample(self.h, material, slot, u, v, _fp(out)):
            raise RuntimeError("bad material or slot")
        return out


def hit_record(p, n, tangent=(1, 0, 0), uv=(0.5, 0.5), front=True, material=0):
    """An nr_hit for the single-vertex probes."""
    h = abi.NrHit()
    h.object_id, h.material_id, h.prim, h.front_face, h.t = 0, material, -1, 1 if front else 0, 1.0
    for k in range(3):
        h.p[k], h.n[k], h.tangent[k] = float(p[k]), float(n[k]), float(tangent[k])
    h.uv[0], h.uv[1] = float(uv[0]), float(uv[1])
    return h


def hash_constants() -> np.ndarray:
    out = (C.c_uint32 * 64)()
    lib().nro_hash_constants(out)
    return np.array(list(out), np.uint32)


class OracleCamera:
    def __init__(self, cam: abi.NrCameraDesc, seed: int):
        self.desc = cam
        self.H, self.W = int(cam.resolution[0]), int(cam.resolution[1])
        self._free = lib().nro_camera_free
        self.h = lib().nro_camera_build(C.byref(cam), seed)

    def __del__(self):
        if getattr(self, "h", None):
            self._free(self.h)
            self.h = None

    def ray(self, x, y, s, seed):
        out = np.zeros(7, np.float32)
        rc = lib().nro_camera_ray(self.h, x, y, s, seed, _fp(out))
        if rc:
            raise ValueError("sample index beyond the pattern")
        return out


def render_pass(scene: OracleScene, cam: OracleCamera, desc: abi.NrRenderDesc, s: int, pixels=None, threads=0,
                counts=False, depth=False):
    """One sample pass: (beauty, diffuse, normal[, depth][, counts]); depth is 1 x H x W."""
    b = np.zeros((3, cam.H, cam.W), np.float32)
    d = np.zeros_like(b)
    nn = np.zeros_like(b)
    z = np.zeros((1, cam.H, cam.W), np.float32)
    px = None if pixels is None else np.ascontiguousarray(pixels, np.int32)
    cnt = np.zeros(5, np.int64)
    rc = lib().nro_render_pass(scene.h, cam.h, C.byref(desc), s,
                               None if px is None else px.ctypes.data_as(C.POINTER(C.c_int32)),
                               0 if px is None else px.size, _fp(b), _fp(d), _fp(nn), _fp(z),
                               cnt.ctypes.data_as(C.POINTER(C.c_int64)) if counts else None, threads)
    if rc:
        raise RuntimeError("oracle render pass failed")
    out = (b, d, nn) + ((z,) if depth else ())
    return out + (cnt,) if counts else out


def render(scene: OracleScene, cam: OracleCamera, desc: abi.NrRenderDesc, threads=0, depth=False):
    """Running means of the passes: (beauty, diffuse, normal[, depth 1 x H x W])."""
    b = np.zeros((3, cam.H, cam.W), np.float32)
    d = np.zeros_like(b)
    nn = np.zeros_like(b)
    z = np.zeros((1, cam.H, cam.W), np.float32)
    rc = lib().nro_render(scene.h, cam.h, C.byref(desc), _fp(b), _fp(d), _fp(nn), _fp(z), threads)
    if rc:
        raise RuntimeError("oracle render failed")
    return (b, d, nn, z) if depth else (b, d, nn)


def viewport(scene: OracleScene, cam: OracleCamera, show_axis_grid: bool, threads=0):
    """trace_viewport_kernel: (beauty, diffuse, world normal), each 3 x H x W."""
    b = np.zeros((3, cam.H, cam.W), np.float32)
    d = np.zeros_like(b)
    nn = np.zeros_like(b)
    if lib().nro_viewport(scene.h, cam.h, int(bool(show_axis_grid)), _fp(b), _fp(d), _fp(nn), threads):
        raise RuntimeError("oracle viewport failed")
    return b, d, nn


def triplanar_grid(ray6):
    """(hit, rgb, alpha) of the viewport's axis grid for one ray (origin, direction)."""
    r = np.ascontiguousarray(ray6, np.float32)
    out = np.zeros(4, np.float32)
    h = lib().nro_triplanar_grid(_fp(r), _fp(out))
    return bool(h), out[:3].copy(), float(out[3])


def center_ray(cam: OracleCamera, x: int, y: int):
    out = np.zeros(6, np.float32)
    lib().nro_center_ray(cam.h, x, y, _fp(out))
    return out


def screen_ray(cam: OracleCamera, su: float, sv: float):
    out = np.zeros(6, np.float32)
    lib().nro_screen_ray(cam.h, su, sv, _fp(out))
    return out


def pick(scene: OracleScene, cam: OracleCamera, su: float, sv: float):
    h = abi.NrHit()
    lib().nro_pick(scene.h, cam.h, su, sv, C.byref(h))
    return h


def selection_mask(scene: OracleScene, cam: OracleCamera, object_id: int, radius: int):
    out = np.zeros((cam.H, cam.W), np.uint8)
    lib().nro_selection_mask(scene.h, cam.h, object_id, radius, out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out


def overlay(img_u8: np.ndarray, mask: np.ndarray, color):
    img = np.ascontiguousarray(img_u8, np.uint8).copy()
    m = np.ascontiguousarray(mask, np.uint8)
    col = np.asarray(color, np.float32)
    _, H, W = img.shape
    lib().nro_overlay(img.ctypes.data_as(C.POINTER(C.c_uint8)), m.ctypes.data_as(C.POINTER(C.c_uint8)), _fp(col), H, W)
    return img


def accumulate(avg: np.ndarray, sample: np.ndarray, s: int):
    a = np.ascontiguousarray(avg, np.float32)
    lib().nro_accumulate(_fp(a), _fp(np.ascontiguousarray(sample, np.float32)), a.size, s)
    return a


def post(chw: np.ndarray, desc: abi.NrPostDesc):
    img = np.ascontiguousarray(chw, np.float32)
    _, H, W = img.shape
    f = np.empty_like(img)
    u8 = np.empty(img.shape, np.uint8)
    lib().nro_post(_fp(img), H, W, C.byref(desc), _fp(f), u8.ctypes.data_as(C.POINTER(C.c_uint8)))
    return u8, f


def pcg_vector2(f: float, seed: int):
    out = np.zeros(2, np.float32)
    lib().nro_pcg_vector2(f, seed, _fp(out))
    return out


def philox_uniform(seed, subseq, offset):
    return lib().nro_philox_uniform(seed, subseq, offset)


def philox_block(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().nro_philox_block(c, k, o)
    return list(o)


def nrooks(n, seed):
    out = np.zeros(2 * n, np.float32)
    lib().nro_nrooks(n, seed, _fp(out))
    return out.reshape(n, 2)


def bvh_build(boxes: np.ndarray):
    b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
    n = b.shape[0]
    nodes = np.zeros((2 * n, 3), np.int32)
    nb = np.zeros((2 * n, 6), np.float32)
    order = np.zeros(n, np.int32)
    k = lib().nro_bvh_build(_fp(b), n, nodes.ctypes.data_as(C.POINTER(C.c_int32)), _fp(nb),
                            order.ctypes.data_as(C.POINTER(C.c_int32)))
    return nodes[:k], nb[:k], order


MATHFN = {"sin": 0, "cos": 1, "acos": 2, "atan2": 3, "log": 4, "pow": 5}


def mathfn(name, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(np.zeros_like(x) if y is None else y, np.float32)
    out = np.empty_like(x)
    lib().nro_mathfn(MATHFN[name], _fp(x), _fp(y), _fp(out), x.size)
    return out
