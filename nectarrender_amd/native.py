"""Loader and thin Python wrapper of libnectar_hip.so (include/nectar_hip.h).

The library is built in-tree (``python -c "import __graft_entry__ as g; g.build()"``).
There is no fallback: if the library is missing or no GPU is visible every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("NECTAR_HIP_LIB") or os.path.join(_HERE, "libnectar_hip.so")
_lib = None


class NectarError(RuntimeError):
    pass


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NectarError(f"HIP library not built: {LIB_PATH} is missing (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "nr_abi_version": (C.c_int, []),
        "nr_last_error": (C.c_char_p, []),
        "nr_create": (vp, [C.c_int]),
        "nr_destroy": (None, [vp]),
        "nr_scene_upload": (C.c_int, [vp, C.POINTER(abi.NrSceneDesc)]),
        "nr_camera_set": (C.c_int, [vp, C.c_int, C.POINTER(abi.NrCameraDesc), C.c_uint32]),
        "nr_aov_reset": (C.c_int, [vp, C.c_int]),
        "nr_render": (C.c_int, [vp, C.c_int, C.POINTER(abi.NrRenderDesc)]),
        "nr_render_views": (C.c_int, [vp, C.POINTER(C.c_int32), C.c_int, C.POINTER(abi.NrRenderDesc)]),
        "nr_render_tiles": (C.c_int, [vp, C.c_int, C.POINTER(abi.NrRenderDesc), C.POINTER(C.c_uint32), C.c_int]),
        "nr_set_tuning": (C.c_int, [vp, C.POINTER(abi.NrTuning)]),
        "nr_get_tuning": (C.c_int, [vp, C.POINTER(abi.NrTuning)]),
        "nr_autotune": (C.c_int, [vp, C.c_int, C.POINTER(abi.NrRenderDesc), C.POINTER(abi.NrTuning)]),
        "nr_tile_error": (C.c_int, [vp, vp, vp, vp, C.c_int32, C.c_int32, C.c_int32, vp]),
        "nr_aov_read": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_size_t]),
        "nr_aov_device_ptr": (vp, [vp, C.c_int, C.c_int]),
        "nr_trace_closest": (C.c_int, [vp, C.POINTER(C.c_float), C.c_int32, C.POINTER(abi.NrHit)]),
        "nr_trace_closest_device": (C.c_int, [vp, vp, C.c_int32, vp]),
        "nr_trace_closest_device_noalloc": (C.c_int, [vp, vp, C.c_int32, vp]),
        "nr_post": (C.c_int, [vp, C.c_int, C.POINTER(abi.NrPostDesc), C.POINTER(C.c_uint8), C.POINTER(C.c_float)]),
        "nr_sync": (C.c_int, [vp]),
        "nr_post_device": (C.c_int, [vp, vp, C.c_int32, C.c_int32, C.POINTER(abi.NrPostDesc), vp, vp]),
        "nr_view_size": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "nr_last_timings": (C.c_int, [vp, C.POINTER(C.c_float), C.c_int]),
        "nr_set_timing": (C.c_int, [vp, C.c_int]),
        "nr_viewport": (C.c_int, [vp, C.c_int, C.c_int]),
        "nr_pick": (C.c_int, [vp, C.c_int, C.c_float, C.c_float, C.POINTER(abi.NrHit)]),
        "nr_selection_mask": (C.c_int, [vp, C.c_int, C.c_int32, C.c_int32, vp]),
        "nr_overlay_device": (C.c_int, [vp, vp, C.c_int32, C.c_int32, vp, C.POINTER(C.c_float)]),
        "nr_readback": (C.c_int, [vp, C.c_int, C.POINTER(abi.NrReadbackDesc), C.POINTER(C.c_uint8)]),
        "nr_view_mask_buffer": (vp, [vp, C.c_int]),
        "nr_set_build_mode": (C.c_int, [vp, C.c_int]),
        "nr_set_trace_layout": (C.c_int, [vp, C.c_int]),
        "nr_set_render_streams": (C.c_int, [vp, C.c_int]),
        "nr_mesh_tree": (C.c_int, [vp, C.c_int, vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "nr_object_set_transform": (C.c_int, [vp, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                              C.POINTER(C.c_float)]),
        "nr_scene_add_object": (C.c_int, [vp, C.POINTER(abi.NrObjectDesc), C.POINTER(abi.NrSceneDesc)]),
        "nr_scene_set_sky": (C.c_int, [vp, C.POINTER(abi.NrSkyDesc)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.nr_abi_version() != 7:
        raise NectarError("ABI version mismatch")
    _lib = L
    return L


def _check(rc: int):
    if rc != abi.NR_OK:
        raise NectarError(f"nectar_hip error {rc}: {lib().nr_last_error().decode()}")


class Context:
    """One HIP device context: a scene, up to 16 views, their AOVs."""

    def __init__(self, device: int = 0):
        L = lib()
        # torch ships its own HIP runtime; when both share a process, torch's must
        # initialise first or it finds no device afterwards
        torch = sys.modules.get("torch")
        if torch is not None and torch.cuda.is_available():
            torch.cuda.init()
        self._h = L.nr_create(device)
        if not self._h:
            raise NectarError(f"nr_create failed: {L.nr_last_error().decode()}")
        self._scene_keep = None
        self.views = {}

    def close(self):
        if self._h:
            lib().nr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, scene: abi.SceneDesc):
        d = scene.to_c()
        self._scene_keep = scene
        _check(lib().nr_scene_upload(self._h, C.byref(d)))
        self._uploaded = scene.counts()

    def add_objects(self, scene: abi.SceneDesc):
        """Uploads the objects registered on `scene` since the last upload (with the images,
        materials and meshes they bring) by nr_scene_add_object, one partial update each."""
        if getattr(self, "_scene_keep", None) is not scene:
            raise NectarError("add_objects needs the scene of the last upload")
        n_obj, n_img, n_mat, n_mesh = self._uploaded
        for i in range(n_obj, len(scene.objects)):
            o = scene.objects[i]
            need_mat, need_mesh = o["material"] + 1, o["mesh"] + 1
            add = scene.tail_c(n_img, n_mat, n_mesh, max(n_mat, need_mat), max(n_mesh, need_mesh))
            obj = scene.object_c(i)
            _check(lib().nr_scene_add_object(self._h, C.byref(obj), C.byref(add)))
            n_img, n_mat, n_mesh = add.n_images + n_img, add.n_materials + n_mat, add.n_meshes + n_mesh
            self._uploaded = (i + 1, n_img, n_mat, n_mesh)

    def set_sky(self, scene: abi.SceneDesc):
        """Uploads `scene.sky` alone (nr_scene_set_sky)."""
        sky = scene.sky_c()
        _check(lib().nr_scene_set_sky(self._h, C.byref(sky)))

    def set_camera(self, view: int, cam: abi.NrCameraDesc, seed: int):
        _check(lib().nr_camera_set(self._h, view, C.byref(cam), seed))
        self.views[view] = (int(cam.resolution[0]), int(cam.resolution[1]))

    def reset(self, view: int = 0):
        _check(lib().nr_aov_reset(self._h, view))

    def render(self, view: int, desc: abi.NrRenderDesc):
        _check(lib().nr_render(self._h, view, C.byref(desc)))

    def set_tuning(self, slots: int = 4, stream_groups: int = 2, shade_wg_per_cu: int = 8, tail_paths: int = 262144,
                   graph_iterations: int = 0):
        _check(lib().nr_set_tuning(self._h, C.byref(abi.NrTuning(slots, stream_groups, shade_wg_per_cu, tail_paths,
                                                                   graph_iterations))))

    @staticmethod
    def _tuning_dict(t: abi.NrTuning) -> dict:
        return {"slots": t.slots, "stream_groups": t.stream_groups, "shade_wg_per_cu": t.shade_wg_per_cu,
                "tail_paths": t.tail_paths, "graph_iterations": t.graph_iterations}

    def get_tuning(self) -> dict:
        t = abi.NrTuning()
        _check(lib().nr_get_tuning(self._h, C.byref(t)))
        return self._tuning_dict(t)

    def autotune(self, view: int, desc: abi.NrRenderDesc) -> dict:
        """Times candidate tunings on `desc`, keeps the fastest (the view's AOVs are cleared)."""
        t = abi.NrTuning()
        _check(lib().nr_autotune(self._h, view, C.byref(desc), C.byref(t)))
        return self._tuning_dict(t)

    def render_tiles(self, view: int, desc: abi.NrRenderDesc, tile_end: np.ndarray, tile_size: int):
        """nr_render with a per-tile sample limit (rows x cols uint32, row major, rows =
        ceil(H / tile_size), cols = ceil(W / tile_size)).  The C entry point reads exactly
        rows x cols values, so a table of another size is refused here."""
        t = np.ascontiguousarray(tile_end, dtype=np.uint32)
        H32, W32 = C.c_int32(), C.c_int32()
        _check(lib().nr_view_size(self._h, view, C.byref(H32), C.byref(W32)))
        H, W = H32.value, W32.value
        ts = int(tile_size)
        if ts <= 0:
            raise ValueError("tile_size must be positive")
        need = -(-H // ts) * -(-W // ts)
        if t.size != need:
            raise ValueError(f"tile_end holds {t.size} limits; a {H}x{W} view in {ts}-pixel tiles needs {need}")
        _check(lib().nr_render_tiles(self._h, view, C.byref(desc), t.ctypes.data_as(C.POINTER(C.c_uint32)),
                                     int(tile_size)))

    def tile_error(self, d_denoised: int, d_judge1: int, d_judge2: int, H: int, W: int, tile_size: int, d_out: int):
        """Per-tile MSE estimate of a denoised display image against two independent plain ones (device ptrs)."""
        _check(lib().nr_tile_error(self._h, C.c_void_p(d_denoised), C.c_void_p(d_judge1), C.c_void_p(d_judge2),
                                   H, W, tile_size, C.c_void_p(d_out)))

    def render_views(self, views, desc: abi.NrRenderDesc):
        """The same passes for several views in one streaming pipeline (nr_render_views)."""
        ids = (C.c_int32 * len(views))(*map(int, views))
        _check(lib().nr_render_views(self._h, ids, len(views), C.byref(desc)))

    def sync(self):
        _check(lib().nr_sync(self._h))

    def read(self, view: int = 0, layer: int = abi.LAYER_BEAUTY) -> np.ndarray:
        """A layer of the view, (C, H, W) float32: C = 1 for DEPTH, 3 otherwise."""
        H, W = self.views[view]
        out = np.empty((1 if layer == abi.LAYER_DEPTH else 3, H, W), np.float32)
        _check(lib().nr_aov_read(self._h, view, layer, out.ctypes.data_as(C.POINTER(C.c_float)), out.size))
        return out

    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        hits = np.empty(r.shape[0], abi.HIT_DTYPE)
        _check(lib().nr_trace_closest(self._h, r.ctypes.data_as(C.POINTER(C.c_float)), r.shape[0],
                                      hits.ctypes.data_as(C.POINTER(abi.NrHit))))
        return hits

    def trace_closest_device(self, d_rays: int, n: int, d_hits: int):
        _check(lib().nr_trace_closest_device(self._h, C.c_void_p(d_rays), n, C.c_void_p(d_hits)))

    def trace_closest_device_noalloc(self, d_rays: int, n: int, d_hits: int):
        """nr_trace_closest_device with its scratch in the library's per-GPU static array."""
        _check(lib().nr_trace_closest_device_noalloc(self._h, C.c_void_p(d_rays), n, C.c_void_p(d_hits)))

    def post(self, view: int, desc: abi.NrPostDesc):
        H, W = self.views[view]
        u8 = np.empty((3, H, W), np.uint8)
        f = np.empty((3, H, W), np.float32)
        _check(lib().nr_post(self._h, view, C.byref(desc), u8.ctypes.data_as(C.POINTER(C.c_uint8)),
                             f.ctypes.data_as(C.POINTER(C.c_float))))
        return u8, f

    def aov_device_ptr(self, view: int, layer: int) -> int:
        p = lib().nr_aov_device_ptr(self._h, view, layer)
        if not p:
            raise NectarError(f"view {view} has no layer {layer}")
        return p

    def post_device(self, d_img: int, H: int, W: int, desc: abi.NrPostDesc, d_u8: int = 0, d_f32: int = 0):
        _check(lib().nr_post_device(self._h, C.c_void_p(d_img), H, W, C.byref(desc), C.c_void_p(d_u8 or None),
                                    C.c_void_p(d_f32 or None)))

    # interactive mode ------------------------------------------------------
    def viewport(self, view: int = 0, show_axis_grid: bool = True):
        """Viewport preview into the view's BEAUTY / DIFFUSE / WORLD_NORMAL layers."""
        _check(lib().nr_viewport(self._h, view, int(bool(show_axis_grid))))

    def pick(self, view: int, su: float, sv: float):
        h = abi.NrHit()
        _check(lib().nr_pick(self._h, view, su, sv, C.byref(h)))
        return h

    def selection_mask_device(self, view: int, object_id: int, radius: int, d_mask: int):
        _check(lib().nr_selection_mask(self._h, view, object_id, radius, C.c_void_p(d_mask)))

    def overlay_device(self, d_img: int, H: int, W: int, d_mask: int, color):
        col = (C.c_float * 3)(*[float(c) for c in color])
        _check(lib().nr_overlay_device(self._h, C.c_void_p(d_img), H, W, C.c_void_p(d_mask), col))

    def readback(self, view: int, post: abi.NrPostDesc, denoiser=None, d_overlay: int = 0,
                 overlay_color=(1.0, 1.0, 1.0)) -> np.ndarray:
        """TransferStream::readback: [denoise] -> post -> [overlay] -> uint8 (3, H, W)."""
        H, W = self.views[view]
        d = abi.NrReadbackDesc()
        d.post = post
        d.denoiser = denoiser
        d.d_overlay = d_overlay or None
        d.overlay_color[:] = [float(c) for c in overlay_color]
        out = np.empty((3, H, W), np.uint8)
        _check(lib().nr_readback(self._h, view, C.byref(d), out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    # scene trees and live edits -------------------------------------------
    def set_render_streams(self, n: int):
        """Stream groups of later nr_render calls (1..4)."""
        _check(lib().nr_set_render_streams(self._h, n))

    def set_trace_layout(self, layout: int):
        """Mesh layout of later trace_closest calls: 0 the binary tree, 1 the 4-wide tree, 2 the 4-wide
        tree with quantised child boxes (fast mode's)."""
        _check(lib().nr_set_trace_layout(self._h, int(layout)))

    def set_build_mode(self, mode: int):
        """Mesh tree builder of later uploads: abi.BUILD_HOST, BUILD_DEVICE or BUILD_AUTO."""
        _check(lib().nr_set_build_mode(self._h, mode))

    def mesh_tree(self, mesh: int):
        """(nodes as abi.NODE_DTYPE, slot -> original triangle, depth) of an uploaded mesh."""
        nn, nt, dep = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().nr_mesh_tree(self._h, mesh, None, None, C.byref(nn), C.byref(nt), C.byref(dep)))
        nodes = np.zeros(nn.value, abi.NODE_DTYPE)
        orig = np.zeros(nt.value, np.int32)
        _check(lib().nr_mesh_tree(self._h, mesh, nodes.ctypes.data_as(C.c_void_p),
                                  orig.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(nn), C.byref(nt), C.byref(dep)))
        return nodes, orig, dep.value

    def set_object_transform(self, index: int, position, rotation, scale):
        f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])
        _check(lib().nr_object_set_transform(self._h, index, f3(position), f3(rotation), f3(scale)))

    def set_timing(self, on: bool):
        _check(lib().nr_set_timing(self._h, int(on)))

    def timings(self):
        buf = (C.c_float * 8)()
        k = lib().nr_last_timings(self._h, buf, 8)
        return list(buf[:k])
