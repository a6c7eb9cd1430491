"""The C-ABI library loads and exports every entry point include/*.h declares (CPU only)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]
LIB = os.path.join(ROOT, "nectarrender_amd", "libnectar_hip.so")


def declared():
    text = "".join(open(h).read() for h in HEADERS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nr_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("nr_create", "nr_destroy", "nr_scene_upload", "nr_camera_set", "nr_render", "nr_render_views", "nr_render_tiles", "nr_tile_error", "nr_set_tuning",
              "nr_get_tuning", "nr_autotune",
              "nr_trace_closest",
              "nr_aov_read", "nr_post", "nr_last_error", "nr_view_size", "nr_unet_create", "nr_unet_forward",
              "nr_unet_train_step", "nr_unet_gradients", "nr_denoise", "nr_viewport", "nr_pick",
              "nr_selection_mask", "nr_overlay_device", "nr_readback", "nr_set_build_mode", "nr_set_trace_layout", "nr_mesh_tree", "nr_trace_closest_device_noalloc",
              "nr_object_set_transform", "nr_view_mask_buffer", "nr_scene_add_object", "nr_scene_set_sky"):
        assert n in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "run __graft_entry__.build() first"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (\w+)$", out, flags=re.M))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_library_loads_and_reports_version():
    L = C.CDLL(LIB)
    assert L.nr_abi_version() == 7
    L.nr_last_error.restype = C.c_char_p
    assert isinstance(L.nr_last_error(), bytes)


def test_null_arguments_fail_cleanly_without_device():
    """Error paths reached before any HIP call: status codes + message, no abort."""
    from nectarrender_amd import abi, native
    L = native.lib()
    assert L.nr_scene_upload(None, None) == 1
    assert b"null" in L.nr_last_error()
    assert L.nr_render(None, 0, None) == 1
    assert L.nr_render_views(None, None, 0, None) == 1
    assert L.nr_render_tiles(None, 0, None, None, 64) == 1
    assert L.nr_set_tuning(None, None) == 1
    assert L.nr_get_tuning(None, None) == 1
    assert L.nr_autotune(None, 0, None, None) == 1
    assert L.nr_tile_error(None, None, None, None, 8, 8, 8, None) == 1
    assert L.nr_trace_closest(None, None, 0, None) == 1
    assert L.nr_trace_closest_device_noalloc(None, None, 0, None) == 1
    assert L.nr_post(None, 0, None, None, None) == 1
    assert L.nr_set_timing(None, 1) == 1
    assert L.nr_sync(None) == 1
    assert L.nr_viewport(None, 0, 1) == 1
    assert L.nr_pick(None, 0, 0.5, 0.5, None) == 1
    assert L.nr_selection_mask(None, 0, 0, 3, None) == 1
    assert L.nr_overlay_device(None, None, 4, 4, None, None) == 1
    assert L.nr_readback(None, 0, None, None) == 1
    assert not L.nr_view_mask_buffer(None, 0)
    assert L.nr_set_build_mode(None, 1) == 1
    assert L.nr_set_trace_layout(None, 1) == 1
    assert L.nr_set_render_streams(None, 2) == 1
    assert L.nr_mesh_tree(None, 0, None, None, None, None, None) == 1
    assert L.nr_object_set_transform(None, 0, None, None, None) == 1
    assert L.nr_scene_add_object(None, None, None) == 1
    assert L.nr_scene_set_sky(None, None) == 1
    from nectarrender_amd import denoise
    D = denoise._lib()
    assert D.nr_unet_param_count() == 1016419
    assert D.nr_unet_forward(None, None, None, None, None, 1, 16, 16, None) == 1
    assert D.nr_unet_train_step(None, None, None, None, None, None, 1, 16, 16, 1e-3, None) == 1
    assert D.nr_denoise(None, 0, None, None) == 1
    assert not D.nr_unet_create(0, 0) or os.path.exists("/dev/kfd")


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible")
def test_create_without_gpu_returns_null_with_message():
    from nectarrender_amd import native
    L = native.lib()
    assert not L.nr_create(0)
    assert L.nr_last_error()
