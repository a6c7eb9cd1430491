"""ctypes wrapper of oracle/_ref/libnr_ioref.so (test infrastructure only).

The library is the reference's vendored stb_image / tinyobjloader built from
the reference tree by `make -C oracle ref`, plus a restatement of the
reference's glue (texture.h:8-47, mesh.h:39-141).  Absent when the reference
tree is not available; tests that need it skip.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libnr_ioref.so")
_lib = None


def available() -> bool:
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB)
        ip = C.POINTER(C.c_int)
        L.ioref_image_u8.restype = C.POINTER(C.c_uint8)
        L.ioref_image_u8.argtypes = [C.c_char_p, ip, ip, ip]
        L.ioref_stbi_u8.restype = C.POINTER(C.c_uint8)
        L.ioref_stbi_u8.argtypes = [C.c_char_p, ip, ip, ip]
        L.ioref_image_f32.restype = C.POINTER(C.c_float)
        L.ioref_image_f32.argtypes = [C.c_char_p, ip, ip, ip]
        L.ioref_obj.restype = C.c_int
        L.ioref_obj.argtypes = [C.c_char_p, ip, ip, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.POINTER(C.c_uint32))]
        L.ioref_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def image_u8(path: str):
    """(planes C x H x W uint8, C) as load_image_uint8 returns them."""
    c, h, w = C.c_int(), C.c_int(), C.c_int()
    p = lib().ioref_image_u8(path.encode(), C.byref(c), C.byref(h), C.byref(w))
    if not p:
        raise RuntimeError("stb could not load " + path)
    n = c.value * h.value * w.value
    out = np.ctypeslib.as_array(p, (n,)).copy().reshape(c.value, h.value, w.value)
    lib().ioref_free(p)
    return out, c.value


def stbi_u8(path: str) -> np.ndarray:
    """stbi_load(path, ..., 0) as the library returns it: H x W x C uint8 (any C)."""
    c, h, w = C.c_int(), C.c_int(), C.c_int()
    p = lib().ioref_stbi_u8(path.encode(), C.byref(c), C.byref(h), C.byref(w))
    if not p:
        raise RuntimeError("stb could not load " + path)
    n = c.value * h.value * w.value
    out = np.ctypeslib.as_array(p, (n,)).copy().reshape(h.value, w.value, c.value)
    lib().ioref_free(p)
    return out


def image_f32(path: str) -> np.ndarray:
    c, h, w = C.c_int(), C.c_int(), C.c_int()
    p = lib().ioref_image_f32(path.encode(), C.byref(c), C.byref(h), C.byref(w))
    if not p:
        raise RuntimeError("stb could not load " + path)
    n = c.value * h.value * w.value
    out = np.ctypeslib.as_array(p, (n,)).copy().reshape(c.value, h.value, w.value)
    lib().ioref_free(p)
    return out


def obj(path: str):
    nv, nt = C.c_int(), C.c_int()
    v, t = C.POINTER(C.c_float)(), C.POINTER(C.c_uint32)()
    if lib().ioref_obj(path.encode(), C.byref(nv), C.byref(nt), C.byref(v), C.byref(t)) != 0:
        raise RuntimeError("tinyobjloader could not load " + path)
    verts = np.ctypeslib.as_array(v, (nv.value * 11,)).copy().reshape(nv.value, 11) if nv.value else np.zeros((0, 11), np.float32)
    tris = np.ctypeslib.as_array(t, (nt.value * 3,)).copy().reshape(nt.value, 3) if nt.value else np.zeros((0, 3), np.uint32)
    lib().ioref_free(v)
    lib().ioref_free(t)
    return verts, tris
