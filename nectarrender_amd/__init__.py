"""MI355X-native NectarRender path-tracing hot path.

libnectar_hip.so (csrc/nectar_hip.hip) is the HIP implementation behind the C
ABI of include/nectar_hip.h.  `_pathtracer` (csrc/bindings.cpp) is the
drop-in for the reference's pybind11 module of the same name; `abi` mirrors
the C ABI in ctypes, `native` loads it, `scenes` builds the benchmark scenes.

``import nectarrender_amd.pathtracer`` (or putting this directory on
``sys.path``) makes the module importable as top-level ``_pathtracer``, the
name the reference's Python package imports.
"""

# PyTorch ships its own copy of the HIP runtime (same soname as ROCm's).  When
# torch is present it is imported before this package's native libraries load,
# so they bind to the runtime torch already loaded and both see the GPU; loading
# ROCm's copy first leaves two runtimes in the process, and whichever starts
# second finds no device.
try:
    import torch as _torch  # noqa: F401
except ImportError:
    pass
