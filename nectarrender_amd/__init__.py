"""MI355X-native NectarRender path-tracing hot path.

libnectar_hip.so (csrc/nectar_hip.hip) is the HIP implementation behind the C
ABI of include/nectar_hip.h; `abi` mirrors that ABI in ctypes, `native` loads
it, `scenes` builds the benchmark scenes.
"""
