"""lightsprintsdk_amd -- MI355X-native ray casting and lightmap gather for LightsprintSDK.

The compute path is liblsgpu.so (hand-written HIP for gfx950, C ABI in
include/lsgpu.h).  This package is the host-side mirror used by tests and the
benchmark: ctypes bindings (native), the RRCollider surface (collider), the
lightmap gather/bake surface (bake) and seeded workload generators (scenes).
"""
from . import native, scenes  # noqa: F401

__all__ = ["native", "scenes"]
