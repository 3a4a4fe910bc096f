"""CPU checks of the drop-in boundary: liblsgpu.so loads without a GPU and
exports every entry point include/lsgpu.h declares; the Python bindings cover
them; argument validation fails cleanly without touching a device."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "lsgpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lsg_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("lsg_collider_create", "lsg_collider_intersect_rrrays", "lsg_collider_intersect_device",
                 "lsg_gather_device", "lsg_scene_create", "lsg_rasterize_texels"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from lightsprintsdk_amd import native
    L = native.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    unbound = [n for n in declared() if n not in native.SIGNATURES]
    assert not unbound, unbound


def test_invalid_arguments_without_device():
    from lightsprintsdk_amd import native
    L = native.lib()
    out = ctypes.c_void_p()
    assert L.lsg_collider_create(None, 0, None, 0, None, None, None) == -3   # out is null
    idx = np.array([0, 1, 5], np.uint32)
    v = np.zeros((3, 3), np.float32)
    code = L.lsg_collider_create(v.ctypes.data_as(ctypes.c_void_p), 3, idx.ctypes.data_as(ctypes.c_void_p), 1,
                                 None, None, ctypes.byref(out))
    assert code == -3 and "out of range" in native.last_error()
    assert L.lsg_collider_info_get(None, None) == -3
    # the allocation-free filter chain checks its size limit on the host, before any device call
    params = native.FilterParams(1.0, 1000.0, (ctypes.c_float * 4)(0, 0, 0, 0), 0)
    big = (native.Lightmap * 1)(native.Lightmap(0, 1025, 1024, 0))
    code = L.lsg_filter_lightmaps_noalloc(big, 1, ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.byref(params), None)
    assert code == -6 and "at most 1048576" in native.last_error()
    many = (native.Lightmap * 4097)()
    assert L.lsg_filter_lightmaps_noalloc(many, 4097, ctypes.c_void_p(16), None, ctypes.byref(params), None) == -6


def test_device_count_is_safe():
    from lightsprintsdk_amd import native
    assert native.device_count() >= 0
