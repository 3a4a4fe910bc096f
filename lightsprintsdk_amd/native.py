"""ctypes bindings of liblsgpu.so (the C ABI declared in include/lsgpu.h).

The library is the product: there is no Python or CPU fallback.  Loading fails
loudly when liblsgpu.so has not been built, and every compute entry point raises
when the library reports an error (including "no gfx950 device")."""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "liblsgpu.so")

P = ctypes.c_void_p
U32 = ctypes.c_uint32
I32 = ctypes.c_int32


class LsgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"liblsgpu error {code}: {msg}")
        self.code = code


class BuildParams(ctypes.Structure):
    _fields_ = [("leaf_size", U32), ("flags", U32)]


class ColliderInfo(ctypes.Structure):
    _fields_ = [("num_triangles", U32), ("num_valid", U32), ("num_nodes", U32), ("root_ref", I32),
                ("depth", U32), ("leaf_size", U32), ("box_min", ctypes.c_float * 3), ("box_max", ctypes.c_float * 3),
                ("node_margin", ctypes.c_float), ("device_bytes", ctypes.c_uint64), ("build_ms", ctypes.c_double),
                ("num_nodes4", U32), ("stack_need", U32)]


class Light(ctypes.Structure):
    _fields_ = [("type", I32), ("position", ctypes.c_float * 3), ("direction", ctypes.c_float * 3),
                ("color", ctypes.c_float * 3), ("cast_shadows", I32)]


class GatherParams(ctypes.Structure):
    _fields_ = [("quality", U32), ("hemisphere", I32), ("ray_length_min", ctypes.c_float),
                ("locality", ctypes.c_float), ("rug_distance", ctypes.c_float), ("light_multiplier", ctypes.c_float),
                ("seed", U32), ("light_rounds", U32), ("environment_multiplier", ctypes.c_float)]


class Lightmap(ctypes.Structure):
    _fields_ = [("rgba_base", U32), ("width", U32), ("height", U32), ("rgbf_base", U32)]


class FilterParams(ctypes.Structure):
    _fields_ = [("smoothing_amount", ctypes.c_float), ("spread_foreground", ctypes.c_float),
                ("background", ctypes.c_float * 4), ("wrap", I32)]


class BakeObject(ctypes.Structure):
    _fields_ = [("tri_uv", P), ("tri_ids", P), ("ntri", U32), ("width", U32), ("height", U32)]


class BakeTuning(ctypes.Structure):
    _fields_ = [("tile", U32), ("waves_per_block", U32), ("blocks_per_cu", U32), ("xcd_queues", U32), ("lane_refill", U32)]


class BakeInfo(ctypes.Structure):
    _fields_ = [("num_texels", U32), ("num_subtexels", U32), ("num_slots", U32), ("num_lightmaps", U32),
                ("num_triangles", U32), ("compute_units", U32), ("grow_steps", U32), ("tuning", BakeTuning),
                ("rgba", P), ("rgbf", P), ("texels", P), ("subtexels", P)]


class Texture(ctypes.Structure):
    _fields_ = [("width", U32), ("height", U32), ("format", I32), ("scaled", I32), ("data", P), ("linear", P),
                ("linear_stored", P)]


BAKE_USE_GRAPH = 1

# name -> (restype, argtypes); restype None means int status checked by check()
SIGNATURES = {
    "lsg_last_error": (ctypes.c_char_p, []),
    "lsg_version": (ctypes.c_char_p, []),
    "lsg_device_count": (ctypes.c_int, []),
    "lsg_set_device": (ctypes.c_int, [ctypes.c_int]),
    "lsg_malloc": (ctypes.c_int, [ctypes.POINTER(P), ctypes.c_size_t]),
    "lsg_free": (ctypes.c_int, [P]),
    "lsg_memcpy_h2d": (ctypes.c_int, [P, P, ctypes.c_size_t]),
    "lsg_memcpy_d2h": (ctypes.c_int, [P, P, ctypes.c_size_t]),
    "lsg_memset": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_size_t]),
    "lsg_synchronize": (ctypes.c_int, []),
    "lsg_stream_create": (ctypes.c_int, [ctypes.POINTER(P)]),
    "lsg_stream_destroy": (ctypes.c_int, [P]),
    "lsg_stream_synchronize": (ctypes.c_int, [P]),
    "lsg_event_create": (ctypes.c_int, [ctypes.POINTER(P)]),
    "lsg_event_destroy": (ctypes.c_int, [P]),
    "lsg_event_record": (ctypes.c_int, [P, P]),
    "lsg_event_elapsed_ms": (ctypes.c_int, [P, P, ctypes.POINTER(ctypes.c_float)]),
    "lsg_collider_create": (ctypes.c_int, [P, U32, P, U32, ctypes.POINTER(BuildParams), P, ctypes.POINTER(P)]),
    "lsg_collider_destroy": (None, [P]),
    "lsg_mesh_hash": (ctypes.c_int, [P, U32, P, U32, P]),
    "lsg_cache_file_name": (ctypes.c_int, [P, U32, ctypes.c_char_p, ctypes.c_char_p, P, U32]),
    "lsg_collider_save": (ctypes.c_int, [P, ctypes.c_char_p]),
    "lsg_collider_load": (ctypes.c_int, [P, U32, P, U32, ctypes.c_char_p, P]),
    "lsg_collider_create_cached": (ctypes.c_int, [P, U32, P, U32, P, P, ctypes.c_char_p, P, P]),
    "lsg_collider_info_get": (ctypes.c_int, [P, ctypes.POINTER(ColliderInfo)]),
    "lsg_collider_download": (ctypes.c_int, [P, P, P]),
    "lsg_bvh_rebuild_sah": (ctypes.c_int, [P, U32, I32, P, P, P, P]),
    "lsg_collider_download_triangles": (ctypes.c_int, [P, P, P, P, P]),
    "lsg_collider_intersect_rrrays": (ctypes.c_int, [P, P, U32, P, ctypes.POINTER(U32)]),
    "lsg_collider_intersect_device": (ctypes.c_int, [P, P, P, U32, P, P, P]),
    "lsg_collider_trace_stats": (ctypes.c_int, [P, P, P, U32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "lsg_scene_create": (ctypes.c_int, [P, P, P, P, P, U32, ctypes.POINTER(P)]),
    "lsg_scene_destroy": (None, [P]),
    "lsg_scene_set_exitance": (ctypes.c_int, [P, P]),
    "lsg_scene_get_exitance": (ctypes.c_int, [P, P]),
    "lsg_scene_exitance_device": (ctypes.c_int, [P, P, P, P, P]),
    "lsg_scene_set_lights": (ctypes.c_int, [P, P, U32]),
    "lsg_scene_set_indirect": (ctypes.c_int, [P, P, P, P]),
    "lsg_scene_set_point_materials": (ctypes.c_int, [P, P, P, P, U32, P, U32, ctypes.c_float]),
    "lsg_gather_device": (ctypes.c_int, [P, P, U32, P, U32, ctypes.POINTER(GatherParams), P, P]),
    "lsg_gather_stats": (ctypes.c_int, [P, P, U32, P, U32, ctypes.POINTER(GatherParams), P, ctypes.POINTER(ctypes.c_double)]),
    "lsg_scene_set_tangents": (ctypes.c_int, [P, P, P]),
    "lsg_scene_set_environment": (ctypes.c_int, [P, I32, U32, U32, P]),
    "lsg_scene_set_environment_blend": (ctypes.c_int, [P, I32, U32, U32, P, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "lsg_gather_layers_device": (ctypes.c_int, [P, P, U32, P, U32, ctypes.POINTER(GatherParams), P, P]),
    "lsg_filter_lightmaps": (ctypes.c_int, [P, U32, P, P, ctypes.POINTER(FilterParams), P]),
    "lsg_filter_plan_create": (ctypes.c_int, [P, U32, ctypes.POINTER(P)]),
    "lsg_filter_plan_destroy": (None, [P]),
    "lsg_filter_plan_run": (ctypes.c_int, [P, P, P, ctypes.POINTER(FilterParams), P]),
    "lsg_filter_plan_set_seams": (ctypes.c_int, [P, P, P]),
    "lsg_unwrap_seams": (ctypes.c_int, [P, U32, P, U32, P, P, P, U32, P]),
    "lsg_filter_lightmaps_noalloc": (ctypes.c_int, [P, U32, P, P, ctypes.POINTER(FilterParams), P]),
    "lsg_bake_create": (ctypes.c_int, [P, P, U32, P, P, ctypes.POINTER(P)]),
    "lsg_bake_destroy": (None, [P]),
    "lsg_bake_info_get": (ctypes.c_int, [P, ctypes.POINTER(BakeInfo)]),
    "lsg_bake_set_tuning": (ctypes.c_int, [P, ctypes.POINTER(BakeTuning)]),
    "lsg_bake_run": (ctypes.c_int, [P, ctypes.POINTER(GatherParams), ctypes.POINTER(GatherParams), ctypes.POINTER(FilterParams), U32, P]),
    "lsg_bake_gather_sample": (ctypes.c_int, [P, ctypes.POINTER(GatherParams), U32, ctypes.POINTER(ctypes.c_float)]),
    "lsg_bake_download": (ctypes.c_int, [P, U32, P]),
    "lsg_bake_set_bounces": (ctypes.c_int, [P, U32, U32]),
    "lsg_bake_set_seams": (ctypes.c_int, [P, U32, P, U32]),
    "lsg_bake_run_layers": (ctypes.c_int, [P, ctypes.POINTER(GatherParams), ctypes.POINTER(GatherParams), ctypes.POINTER(FilterParams), U32, P]),
    "lsg_bake_download_layer": (ctypes.c_int, [P, U32, U32, P]),
    "lsg_bake_download_layer_rgba": (ctypes.c_int, [P, U32, U32, P]),
    "lsg_format_bytes": (U32, [ctypes.c_int32]),
    "lsg_lightmap_encode": (ctypes.c_int, [P, U32, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_float, P, P, P]),
    "lsg_stitch_vertices": (ctypes.c_int, [P, U32, P, U32, P, ctypes.c_float, ctypes.c_float, P, P, P]),
    "lsg_vertex_plan_create": (ctypes.c_int, [P, P, P, U32, P]),
    "lsg_vertex_plan_destroy": (None, [P]),
    "lsg_vertex_plan_corners": (ctypes.c_int, [P, P, P]),
    "lsg_vertex_plan_smooth": (ctypes.c_int, [P, P, P, P]),
    "lsg_vertex_plan_gather": (ctypes.c_int, [P, P, U32, P, P]),
    "lsg_vertex_plan_triangle_data": (ctypes.c_int, [P, U32, P]),
    "lsg_mlp_num_params": (U32, []),
    "lsg_mlp_create": (ctypes.c_int, [U32, ctypes.POINTER(P)]),
    "lsg_mlp_destroy": (None, [P]),
    "lsg_mlp_set_params": (ctypes.c_int, [P, P]),
    "lsg_mlp_get_params": (ctypes.c_int, [P, P]),
    "lsg_mlp_get_grads": (ctypes.c_int, [P, P]),
    "lsg_mlp_forward": (ctypes.c_int, [P, P, U32, P, P]),
    "lsg_mlp_train_step": (ctypes.c_int, [P, P, P, U32, ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_float), P]),
    "lsg_mlp_gemm_test": (ctypes.c_int, [P, P, P, U32, U32, U32]),
    "lsg_gather_parts_device": (ctypes.c_int, [P, P, U32, P, ctypes.POINTER(GatherParams), P, P, P]),
    "lsg_learned_scale": (ctypes.c_int, [P, U32, ctypes.POINTER(ctypes.c_float)]),
    "lsg_learned_features": (ctypes.c_int, [P, P, U32, P, P, P, ctypes.c_float, P, U32, U32, P, P, P]),
    "lsg_learned_decide": (ctypes.c_int, [P, P, P, P, U32, ctypes.c_float, ctypes.c_float, P, P, P, P]),
    "lsg_learned_bound": (ctypes.c_int, [P, P, P, P, P, U32, U32, ctypes.c_float, ctypes.c_float, P,
                                         ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), P]),
    "lsg_rasterize_texels": (ctypes.c_int, [P, P, U32, U32, U32, U32, P, ctypes.POINTER(U32), P, ctypes.POINTER(U32), U32]),
}

_lib = None


def lib():
    """Loads liblsgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(hipcc --offload-arch=gfx950); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return (lib().lsg_last_error() or b"").decode(errors="replace")


def check(code: int, what: str = "") -> None:
    if code != 0:
        raise LsgError(code, f"{what}: {last_error()}" if what else last_error())


def device_count() -> int:
    return int(lib().lsg_device_count())
