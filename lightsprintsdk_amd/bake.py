"""Lightmap bake over liblsgpu: the RRSolver::updateLightmaps gather backend.

Mirrors the reference's call chain for a static scene (RRSolver/lightmap.cpp:943-1163):
  1. texel rasterisation of every lightmapped object   (lsg_rasterize_texels,
     enumerateTexelsPartial, lightmap.cpp:83-335)
  2. first gather of direct light per triangle          (lsg_gather_device with one
     subtexel covering each triangle, gatherPerTrianglePhysical, gather.cpp:843-952,
     at quality/2 as updateSolverIndirectIllumination does, gather.cpp:1085-1088)
  3. exitance table = direct irradiance * albedo + emittance (what
     PathtracerWorker::getIncidentRadiance reads at a hit, pathtracer.cpp:235-316;
     the reference's radiosity bounce on top of it is not part of this path)
  4. final gather of every texel                        (lsg_gather_device,
     processTexel, gather.cpp:600-836)
Everything between the host arrays and the RGBA texels runs on the GPU.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import native as N
from .collider import RRCollider, DeviceBuffer
from . import scenes as S

TEXEL_DTYPE = np.dtype([("out_index", "<u4"), ("sub_begin", "<u4"), ("sub_count_flags", "<u4"), ("key", "<u4")])
SUBTEXEL_DTYPE = np.dtype([("triangle", "<u4"), ("uv", "<f4", 6), ("area", "<f4")])
LIGHT_DTYPE = np.dtype([("type", "<i4"), ("position", "<f4", 3), ("direction", "<f4", 3), ("color", "<f4", 3),
                        ("cast_shadows", "<i4")])


def rasterize(tri_uv: np.ndarray, tri_ids: np.ndarray, width: int, height: int, object_id: int, out_base: int = 0):
    """SubTexels of one object's lightmap: (texels TEXEL_DTYPE, subtexels SUBTEXEL_DTYPE)."""
    uv = np.ascontiguousarray(tri_uv, dtype=np.float32)
    ids = np.ascontiguousarray(tri_ids, dtype=np.uint32)
    nt = ctypes.c_uint32(0)
    ns = ctypes.c_uint32(0)
    L = N.lib()
    N.check(L.lsg_rasterize_texels(uv.ctypes.data_as(ctypes.c_void_p), ids.ctypes.data_as(ctypes.c_void_p), len(ids),
                                   width, height, object_id, None, ctypes.byref(nt), None, ctypes.byref(ns), out_base),
            "rasterize(count)")
    tex = np.zeros(nt.value, TEXEL_DTYPE)
    sub = np.zeros(ns.value, SUBTEXEL_DTYPE)
    N.check(L.lsg_rasterize_texels(uv.ctypes.data_as(ctypes.c_void_p), ids.ctypes.data_as(ctypes.c_void_p), len(ids),
                                   width, height, object_id, tex.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nt),
                                   sub.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ns), out_base), "rasterize")
    return tex, sub


# lsg_point_material (include/lsgpu.h): a material read at the hit point
POINT_MATERIAL_DTYPE = np.dtype([("diffuse_texture", "<i4"), ("emissive_texture", "<i4"), ("transmittance_texture", "<i4"),
                                 ("flags", "<u4"), ("threshold", "<f4"), ("min_quality", "<u4"), ("albedo", "<f4", 3),
                                 ("emittance", "<f4", 3), ("emittance_multiplier", "<f4"), ("emit_sides", "<u4")])
PM_TRANSMITTANCE_IN_ALPHA, PM_KEYED, PM_INVERTED = 1, 2, 4
NO_MATERIAL = 0xFFFFFFFF
# RRBufferFormat of a texture (include/lsgpu.h LSG_FORMAT_*)
FORMAT_RGB, FORMAT_BGR, FORMAT_RGBA, FORMAT_RGBF, FORMAT_RGBAF, FORMAT_LUMINANCE, FORMAT_LUMINANCEF = 0, 1, 2, 3, 4, 9, 10
_FORMAT_DTYPE = {0: (np.uint8, 3), 1: (np.uint8, 3), 2: (np.uint8, 4), 3: (np.float32, 3), 4: (np.float32, 4),
                 9: (np.uint8, 1), 10: (np.float32, 1)}


def texture(fmt: int, data, scaled: bool = True) -> dict:
    """A texture as RRBufferInMemory stores it: data [height, width, channels] of the
    format's element type (bytes for RGB, BGR, RGBA, LUMINANCE; float32 otherwise)."""
    dt, ch = _FORMAT_DTYPE[fmt]
    a = np.ascontiguousarray(data, dt)
    if a.ndim == 2:
        a = a[:, :, None]
    if a.ndim != 3 or a.shape[2] != ch:
        raise ValueError(f"format {fmt} takes [height, width, {ch}] {np.dtype(dt).name} texels, got {a.shape} {a.dtype}")
    return dict(format=fmt, scaled=bool(scaled), data=a, width=a.shape[1], height=a.shape[0])


def lights_array(lights) -> np.ndarray:
    a = np.zeros(len(lights), LIGHT_DTYPE)
    for k, l in enumerate(lights):
        a[k]["type"] = l["type"]
        a[k]["position"] = l.get("position", (0, 0, 0))
        a[k]["direction"] = l.get("direction", (0, 0, -1))
        a[k]["color"] = l["color"]
        a[k]["cast_shadows"] = l.get("cast_shadows", 1)
    return a


class GatherParams:
    """RRSolver::UpdateParameters(quality) defaults (RRSolver.h:570-592)."""

    def __init__(self, quality=1024, hemisphere=1, ray_length_min=1e-6, locality=100000.0, rug_distance=0.001,
                 light_multiplier=1.0, seed=1, light_rounds=0, environment_multiplier=0.0):
        self.quality = quality
        self.hemisphere = hemisphere
        self.ray_length_min = ray_length_min
        self.locality = locality
        self.rug_distance = rug_distance
        self.light_multiplier = light_multiplier
        self.seed = seed
        self.light_rounds = light_rounds      # 0: quality/10+1 (gather.cpp:396)
        self.environment_multiplier = environment_multiplier   # indirect.environmentMultiplier

    def c(self):
        return N.GatherParams(self.quality, self.hemisphere, self.ray_length_min, self.locality, self.rug_distance,
                              self.light_multiplier, self.seed, self.light_rounds, self.environment_multiplier)

    def as_dict(self):
        return dict(quality=self.quality, hemisphere=self.hemisphere, ray_length_min=self.ray_length_min,
                    locality=self.locality, rug_distance=self.rug_distance, light_multiplier=self.light_multiplier,
                    seed=self.seed, light_rounds=self.light_rounds, environment_multiplier=self.environment_multiplier)


class FilterParams:
    """RRSolver::FilteringParameters defaults (RRSolver.h:614-620)."""

    def __init__(self, smoothing_amount=1.0, spread_foreground=1000.0, background=(0, 0, 0, 0), wrap=False):
        self.smoothing_amount = smoothing_amount
        self.spread_foreground = spread_foreground
        self.background = background
        self.wrap = wrap

    def c(self):
        return N.FilterParams(self.smoothing_amount, self.spread_foreground, (ctypes.c_float * 4)(*self.background),
                              1 if self.wrap else 0)


class GatherScene:
    """Collider + what the final gather reads at a hit, resident on the GPU."""

    def __init__(self, verts, idx, normals, side_bits, exitance, lights, collider: RRCollider | None = None):
        self.verts = np.ascontiguousarray(verts, np.float32)
        self.idx = np.ascontiguousarray(idx, np.uint32)
        self.col = collider or RRCollider.create(self.verts, self.idx)
        self.normals = np.ascontiguousarray(normals, np.float32)
        self.side_bits = np.ascontiguousarray(side_bits, np.uint8)
        ex = np.ascontiguousarray(exitance, np.float32).reshape(len(self.idx), 8)
        self.lights = lights_array(lights) if not isinstance(lights, np.ndarray) else np.ascontiguousarray(lights, LIGHT_DTYPE)
        h = ctypes.c_void_p()
        N.check(N.lib().lsg_scene_create(self.col._h, self.normals.ctypes.data_as(ctypes.c_void_p),
                                         self.side_bits.ctypes.data_as(ctypes.c_void_p), ex.ctypes.data_as(ctypes.c_void_p),
                                         self.lights.ctypes.data_as(ctypes.c_void_p), len(self.lights), ctypes.byref(h)),
                "lsg_scene_create")
        self._h = h.value

    def set_exitance(self, exitance):
        ex = np.ascontiguousarray(exitance, np.float32).reshape(len(self.idx), 8)
        N.check(N.lib().lsg_scene_set_exitance(self._h, ex.ctypes.data_as(ctypes.c_void_p)), "set_exitance")

    def get_exitance(self):
        out = np.zeros((len(self.idx), 8), np.float32)
        N.check(N.lib().lsg_scene_get_exitance(self._h, out.ctypes.data_as(ctypes.c_void_p)), "get_exitance")
        return out

    def set_lights(self, lights):
        self.lights = lights_array(lights) if not isinstance(lights, np.ndarray) else np.ascontiguousarray(lights, LIGHT_DTYPE)
        N.check(N.lib().lsg_scene_set_lights(self._h, self.lights.ctypes.data_as(ctypes.c_void_p), len(self.lights)),
                "set_lights")

    def gather_device(self, dtex: DeviceBuffer, ntex: int, dsub: DeviceBuffer, nsub: int, params: GatherParams,
                      dout: DeviceBuffer, stream=None):
        p = params.c()
        N.check(N.lib().lsg_gather_device(self._h, dtex.ptr, ntex, dsub.ptr, nsub, ctypes.byref(p), dout.ptr, stream),
                "lsg_gather_device")

    ENV_NONE, ENV_CUBE, ENV_PANORAMA = 0, 1, 2

    def set_environment(self, kind=0, width=0, height=0, rgb=None):
        """RRSolver::setEnvironment: kind ENV_CUBE (6 faces width x width, +x,-x,+y,-y,+z,-z)
        or ENV_PANORAMA (width x height), rgb linear float texels; ENV_NONE clears it."""
        if not kind:
            N.check(N.lib().lsg_scene_set_environment(self._h, 0, 0, 0, None), "set_environment")
            self.env = None
            return
        rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
        N.check(N.lib().lsg_scene_set_environment(self._h, kind, width, height, rgb.ctypes.data_as(ctypes.c_void_p)),
                "set_environment")
        self.env = dict(type=kind, width=width, height=height, rgb=rgb)

    def _oracle_env(self):
        """The environment as the oracle takes it (with the blend's second environment)."""
        env = getattr(self, "env", None)
        blend = getattr(self, "env_blend", None)
        if blend is None:
            return env
        env = dict(env) if env else dict(type=0, width=0, height=0, rgb=np.zeros((0, 3), np.float32))
        env["blend"] = blend
        return env

    def set_environment_blend(self, kind1=0, width1=0, height1=0, rgb1=None, angle0=0.0, angle1=0.0, blend=0.0):
        """RRSolver::setEnvironment(env0, env1, angle0, angle1) with setEnvironmentBlendFactor(blend):
        the second environment (kind1 ENV_NONE: none), the yaw of each and the blend factor;
        rays that leave the scene see env0 rotated * (1 - blend) + env1 rotated * blend
        (lsg_scene_set_environment_blend)."""
        env1 = None
        ptr = None
        if kind1:
            rgb1 = np.ascontiguousarray(rgb1, np.float32).reshape(-1, 3)
            env1 = dict(type=kind1, width=width1, height=height1, rgb=rgb1)
            ptr = rgb1.ctypes.data_as(ctypes.c_void_p)
        N.check(N.lib().lsg_scene_set_environment_blend(self._h, kind1, width1, height1, ptr, float(angle0), float(angle1),
                                                        float(blend)), "set_environment_blend")
        self.env_blend = dict(env1=env1, angle0=float(angle0), angle1=float(angle1), blend=float(blend))

    def set_tangents(self, tangents=None, bitangents=None):
        """Per-vertex tangent frames for directional lightmaps (RRMesh::TangentBasis);
        None = each vertex's frame built from its normal (buildBasisFromNormal)."""
        if tangents is None:
            N.check(N.lib().lsg_scene_set_tangents(self._h, None, None), "set_tangents")
            self.tangents = self.bitangents = None
            return
        self.tangents = np.ascontiguousarray(tangents, np.float32).reshape(len(self.verts), 3)
        self.bitangents = np.ascontiguousarray(bitangents, np.float32).reshape(len(self.verts), 3)
        N.check(N.lib().lsg_scene_set_tangents(self._h, self.tangents.ctypes.data_as(ctypes.c_void_p),
                                               self.bitangents.ctypes.data_as(ctypes.c_void_p)), "set_tangents")

    def set_indirect(self, corners=None, albedo=None, sides=None):
        """Per-corner indirect irradiance that hemisphere hits interpolate at their
        hitPoint2d and multiply by the albedo (pathtracer.cpp:326-348, rrcore.cpp:262-273):
        corners [ntri, 3 corners, 3] linear RGB, albedo [ntri, 3], sides [ntri] (bit 0 front,
        bit 1 back take the term; None = both).  corners None removes the term."""
        L = N.lib()
        if corners is None:
            N.check(L.lsg_scene_set_indirect(self._h, None, None, None), "set_indirect")
            self.ind = None
            self.side_bits = self.side_bits & np.uint8(~(8 | 16) & 0xFF)
            return
        c = np.ascontiguousarray(corners, np.float32).reshape(len(self.idx), 9)
        a = np.ascontiguousarray(albedo, np.float32).reshape(len(self.idx), 3)
        sd = None if sides is None else np.ascontiguousarray(sides, np.uint8).reshape(len(self.idx))
        N.check(L.lsg_scene_set_indirect(self._h, c.ctypes.data_as(ctypes.c_void_p), a.ctypes.data_as(ctypes.c_void_p),
                                         sd.ctypes.data_as(ctypes.c_void_p) if sd is not None else None), "set_indirect")
        # the records as the library lays them out: corner k rgb, then albedo channel k
        rec = np.zeros((len(self.idx), 3, 4), np.float32)
        rec[:, :, :3] = c.reshape(-1, 3, 3)
        rec[:, :, 3] = a
        self.ind = rec.reshape(-1, 12)
        b = np.full(len(self.idx), 3, np.uint8) if sd is None else sd
        self.side_bits = (self.side_bits & np.uint8(~(8 | 16) & 0xFF)) | np.where(b & 1, 8, 0).astype(np.uint8) \
            | np.where(b & 2, 16, 0).astype(np.uint8)

    def set_point_materials(self, tri_material=None, tri_uv=None, materials=None, textures=(), gamma=0.45):
        """Point materials (RRObject::getPointMaterial): tri_material [ntri] uint32 material
        index or NO_MATERIAL, tri_uv [ntri, 6] material uv of the corners, materials a
        POINT_MATERIAL_DTYPE array, textures a list of texture() dicts, gamma the solver's
        sRGB power (0 = no colour space).  tri_material None removes them.  A triangle with
        a material keeps its direct irradiance in the exitance table."""
        L = N.lib()
        if tri_material is None:
            N.check(L.lsg_scene_set_point_materials(self._h, None, None, None, 0, None, 0, 0.0), "set_point_materials")
            self.pm = None
            return
        tm = np.ascontiguousarray(tri_material, np.uint32).reshape(len(self.idx))
        uv = np.ascontiguousarray(tri_uv, np.float32).reshape(len(self.idx), 6)
        mats = np.ascontiguousarray(materials, POINT_MATERIAL_DTYPE).reshape(-1)
        texs = [t if isinstance(t, dict) else texture(*t) for t in textures]
        arr = (N.Texture * max(len(texs), 1))()
        for k, t in enumerate(texs):
            arr[k] = N.Texture(t["width"], t["height"], t["format"], int(t["scaled"]), t["data"].ctypes.data_as(ctypes.c_void_p),
                               None, None)
        N.check(L.lsg_scene_set_point_materials(self._h, tm.ctypes.data_as(ctypes.c_void_p), uv.ctypes.data_as(ctypes.c_void_p),
                                                mats.ctypes.data_as(ctypes.c_void_p), len(mats),
                                                ctypes.cast(arr, ctypes.c_void_p), len(texs), float(gamma)),
                "set_point_materials")
        self.pm = dict(tri_mat=tm, tri_uv=uv, mats=mats, textures=texs, gamma=float(gamma))

    def gather_layers_device(self, dtex: DeviceBuffer, ntex: int, dsub: DeviceBuffer, nsub: int, params: GatherParams,
                             douts, stream=None):
        """processTexel into the output layers of updateLightmap: douts = 5 DeviceBuffers
        (lightmap, directions 1-3, bent normals) or None for a layer not wanted."""
        p = params.c()
        ptrs = (ctypes.c_void_p * 5)(*[(d.ptr if d is not None else None) for d in douts])
        N.check(N.lib().lsg_gather_layers_device(self._h, dtex.ptr, ntex, dsub.ptr, nsub, ctypes.byref(p), ptrs, stream),
                "lsg_gather_layers_device")

    def gather_layers(self, texels, subs, params: GatherParams, layers=(0, 1, 2, 3, 4), nout=None):
        """Host convenience: returns {layer: float32 [nout,4]}."""
        if nout is None:
            nout = int(texels["out_index"].max()) + 1 if len(texels) else 0
        dt = DeviceBuffer.from_array(texels)
        ds = DeviceBuffer.from_array(subs)
        outs = [None] * 5
        for k in layers:
            outs[k] = DeviceBuffer(16 * max(nout, 1))
            N.check(N.lib().lsg_memset(outs[k].ptr, 0, 16 * max(nout, 1)), "memset")
        self.gather_layers_device(dt, len(texels), ds, len(subs), params, outs)
        N.check(N.lib().lsg_synchronize(), "sync")
        return {k: outs[k].download(np.float32, (nout, 4)) for k in layers}

    def gather(self, texels, subs, params: GatherParams, nout=None):
        """Host convenience: uploads texels/subtexels, gathers, returns float32 [nout,4]."""
        if nout is None:
            nout = int(texels["out_index"].max()) + 1 if len(texels) else 0
        dt = DeviceBuffer.from_array(texels)
        ds = DeviceBuffer.from_array(subs)
        do = DeviceBuffer(16 * max(nout, 1))
        N.check(N.lib().lsg_memset(do.ptr, 0, 16 * max(nout, 1)), "memset")
        self.gather_device(dt, len(texels), ds, len(subs), params, do)
        N.check(N.lib().lsg_synchronize(), "sync")
        return do.download(np.float32, (nout, 4))

    def oracle_arrays(self, exitance):
        return dict(verts=self.verts, idx=self.idx, normals=self.normals, side_bits=self.side_bits,
                    tangents=getattr(self, "tangents", None), bitangents=getattr(self, "bitangents", None),
                    env=self._oracle_env(), ind=getattr(self, "ind", None), pm=getattr(self, "pm", None),
                    exitance=np.ascontiguousarray(exitance, np.float32).reshape(len(self.idx), 8),
                    lights=np.array([(l["type"], l["position"], l["direction"], l["color"], l["cast_shadows"])
                                     for l in self.lights], dtype=[("type", "<i4"), ("pos", "<f4", 3), ("dir", "<f4", 3),
                                                                   ("color", "<f4", 3), ("cast_shadows", "<i4")]))

    def close(self):
        if getattr(self, "_h", None):
            N.lib().lsg_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def triangle_texels(ntri: int, tri_ids=None):
    """One texel per triangle whose single subtexel is the whole triangle: the
    first-gather work unit of gatherPerTrianglePhysical (gather.cpp:905-940)."""
    ids = np.arange(ntri, dtype=np.uint32) if tri_ids is None else np.asarray(tri_ids, np.uint32)
    tex = np.zeros(len(ids), TEXEL_DTYPE)
    tex["out_index"] = np.arange(len(ids))
    tex["sub_begin"] = np.arange(len(ids))
    tex["sub_count_flags"] = 1 | (255 << 16)
    tex["key"] = 0x80000000 | ids
    sub = np.zeros(len(ids), SUBTEXEL_DTYPE)
    sub["triangle"] = ids
    sub["uv"] = np.array([0, 0, 1, 0, 0, 1], np.float32)
    sub["area"] = 1.0
    return tex, sub


class SceneBake:
    """A whole BakeScene on the GPU: rasterised texels of every lightmapped object,
    the exitance table and the per-object output slots."""

    def __init__(self, scene: S.BakeScene, params: GatherParams | None = None):
        self.scene = scene
        self.params = params or GatherParams(ray_length_min=S.minimal_safe_distance(scene.verts))
        nt = scene.ntri
        ex0 = np.zeros((nt, 8), np.float32)
        ex0[:, 0:3] = scene.emittance
        self.gs = GatherScene(scene.verts, scene.idx, scene.normals, scene.side_bits, ex0, scene.lights)
        texs, subs = [], []
        base = 0
        sub_base = 0
        self.slots = []
        bases = np.cumsum([0] + [ob.width * ob.height for ob in scene.objects])
        # every object's lightmap is rasterised on the GPU (lsg_rasterize_texels), in object order
        def one(k):
            ob = scene.objects[k]
            ids = np.arange(ob.tri_begin, ob.tri_begin + ob.tri_count, dtype=np.uint32)
            return rasterize(scene.tri_uv[ob.tri_begin:ob.tri_begin + ob.tri_count], ids, ob.width, ob.height, k,
                             int(bases[k]))
        parts = [one(k) for k in range(len(scene.objects))]
        for k, ob in enumerate(scene.objects):
            t, s = parts[k]
            t["sub_begin"] += sub_base
            texs.append(t)
            subs.append(s)
            self.slots.append((base, ob.width, ob.height))
            base += ob.width * ob.height
            sub_base += len(s)
        del parts
        self.texels = np.concatenate(texs) if texs else np.zeros(0, TEXEL_DTYPE)
        self.subs = np.concatenate(subs) if subs else np.zeros(0, SUBTEXEL_DTYPE)
        self.nout = base
        self.maps = (N.Lightmap * max(len(self.slots), 1))()
        for k, (b, w, h) in enumerate(self.slots):
            self.maps[k] = N.Lightmap(b, w, h, b)
        self.filtering = FilterParams()
        self.d_rgbf = DeviceBuffer(12 * max(self.nout, 1))
        self.d_tex = DeviceBuffer.from_array(self.texels)
        self.d_sub = DeviceBuffer.from_array(self.subs)
        self.d_out = DeviceBuffer(16 * max(self.nout, 1))
        N.check(N.lib().lsg_memset(self.d_out.ptr, 0, 16 * max(self.nout, 1)), "memset")
        ttex, tsub = triangle_texels(nt)
        self.d_ttex = DeviceBuffer.from_array(ttex)
        self.d_tsub = DeviceBuffer.from_array(tsub)
        self.d_tout = DeviceBuffer(16 * max(nt, 1))
        self.first_params = GatherParams(quality=max(self.params.quality // 2, 1), hemisphere=0,
                                         ray_length_min=self.params.ray_length_min, locality=self.params.locality,
                                         rug_distance=self.params.rug_distance,
                                         light_multiplier=self.params.light_multiplier, seed=self.params.seed ^ 0x5bd1e995)
        self.d_albedo = DeviceBuffer.from_array(np.ascontiguousarray(scene.albedo, np.float32))
        self.d_emit = DeviceBuffer.from_array(np.ascontiguousarray(scene.emittance, np.float32))

    @property
    def exitance(self):
        """The scene's current exitance table, float32 [ntri, 8], read back from the GPU."""
        return self.gs.get_exitance()

    def first_gather(self, stream=None):
        """Per-triangle direct irradiance -> exitance table, on the GPU.  The direct
        illumination includes the environment (updateSolverDirectIllumination reads
        indirect.environmentMultiplier, gather.cpp:958-975)."""
        nt = self.scene.ntri
        self.first_params.environment_multiplier = self.params.environment_multiplier
        L = N.lib()
        # hemisphere rays of the first gather see emitters only, not an earlier result
        N.check(L.lsg_scene_exitance_device(self.gs._h, None, self.d_albedo.ptr, self.d_emit.ptr, stream), "exitance")
        self.gs.gather_device(self.d_ttex, nt, self.d_tsub, nt, self.first_params, self.d_tout, stream)
        N.check(L.lsg_scene_exitance_device(self.gs._h, self.d_tout.ptr, self.d_albedo.ptr, self.d_emit.ptr, stream),
                "exitance")

    def final_gather(self, stream=None):
        self.gs.gather_device(self.d_tex, len(self.texels), self.d_sub, len(self.subs), self.params, self.d_out, stream)

    def _filter_plan(self):
        """The lightmap batch and its filter scratch, kept in HBM between bakes."""
        if getattr(self, "_fplan", None) is None:
            h = ctypes.c_void_p()
            N.check(N.lib().lsg_filter_plan_create(self.maps, len(self.slots), ctypes.byref(h)), "lsg_filter_plan_create")
            self._fplan = h.value
        return self._fplan

    def set_seams(self, seams):
        """Unwrap seams per lightmap (a list of float32 [n, 8] arrays, object_seams), smoothed
        by every later filter(); None removes them."""
        plan = self._filter_plan()
        if seams is None:
            N.check(N.lib().lsg_filter_plan_set_seams(plan, None, None), "lsg_filter_plan_set_seams")
            return
        assert len(seams) == len(self.slots)
        arrs = [np.ascontiguousarray(a, np.float32).reshape(-1, 8) for a in seams]
        begin = np.cumsum([0] + [len(a) for a in arrs]).astype(np.uint32)
        allv = np.ascontiguousarray(np.concatenate(arrs) if arrs else np.zeros((0, 8), np.float32))
        N.check(N.lib().lsg_filter_plan_set_seams(plan, allv.ctypes.data_as(ctypes.c_void_p) if len(allv) else None,
                                                  begin.ctypes.data_as(ctypes.c_void_p)), "lsg_filter_plan_set_seams")

    def filter(self, stream=None, rgba=None, rgbf=None):
        """updateLightmap's filter chain + RGBF writeback on every lightmap (GPU)."""
        fp = self.filtering.c()
        N.check(N.lib().lsg_filter_plan_run(self._filter_plan(), (rgba or self.d_out).ptr, (rgbf or self.d_rgbf).ptr,
                                            ctypes.byref(fp), stream), "lsg_filter_plan_run")

    def __del__(self):
        try:
            if getattr(self, "_fplan", None):
                N.lib().lsg_filter_plan_destroy(self._fplan)
                self._fplan = None
        except Exception:
            pass

    def clear(self):
        """Empty every lightmap slot (the filters grow into uncovered slots in place, and
        the reference gathers each updateLightmaps pass into a fresh buffer)."""
        N.check(N.lib().lsg_memset(self.d_out.ptr, 0, 16 * max(self.nout, 1)), "memset")

    def bake(self, stream=None):
        """One updateLightmaps pass after the first gather: final gather + filters."""
        self.clear()
        self.final_gather(stream)
        self.filter(stream)

    def bake_layers(self, layers=(0, 1, 2, 3, 4), stream=None):
        """updateLightmaps with directional and bent-normal layers: one layered final
        gather, then the filter chain + RGBF writeback on each layer separately
        (lightmap.cpp:806-876).  Returns {layer: [per-object RGBF maps]}."""
        n = max(self.nout, 1)
        outs, rgbfs = [None] * 5, [None] * 5
        for k in layers:
            outs[k] = self.d_out if k == 0 else DeviceBuffer(16 * n)
            rgbfs[k] = self.d_rgbf if k == 0 else DeviceBuffer(12 * n)
            N.check(N.lib().lsg_memset(outs[k].ptr, 0, 16 * n), "memset")
        self.gs.gather_layers_device(self.d_tex, len(self.texels), self.d_sub, len(self.subs), self.params, outs, stream)
        res = {}
        for k in layers:
            self.filter(stream, rgba=outs[k], rgbf=rgbfs[k])
        N.check(N.lib().lsg_synchronize(), "sync")
        for k in layers:
            out = rgbfs[k].download(np.float32, (n, 3))
            res[k] = [out[b:b + w * h].reshape(h, w, 3) for b, w, h in self.slots]
        return res

    def lightmaps_rgbf(self):
        N.check(N.lib().lsg_synchronize(), "sync")
        out = self.d_rgbf.download(np.float32, (max(self.nout, 1), 3))
        return [out[b:b + w * h].reshape(h, w, 3) for b, w, h in self.slots]

    def lightmaps_rgba(self):
        N.check(N.lib().lsg_synchronize(), "sync")
        out = self.d_out.download(np.float32, (max(self.nout, 1), 4))
        maps = []
        for base, w, h in self.slots:
            maps.append(out[base:base + w * h].reshape(h, w, 4))
        return maps


# ---------------------------------------------------------------------------------------
# per-vertex outputs (updateLightmap into vertex buffers, lightmap.cpp:766-800)

def stitch_vertices(verts, idx, corner_normals, weld_distance=0.0, max_smooth_angle=0.01):
    """The solver's vertex stitching (lsg_stitch_vertices, RRMesh::createOptimizedVertices
    with SmoothingParameters' vertexWeldDistance and |maxSmoothAngle|, defaults 0 and 0.01):
    (group of every vertex, 0xFFFFFFFF when unused; representative vertex of every group)."""
    v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    i = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
    n = np.ascontiguousarray(corner_normals, np.float32).reshape(-1, 9)
    assert len(n) == len(i)
    group = np.empty(len(v), np.uint32)
    rep = np.empty(max(len(v), 1), np.uint32)
    ng = ctypes.c_uint32()
    N.check(N.lib().lsg_stitch_vertices(v.ctypes.data_as(ctypes.c_void_p), len(v), i.ctypes.data_as(ctypes.c_void_p), len(i),
                                        n.ctypes.data_as(ctypes.c_void_p), float(weld_distance), float(abs(max_smooth_angle)),
                                        group.ctypes.data_as(ctypes.c_void_p), rep.ctypes.data_as(ctypes.c_void_p),
                                        ctypes.byref(ng)), "lsg_stitch_vertices")
    return group, rep[:ng.value].copy()


def unwrap_seams(verts, idx, corner_normals, tri_uv) -> np.ndarray:
    """The unwrap seams lightmapSmooth smooths for one object (lsg_unwrap_seams,
    UnwrapSeams, RRBuffer.cpp:536-625): float32 [n, 8] per seam, the edge's two ends in
    the first triangle's unwrap (u0, v0, u1, v1), then in the second's."""
    v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    i = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
    n = np.ascontiguousarray(corner_normals, np.float32).reshape(-1, 9)
    uv = np.ascontiguousarray(tri_uv, np.float32).reshape(-1, 6)
    assert len(n) == len(i) and len(uv) == len(i)
    args = [v.ctypes.data_as(ctypes.c_void_p), len(v), i.ctypes.data_as(ctypes.c_void_p), len(i),
            n.ctypes.data_as(ctypes.c_void_p), uv.ctypes.data_as(ctypes.c_void_p)]
    cnt = ctypes.c_uint32()
    N.check(N.lib().lsg_unwrap_seams(*args, None, 0, ctypes.byref(cnt)), "lsg_unwrap_seams")
    out = np.zeros((cnt.value, 8), np.float32)
    if cnt.value:
        N.check(N.lib().lsg_unwrap_seams(*args, out.ctypes.data_as(ctypes.c_void_p), cnt.value, ctypes.byref(cnt)),
                "lsg_unwrap_seams")
    return out


def object_seams(scene: S.BakeScene, k: int) -> np.ndarray:
    """The unwrap seams of lightmapped object k of a scene: its own triangles, their
    vertices and per-vertex normals (the object's mesh, as lightmapSmooth is given it)."""
    ob = scene.objects[k]
    tri = np.ascontiguousarray(scene.idx[ob.tri_begin:ob.tri_begin + ob.tri_count], np.uint32)
    used, inv = np.unique(tri, return_inverse=True)
    return unwrap_seams(scene.verts[used], inv.reshape(-1, 3), scene.normals[tri].reshape(-1, 9),
                        scene.tri_uv[ob.tri_begin:ob.tri_begin + ob.tri_count])


class VertexPlan:
    """Per-vertex lighting of a scene (lsg_vertex_plan_*): the final gather of one texel per
    triangle (gatherPerTrianglePhysical, gather.cpp:843-952) smoothed into the stitched
    vertices (updateVertexBufferFromPerTriangleDataPhysical, vertexBuffer.cpp:262-290)."""

    def __init__(self, gs: GatherScene, scene: S.BakeScene, weld_distance=0.0, max_smooth_angle=0.01):
        self.gs, self.scene = gs, scene
        idx = np.ascontiguousarray(scene.idx, np.uint32)
        self.vertex_group, self.group_rep = stitch_vertices(scene.verts, idx, scene.normals[idx].reshape(-1, 9),
                                                            weld_distance, max_smooth_angle)
        self.ngroups = len(self.group_rep)
        self.corner_group = np.ascontiguousarray(self.vertex_group[idx].reshape(-1), np.uint32)
        self.group_pos = np.ascontiguousarray(scene.verts[self.group_rep], np.float32).reshape(-1, 3)
        h = ctypes.c_void_p()
        N.check(N.lib().lsg_vertex_plan_create(gs._h, self.corner_group.ctypes.data_as(ctypes.c_void_p),
                                               self.group_pos.ctypes.data_as(ctypes.c_void_p), self.ngroups,
                                               ctypes.byref(h)), "lsg_vertex_plan_create")
        self._h = h.value
        self._out = {}

    def corners(self):
        """(group each corner counts for, 0xFFFFFFFF = none; its power = angle x area)"""
        nc = 3 * self.scene.ntri
        key = np.empty(max(nc, 1), np.uint32)
        power = np.empty(max(nc, 1), np.float32)
        N.check(N.lib().lsg_vertex_plan_corners(self._h, key.ctypes.data_as(ctypes.c_void_p),
                                                power.ctypes.data_as(ctypes.c_void_p)), "corners")
        return key[:nc], power[:nc]

    def _buf(self, k):
        if k not in self._out:
            self._out[k] = DeviceBuffer(16 * max(self.ngroups, 1))
        return self._out[k]

    def gather(self, params: GatherParams, layers=(0,), stream=None):
        """The per-triangle final gather of each layer smoothed into the groups: {layer: [ngroups, 4]}."""
        outs = (ctypes.c_void_p * 5)()           # one slot per LSG_LAYER_*
        mask = 0
        for k in layers:
            mask |= 1 << k
            outs[k] = self._buf(k).ptr
        p = params.c()
        N.check(N.lib().lsg_vertex_plan_gather(self._h, ctypes.byref(p), mask, outs, stream), "lsg_vertex_plan_gather")
        N.check(N.lib().lsg_synchronize(), "sync")
        return {k: self._buf(k).download(np.float32, (self.ngroups, 4)) for k in layers}

    def triangle_data(self, layer=0):
        out = np.empty((max(self.scene.ntri, 1), 4), np.float32)
        N.check(N.lib().lsg_vertex_plan_triangle_data(self._h, layer, out.ctypes.data_as(ctypes.c_void_p)), "triangle data")
        return out[:self.scene.ntri]

    def smooth(self, tri_data):
        """Smooths host per-triangle values [ntri, 4] into the groups: [ngroups, 4]."""
        d = DeviceBuffer.from_array(np.ascontiguousarray(tri_data, np.float32).reshape(-1, 4))
        o = DeviceBuffer(16 * max(self.ngroups, 1))
        N.check(N.lib().lsg_vertex_plan_smooth(self._h, d.ptr, o.ptr, None), "lsg_vertex_plan_smooth")
        N.check(N.lib().lsg_synchronize(), "sync")
        return o.download(np.float32, (self.ngroups, 4))

    def object_vertices(self, k, group_values):
        """Object k's vertex buffer: (its vertex indices in ascending order, [n, 3] values).  A
        vertex takes its group's value when a triangle the smoothing counts uses it, else 0
        (the solver's vertex lookup, vertexBuffer.cpp:21-76)."""
        ob = self.scene.objects[k]
        tri = np.arange(ob.tri_begin, ob.tri_begin + ob.tri_count)
        key, _ = self.corners()
        corners = (3 * tri[:, None] + np.arange(3)).reshape(-1)
        verts = self.scene.idx[tri].reshape(-1)
        ids = np.unique(verts)
        counted = np.zeros(len(self.scene.verts), bool)
        counted[verts[key[corners] != 0xFFFFFFFF]] = True
        vals = np.zeros((len(ids), 3), np.float32)
        m = counted[ids]
        vals[m] = group_values[self.vertex_group[ids[m]], :3]
        return ids, vals

    def close(self):
        if getattr(self, "_h", None):
            N.lib().lsg_vertex_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def box_room_scene(lightmap_size=256):
    """The box room: closed 10x10x4 m room + 20 seeded boxes, ~50k triangles, 21 lightmaps."""
    return S.box_room(seed=3, n_boxes=20, target_tris=50_000, lightmap_size=lightmap_size)


# ---------------------------------------------------------------------------------------
# whole-scene bake plan (lsg_bake_*) and its autotuner

class BakePlan:
    """All lightmaps of a BakeScene baked by one plan: persistent final gather over
    every texel of every object, optional hipGraph capture/replay."""

    def __init__(self, scene: S.BakeScene, params: GatherParams | None = None, filtering: "FilterParams | None" = None):
        self.scene = scene
        self.params = params or GatherParams(ray_length_min=S.minimal_safe_distance(scene.verts))
        self.first_params = GatherParams(quality=max(self.params.quality // 2, 1), hemisphere=0,
                                         ray_length_min=self.params.ray_length_min, locality=self.params.locality,
                                         rug_distance=self.params.rug_distance,
                                         light_multiplier=self.params.light_multiplier, seed=self.params.seed ^ 0x5bd1e995)
        self.filtering = filtering or FilterParams()
        ex0 = np.zeros((scene.ntri, 8), np.float32)
        ex0[:, 0:3] = scene.emittance
        self.gs = GatherScene(scene.verts, scene.idx, scene.normals, scene.side_bits, ex0, scene.lights)
        self._keep = []
        objs = (N.BakeObject * max(len(scene.objects), 1))()
        for k, ob in enumerate(scene.objects):
            uv = np.ascontiguousarray(scene.tri_uv[ob.tri_begin:ob.tri_begin + ob.tri_count], np.float32)
            ids = np.arange(ob.tri_begin, ob.tri_begin + ob.tri_count, dtype=np.uint32)
            self._keep += [uv, ids]
            objs[k] = N.BakeObject(uv.ctypes.data_as(ctypes.c_void_p), ids.ctypes.data_as(ctypes.c_void_p), ob.tri_count,
                                   ob.width, ob.height)
        alb = np.ascontiguousarray(scene.albedo, np.float32)
        emi = np.ascontiguousarray(scene.emittance, np.float32)
        h = ctypes.c_void_p()
        N.check(N.lib().lsg_bake_create(self.gs._h, objs, len(scene.objects), alb.ctypes.data_as(ctypes.c_void_p),
                                        emi.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h)), "lsg_bake_create")
        self._h = h.value

    def info(self) -> N.BakeInfo:
        inf = N.BakeInfo()
        N.check(N.lib().lsg_bake_info_get(self._h, ctypes.byref(inf)), "bake info")
        return inf

    def set_tuning(self, tile=0, waves_per_block=0, blocks_per_cu=0, xcd_queues=0, lane_refill=0):
        t = N.BakeTuning(tile, waves_per_block, blocks_per_cu, xcd_queues, lane_refill)
        N.check(N.lib().lsg_bake_set_tuning(self._h, ctypes.byref(t)), "set_tuning")

    def run(self, first=True, graph=False, filtering=True, stream=None):
        self.first_params.environment_multiplier = self.params.environment_multiplier
        fp = self.first_params.c()
        p = self.params.c()
        flt = self.filtering.c()
        N.check(N.lib().lsg_bake_run(self._h, ctypes.byref(fp) if first else None, ctypes.byref(p),
                                     ctypes.byref(flt) if filtering else None, N.BAKE_USE_GRAPH if graph else 0, stream),
                "lsg_bake_run")

    def set_seams(self, k, seams):
        """Unwrap seams of object k's lightmap (float32 [n, 8], object_seams), smoothed by
        the filter chain of later runs; an empty array removes them."""
        a = np.ascontiguousarray(seams, np.float32).reshape(-1, 8)
        N.check(N.lib().lsg_bake_set_seams(self._h, k, a.ctypes.data_as(ctypes.c_void_p) if len(a) else None, len(a)),
                "lsg_bake_set_seams")

    def set_bounces(self, bounces=0, quality=0):
        """Radiosity on the GPU (lsg_bake_set_bounces): after the first gather, `bounces`
        hemisphere-only per-triangle gathers of `quality` rays (0 = the first gather's)
        spread the light between triangles before the final gather."""
        N.check(N.lib().lsg_bake_set_bounces(self._h, bounces, quality), "set_bounces")
        self.bounces, self.bounce_quality = bounces, quality

    def exitance(self):
        """The scene's current exitance table, float32 [ntri, 8] (front rgb+pad, back rgb+pad)."""
        return self.gs.get_exitance()

    def run_layers(self, layers=(0, 1, 2, 3, 4), first=True, filtering=True, stream=None):
        """The plan's bake with directional and bent-normal layers (lsg_bake_run_layers)."""
        self.first_params.environment_multiplier = self.params.environment_multiplier
        fp = self.first_params.c()
        p = self.params.c()
        flt = self.filtering.c()
        mask = 0
        for k in layers:
            mask |= 1 << k
        N.check(N.lib().lsg_bake_run_layers(self._h, ctypes.byref(fp) if first else None, ctypes.byref(p),
                                            ctypes.byref(flt) if filtering else None, mask, stream), "lsg_bake_run_layers")

    def lightmap_layer(self, k, layer):
        ob = self.scene.objects[k]
        out = np.zeros((ob.height, ob.width, 3), np.float32)
        N.check(N.lib().lsg_bake_download_layer(self._h, k, layer, out.ctypes.data_as(ctypes.c_void_p)), "bake download")
        return out

    def lightmap_layer_rgba(self, k, layer=0):
        """Object k's layer as the filters left it, linear RGBA with the background's alpha
        (lsg_bake_download_layer_rgba)."""
        ob = self.scene.objects[k]
        out = np.zeros((ob.height, ob.width, 4), np.float32)
        N.check(N.lib().lsg_bake_download_layer_rgba(self._h, k, layer, out.ctypes.data_as(ctypes.c_void_p)), "bake download")
        return out

    def sample_ms(self, stride):
        ms = ctypes.c_float()
        p = self.params.c()
        N.check(N.lib().lsg_bake_gather_sample(self._h, ctypes.byref(p), stride, ctypes.byref(ms)), "gather_sample")
        return ms.value

    def lightmap(self, k):
        ob = self.scene.objects[k]
        out = np.zeros((ob.height, ob.width, 3), np.float32)
        N.check(N.lib().lsg_bake_download(self._h, k, out.ctypes.data_as(ctypes.c_void_p)), "bake download")
        return out

    def rgba(self):
        inf = self.info()
        buf = np.empty((max(inf.num_slots, 1), 4), np.float32)
        N.check(N.lib().lsg_synchronize(), "sync")
        N.check(N.lib().lsg_memcpy_d2h(buf.ctypes.data_as(ctypes.c_void_p), inf.rgba, buf.nbytes), "d2h")
        return buf

    def close(self):
        if getattr(self, "_h", None):
            N.lib().lsg_bake_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


TUNING_CANDIDATES = [
    dict(tile=t, waves_per_block=w, blocks_per_cu=b, xcd_queues=x)
    for t in (1, 4) for w in (4, 8) for b in (2, 4, 8) for x in (1, 8)
    if w * b <= 32
]


def autotune(plan: BakePlan, sample_texels: int = 60000, candidates=None, cache_path: str | None = None):
    """Times the persistent final gather on an evenly strided texel sample for each
    launch shape and keeps the fastest.  The choice is written to cache_path (JSON,
    keyed by texel count, triangle count and quality) and reused when present."""
    import json
    import os
    info = plan.info()
    key = f"tex{info.num_texels}_tri{info.num_triangles}_q{plan.params.quality}_cu{info.compute_units}"
    if cache_path and os.path.exists(cache_path):
        try:
            cache = json.load(open(cache_path))
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            plan.set_tuning(**cache[key]["tuning"])
            return cache[key]
    else:
        cache = {}
    stride = max(1, info.num_texels // max(sample_texels, 1))
    results = []
    for c in candidates or TUNING_CANDIDATES:
        plan.set_tuning(**c)
        ms = min(plan.sample_ms(stride) for _ in range(2))
        results.append((ms, c))
    results.sort(key=lambda r: r[0])
    best_ms, best = results[0]
    # then the lane schedule on the best launch shape: 64 shooters at a time (1) or
    # lanes taking shooters as they free up (2, the default the shapes were timed with)
    if candidates is None:
        alt = dict(best, lane_refill=1)
        plan.set_tuning(**alt)
        ms = min(plan.sample_ms(stride) for _ in range(2))
        results.append((ms, alt))
        if ms < best_ms:
            best_ms, best = ms, alt
        results.sort(key=lambda r: r[0])
    plan.set_tuning(**best)
    entry = dict(tuning=best, sample_ms=round(best_ms, 4), stride=stride,
                 worst_ms=round(results[-1][0], 4), tried=len(results))
    if cache_path:
        cache[key] = entry
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return entry
