"""Python mirror of the RRCollider plug-in surface over liblsgpu.

Mirrors include/Lightsprint/RRCollider.h: RRCollider.create(mesh, objects,
technique, aborting, cacheLocation, buildParams) returns None on failure
(RRCollider.cpp:216-239 falls back to IT_LINEAR; here the caller decides),
intersect_batch() is the re-added RRCollider::intersectBatch
(doc/doxygen/doc.news:3212) over arrays of rays.

Rays are float32 [n,8] = (origin.xyz, rayLengthMin, dir.xyz, rayLengthMax).
Hits come back as a structured array (triangle, distance, point2d, side).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import native as N

IT_LINEAR, IT_BSP_COMPACT, IT_BSP_FAST, IT_BSP_FASTER, IT_BSP_FASTEST, IT_BVH_COMPACT, IT_BVH_FAST, IT_VERIFICATION = range(8)
IT_HIP_LBVH = 8   # technique number registered for this backend

FILL_DISTANCE, FILL_POINT3D, FILL_POINT2D, FILL_PLANE, FILL_TRIANGLE, FILL_SIDE = (1 << k for k in range(6))

HIT_DTYPE = np.dtype([("tri", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])

# RRRay, include/Lightsprint/RRCollider.h:85-137 (LP64, 128 bytes, align 16)
RRRAY_DTYPE = np.dtype({
    "names": ["rayOrigin", "rayDir", "rayLengthMin", "rayLengthMax", "rayFlags", "hitDistance", "hitTriangle",
              "hitPoint2d", "hitPlane", "hitPoint3d", "hitFrontSide", "hitPadding1", "hitPadding2", "hitObject",
              "collisionHandler"],
    "formats": [("<f4", 4), ("<f4", 4), "<f4", "<f4", "<u4", "<f4", "<u4", ("<f4", 2), ("<f4", 4), ("<f4", 3), "u1",
                ("<f4", 3), ("<f4", 2), "<u8", "<u8"],
    "offsets": [0, 16, 32, 36, 40, 44, 48, 52, 60, 76, 88, 92, 104, 112, 120],
    "itemsize": 128,
})


def make_rrrays(rays: np.ndarray, flags: int = 0x3F) -> np.ndarray:
    """RRRay records as RRRay::RRRay() initialises them (RRRay.cpp:19-25), then
    filled from float32 [n,8] rays."""
    out = np.zeros(len(rays), RRRAY_DTYPE)
    out["rayOrigin"][:, :3] = rays[:, 0:3]
    out["rayOrigin"][:, 3] = 1.0
    out["rayDir"][:, :3] = rays[:, 4:7]
    out["rayLengthMin"] = rays[:, 3]
    out["rayLengthMax"] = rays[:, 7]
    out["rayFlags"] = flags
    return out


def rebuild_sah(tree: dict) -> dict:
    """lsg_bvh_rebuild_sah on a tree dict of download_tree's form (host only, no GPU):
    the inner nodes rebuilt over the same leaves by binned SAH."""
    nodes = np.ascontiguousarray(tree["nodes"], dtype=np.float32).reshape(-1, 16)
    out = np.zeros((max(len(nodes), 1), 16), np.float32)
    nn, root, depth = ctypes.c_uint32(), ctypes.c_int32(), ctypes.c_uint32()
    N.check(N.lib().lsg_bvh_rebuild_sah(nodes.ctypes.data_as(ctypes.c_void_p), len(nodes), int(tree["root"]),
                                        out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nn), ctypes.byref(root),
                                        ctypes.byref(depth)), "lsg_bvh_rebuild_sah")
    return dict(tree, nodes=out[:nn.value], root=int(root.value), depth=int(depth.value))


class DeviceBuffer:
    """Device allocation owned by Python (freed on close/GC)."""

    def __init__(self, nbytes: int):
        p = ctypes.c_void_p()
        N.check(N.lib().lsg_malloc(ctypes.byref(p), int(nbytes)), "lsg_malloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        N.check(N.lib().lsg_memcpy_h2d(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes), "h2d")

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        N.check(N.lib().lsg_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes), "d2h")
        return out

    def close(self):
        if getattr(self, "ptr", None):
            N.lib().lsg_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceRays:
    """Rays resident in HBM as two float4 arrays plus a hit array."""

    def __init__(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        self.n = len(rays)
        self.orig = DeviceBuffer.from_array(np.ascontiguousarray(rays[:, 0:4]))
        self.dir = DeviceBuffer.from_array(np.ascontiguousarray(rays[:, 4:8]))
        self.hits = DeviceBuffer(16 * max(self.n, 1))
        self.front = DeviceBuffer(max(self.n, 1))

    def download_hits(self):
        h = self.hits.download(HIT_DTYPE, (self.n,))
        f = self.front.download(np.uint8, (self.n,))
        return h, f


def mesh_hash(verts: np.ndarray, idx: np.ndarray) -> bytes:
    """RRMesh::getHash's SHA-1 of a mesh (lsg_mesh_hash)."""
    v = np.ascontiguousarray(verts, dtype=np.float32)
    i = np.ascontiguousarray(idx, dtype=np.uint32)
    out = (ctypes.c_uint8 * 20)()
    N.check(N.lib().lsg_mesh_hash(v.ctypes.data_as(ctypes.c_void_p), len(v), i.ctypes.data_as(ctypes.c_void_p), len(i), out),
            "lsg_mesh_hash")
    return bytes(out)


def cache_file_name(hash20: bytes, version: int, location=None, extension=None) -> str:
    """RRHash::getFileName of a hash (lsg_cache_file_name)."""
    h = (ctypes.c_uint8 * 20).from_buffer_copy(hash20)
    out = ctypes.create_string_buffer(4096)
    N.check(N.lib().lsg_cache_file_name(h, version, None if location is None else str(location).encode(),
                                        None if extension is None else str(extension).encode(), out, 4096),
            "lsg_cache_file_name")
    return out.value.decode()


class RRCollider:
    """Collider built by liblsgpu (technique IT_HIP_LBVH)."""

    def __init__(self, handle, mesh_verts, mesh_idx):
        self._h = handle
        self.verts = mesh_verts
        self.idx = mesh_idx

    @staticmethod
    def create(verts: np.ndarray, idx: np.ndarray, technique: int = IT_HIP_LBVH, aborting=None,
               cacheLocation="*", buildParams=None, leaf_size: int = 0, lbvh_only: bool = False):
        """RRCollider::create for a mesh given as float32 [nv,3] / uint32 [nt,3].  lbvh_only keeps
        the LBVH's inner nodes instead of rebuilding them by SAH.  cacheLocation: a directory
        prefix for tree files (lsg_collider_create_cached), None for the temp directory's
        "Lightsprint/" as the reference's default; "*" (this binding's default) builds without
        a cache.  from_cache tells whether the tree came from a file."""
        v = np.ascontiguousarray(verts, dtype=np.float32)
        i = np.ascontiguousarray(idx, dtype=np.uint32)
        h = ctypes.c_void_p()
        bp = N.BuildParams(leaf_size, 1 if lbvh_only else 0)
        ab = None
        if aborting is not None:
            ab = ctypes.cast(ctypes.pointer(aborting), ctypes.c_void_p)
        fc = ctypes.c_int32(0)
        if cacheLocation == "*":
            code = N.lib().lsg_collider_create(v.ctypes.data_as(ctypes.c_void_p), len(v), i.ctypes.data_as(ctypes.c_void_p),
                                               len(i), ctypes.byref(bp), ab, ctypes.byref(h))
        else:
            loc = None if cacheLocation is None else str(cacheLocation).encode()
            code = N.lib().lsg_collider_create_cached(v.ctypes.data_as(ctypes.c_void_p), len(v), i.ctypes.data_as(ctypes.c_void_p),
                                                      len(i), ctypes.byref(bp), ab, loc, ctypes.byref(h), ctypes.byref(fc))
        N.check(code, "lsg_collider_create")
        col = RRCollider(h.value, v, i)
        col.from_cache = bool(fc.value)
        return col

    @staticmethod
    def load(verts: np.ndarray, idx: np.ndarray, path: str):
        """A collider from a tree file written by save() for the same mesh (lsg_collider_load)."""
        v = np.ascontiguousarray(verts, dtype=np.float32)
        i = np.ascontiguousarray(idx, dtype=np.uint32)
        h = ctypes.c_void_p()
        N.check(N.lib().lsg_collider_load(v.ctypes.data_as(ctypes.c_void_p), len(v), i.ctypes.data_as(ctypes.c_void_p), len(i),
                                          str(path).encode(), ctypes.byref(h)), "lsg_collider_load")
        col = RRCollider(h.value, v, i)
        col.from_cache = True
        return col

    def save(self, path: str):
        N.check(N.lib().lsg_collider_save(self._h, str(path).encode()), "lsg_collider_save")

    def info(self) -> N.ColliderInfo:
        inf = N.ColliderInfo()
        N.check(N.lib().lsg_collider_info_get(self._h, ctypes.byref(inf)), "info")
        return inf

    def getTechnique(self):
        return IT_HIP_LBVH

    def getMemoryOccupied(self):
        return int(self.info().device_bytes)

    def download_tree(self):
        inf = self.info()
        nodes = np.zeros((max(inf.num_nodes, 1), 16), np.float32)
        prim = np.zeros(max(inf.num_valid, 1), np.uint32)
        N.check(N.lib().lsg_collider_download(self._h, nodes.ctypes.data_as(ctypes.c_void_p),
                                              prim.ctypes.data_as(ctypes.c_void_p)), "download")
        return dict(nodes=nodes[:inf.num_nodes], prim=prim[:inf.num_valid], root=int(inf.root_ref),
                    depth=int(inf.depth), nvalid=int(inf.num_valid))

    def intersect_device(self, rays: DeviceRays, stream=None):
        N.check(N.lib().lsg_collider_intersect_device(self._h, rays.orig.ptr, rays.dir.ptr, rays.n, rays.hits.ptr,
                                                      rays.front.ptr, stream), "intersect_device")

    def intersect_batch(self, rays: np.ndarray):
        """Closest hits of float32 [n,8] rays; returns (hits structured array, front uint8)."""
        d = DeviceRays(rays)
        self.intersect_device(d)
        N.check(N.lib().lsg_synchronize(), "sync")
        return d.download_hits()

    def intersect_rrrays(self, rrrays: np.ndarray):
        """intersectBatch over RRRay records (in place); returns hit mask."""
        assert rrrays.dtype == RRRAY_DTYPE and rrrays.flags["C_CONTIGUOUS"]
        mask = np.zeros(len(rrrays), np.uint8)
        nh = ctypes.c_uint32(0)
        N.check(N.lib().lsg_collider_intersect_rrrays(self._h, rrrays.ctypes.data_as(ctypes.c_void_p), len(rrrays),
                                                      mask.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nh)),
                "intersect_rrrays")
        return mask

    def trace_stats(self, rays: DeviceRays):
        nv = ctypes.c_double()
        nt = ctypes.c_double()
        N.check(N.lib().lsg_collider_trace_stats(self._h, rays.orig.ptr, rays.dir.ptr, rays.n, ctypes.byref(nv),
                                                 ctypes.byref(nt)), "trace_stats")
        return nv.value / max(rays.n, 1), nt.value / max(rays.n, 1)

    def close(self):
        if self._h:
            N.lib().lsg_collider_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
