"""ctypes bindings of the CPU oracle (oracle/liblsg_oracle.so) and a runner for
the compiled reference collider (oracle/_ref/ref_collider).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes
import json
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liblsg_oracle.so")
REF_BIN = os.path.join(ORACLE_DIR, "_ref", "ref_collider")


def _sources():
    """The oracle's C files, as oracle/Makefile lists them (planted-mistake tests rebuild
    all of them with one file changed)."""
    with open(os.path.join(ORACLE_DIR, "Makefile")) as f:
        for line in f:
            if line.startswith("ORACLE_SRC"):
                return line.split(":=", 1)[1].split()
    raise RuntimeError("ORACLE_SRC not found in oracle/Makefile")


SOURCES = _sources()

HIT_DTYPE = np.dtype([("tri", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("p", "<f4", 3),
                      ("plane", "<f4", 4), ("front", "<i4")])
TRI_DTYPE = np.dtype([("n", "<f4", 4), ("s", "<f4", 3), ("ireal", "<f4"), ("r", "<f4", 3),
                      ("ir", "<f4", 3), ("l", "<f4", 3), ("byte", "<i4")])


class Stats(ctypes.Structure):
    _fields_ = [("node_visits", ctypes.c_double), ("tri_tests", ctypes.c_double), ("rays", ctypes.c_double)]


class GParams(ctypes.Structure):
    _fields_ = [("quality", ctypes.c_uint32), ("hemisphere", ctypes.c_int32), ("ray_length_min", ctypes.c_float),
                ("locality", ctypes.c_float), ("rug_distance", ctypes.c_float), ("light_multiplier", ctypes.c_float),
                ("seed", ctypes.c_uint32), ("light_rounds", ctypes.c_uint32), ("environment_multiplier", ctypes.c_float)]


class Env(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("w", ctypes.c_uint32), ("h", ctypes.c_uint32), ("rgb", ctypes.c_void_p),
                ("env1", ctypes.c_void_p), ("angle0", ctypes.c_float), ("angle1", ctypes.c_float), ("blend", ctypes.c_float),
                ("blended", ctypes.c_int32)]


ENV_CUBE, ENV_PANORAMA = 1, 2


def _env(scene_arrays):
    """scene_arrays['env'] = dict(type, width, height, rgb) or absent -> (Env or None, keep-alive)."""
    e = scene_arrays.get("env")
    if not e:
        return None, None
    rgb = np.ascontiguousarray(e["rgb"], dtype=np.float32)
    ev = Env(e["type"], e["width"], e["height"], rgb.ctypes.data_as(ctypes.c_void_p))
    keep = [rgb]
    b = e.get("blend")       # dict(env1=dict or None, angle0, angle1, blend): RRSolver's second environment
    if b:
        if b.get("env1"):
            e1, k1 = _env(dict(env=b["env1"]))
            keep += [e1, k1]
            ev.env1 = ctypes.cast(ctypes.pointer(e1), ctypes.c_void_p)
        ev.angle0, ev.angle1, ev.blend, ev.blended = b["angle0"], b["angle1"], b["blend"], 1
    return ev, keep


def env_index(env, d):
    """Texel index the oracle's environment lookup picks for direction d."""
    L = lib()
    L.lsgo_env_index.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.lsgo_env_index.restype = ctypes.c_uint32
    ev, keep = _env(dict(env=env))
    dd = np.ascontiguousarray(d, dtype=np.float32)
    return int(L.lsgo_env_index(ctypes.byref(ev), _p(dd)))


class GStats(ctypes.Structure):
    _fields_ = [("hemi_rays", ctypes.c_double), ("shadow_rays", ctypes.c_double), ("series", ctypes.c_double)]


LIGHT_DTYPE = np.dtype([("type", "<i4"), ("pos", "<f4", 3), ("dir", "<f4", 3), ("color", "<f4", 3), ("cast_shadows", "<i4")])

_lib = None


def have_oracle() -> bool:
    return os.path.exists(ORACLE_LIB)


def build_oracle(ref: bool = True) -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    if ref and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-j8", "-C", ORACLE_DIR, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not have_oracle():
            build_oracle(ref=False)
        L = ctypes.CDLL(ORACLE_LIB)
        P = ctypes.c_void_p
        U = ctypes.c_uint32
        L.lsgo_srlnp_setup.argtypes = [P, P, P, P]
        L.lsgo_srlnp_hit.argtypes = [P, P, P, ctypes.c_float, ctypes.c_float, P, P, P, P]
        L.lsgo_srlnp_hit.restype = ctypes.c_int
        L.lsgo_mesh_aabb.argtypes = [P, U, P, P]
        L.lsgo_box_init.argtypes = [P, P]
        L.lsgo_box_intersect.argtypes = [P, P, P, P, ctypes.c_float, ctypes.c_float, P, P]
        L.lsgo_box_intersect.restype = ctypes.c_int
        L.lsgo_delta_bsp.argtypes = [P, P]
        L.lsgo_delta_bsp.restype = ctypes.c_float
        L.lsgo_intersect_linear_srlnp.argtypes = [P, U, P, U, P, U, P, ctypes.c_int]
        L.lsgo_intersect_mt.argtypes = [P, U, P, U, P, U, P, ctypes.c_int, ctypes.c_int]
        L.lsgo_bvh_traverse.argtypes = [P, U, ctypes.c_int32, P, U, P, U, P, U, P, U, P, P, ctypes.c_int]
        L.lsgo_lbvh_build.argtypes = [P, U, P, U, U, P, P, P, P]
        L.lsgo_lbvh_build.restype = U
        L.lsgo_filler_triangle_point.argtypes = [U, P, P]
        L.lsgo_filler_circle_point.argtypes = [P, P, P, P]
        L.lsgo_filler_circle_point.restype = U
        L.lsgo_gather.argtypes = [P, U, P, U, P, P, P, P, U, P, P, U, P, P, ctypes.c_int32, P, U, P, P, ctypes.c_int]
        L.lsgo_texel_key.argtypes = [U, U]
        L.lsgo_texel_key.restype = U
        L.lsgo_jitter.argtypes = [U, U, P, P]
        L.lsgo_subtexel_walk.argtypes = [P, U, P, U, P]
        L.lsgo_subtexel_walk.restype = None
        L.lsgo_shots_before.argtypes = [U, U, U]
        L.lsgo_shots_before.restype = U
        L.lsgo_vertex_corners.argtypes = [P, P, U, P, P, P, P]
        L.lsgo_vertex_corners.restype = None
        L.lsgo_smooth_vertices.argtypes = [P, P, U, P, U, P]
        L.lsgo_smooth_vertices.restype = None
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _mesh(verts, idx):
    v = np.ascontiguousarray(verts, dtype=np.float32)
    i = np.ascontiguousarray(idx, dtype=np.uint32)
    return v, i


def srlnp_setup(v0, v1, v2):
    out = np.zeros(1, TRI_DTYPE)
    a, b, c = (np.ascontiguousarray(x, dtype=np.float32) for x in (v0, v1, v2))
    lib().lsgo_srlnp_setup(_p(a), _p(b), _p(c), _p(out))
    return out[0]


def mesh_box(verts):
    mn = np.zeros(3, np.float32)
    mx = np.zeros(3, np.float32)
    v = np.ascontiguousarray(verts, dtype=np.float32)
    lib().lsgo_mesh_aabb(_p(v), len(v), _p(mn), _p(mx))
    lib().lsgo_box_init(_p(mn), _p(mx))
    return mn, mx


def box_intersect(mn, mx, o, d, rmin, rmax):
    lo = np.zeros(1, np.float32)
    hi = np.zeros(1, np.float32)
    mn, mx, o, d = (np.ascontiguousarray(x, dtype=np.float32) for x in (mn, mx, o, d))
    r = lib().lsgo_box_intersect(_p(mn), _p(mx), _p(o), _p(d), rmin, rmax, _p(lo), _p(hi))
    return bool(r), float(lo[0]), float(hi[0])


def intersect_linear_srlnp(verts, idx, rays, threads=0):
    v, i = _mesh(verts, idx)
    r = np.ascontiguousarray(rays, dtype=np.float32)
    out = np.zeros(len(r), HIT_DTYPE)
    lib().lsgo_intersect_linear_srlnp(_p(v), len(v), _p(i), len(i), _p(r), len(r), _p(out), threads)
    return out


def intersect_mt(verts, idx, rays, fp64=False, threads=0):
    v, i = _mesh(verts, idx)
    r = np.ascontiguousarray(rays, dtype=np.float32)
    out = np.zeros(len(r), HIT_DTYPE)
    lib().lsgo_intersect_mt(_p(v), len(v), _p(i), len(i), _p(r), len(r), _p(out), 1 if fp64 else 0, threads)
    return out


def lbvh_build(verts, idx, leaf_size=4):
    v, i = _mesh(verts, idx)
    nt = len(i)
    nodes = np.zeros((max(nt - 1, 1), 16), np.float32)
    prim = np.zeros(max(nt, 1), np.uint32)
    root = ctypes.c_int32(0)
    depth = ctypes.c_int32(0)
    n = lib().lsgo_lbvh_build(_p(v), len(v), _p(i), nt, leaf_size, _p(nodes), _p(prim),
                              ctypes.byref(root), ctypes.byref(depth))
    return dict(nodes=nodes[:max(n - 1, 0)], prim=prim[:n], root=int(root.value), depth=int(depth.value), nvalid=int(n))


def bvh_traverse(tree, verts, idx, rays, threads=0):
    v, i = _mesh(verts, idx)
    r = np.ascontiguousarray(rays, dtype=np.float32)
    nodes = np.ascontiguousarray(tree["nodes"], dtype=np.float32)
    prim = np.ascontiguousarray(tree["prim"], dtype=np.uint32)
    out = np.zeros(len(r), HIT_DTYPE)
    st = Stats()
    lib().lsgo_bvh_traverse(_p(nodes), len(nodes), tree["root"], _p(prim), len(prim), _p(v), len(v), _p(i), len(i),
                            _p(r), len(r), _p(out), ctypes.byref(st), threads)
    return out, st


def filler_triangle_point(n):
    a = np.zeros(1, np.float32)
    b = np.zeros(1, np.float32)
    lib().lsgo_filler_triangle_point(n, _p(a), _p(b))
    return float(a[0]), float(b[0])


# ---- compiled reference collider --------------------------------------------------------

def have_reference() -> bool:
    return os.path.exists(REF_BIN)


def reference_intersect(verts, idx, rays, technique=3, threads=1, repeats=1, flags=None):
    """Runs oracle/_ref/ref_collider; returns (hits, timing dict)."""
    v, i = _mesh(verts, idx)
    r = np.ascontiguousarray(rays, dtype=np.float32)
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.array([0x3152534C, len(v), len(i), len(r)], np.uint32).tobytes())
            f.write(v.tobytes())
            f.write(i.tobytes())
            f.write(r.tobytes())
        args = [REF_BIN, str(technique), fin, fout, str(threads), str(repeats)]
        if flags is not None:
            args.append(str(flags))
        env = dict(os.environ)
        env["OMP_NUM_THREADS"] = str(threads)
        p = subprocess.run(args, check=True, capture_output=True, text=True, env=env)
        timing = json.loads(p.stdout.strip().splitlines()[-1])
        hits = np.fromfile(fout, dtype=HIT_DTYPE)
    return hits, timing


# ---- gather --------------------------------------------------------------------------------

def gather(scene_arrays, params, texels, subs, tree=None, threads=0):
    """scene_arrays: dict verts, idx, normals, side_bits, exitance[nt,8], lights (LIGHT_DTYPE array)."""
    v, i = _mesh(scene_arrays["verts"], scene_arrays["idx"])
    vn = np.ascontiguousarray(scene_arrays["normals"], dtype=np.float32)
    sb = np.ascontiguousarray(scene_arrays["side_bits"], dtype=np.uint8)
    ex = np.ascontiguousarray(scene_arrays["exitance"], dtype=np.float32)
    lights = np.ascontiguousarray(scene_arrays["lights"], dtype=LIGHT_DTYPE)
    tx = np.ascontiguousarray(texels, dtype=np.uint32)
    sbs = np.ascontiguousarray(subs).view(np.float32)
    nout = int(tx[:, 0].max()) + 1 if len(tx) else 0
    out = np.zeros((max(nout, 1), 4), np.float64)
    gp = GParams(params["quality"], params["hemisphere"], params["ray_length_min"], params["locality"],
                 params["rug_distance"], params["light_multiplier"], params["seed"], params.get("light_rounds", 0),
                 params.get("environment_multiplier", 0.0))
    gs = GStats()
    if tree is not None:
        nodes = np.ascontiguousarray(tree["nodes"], dtype=np.float32)
        prim = np.ascontiguousarray(tree["prim"], dtype=np.uint32)
        np_, pp_, root = _p(nodes), _p(prim), tree["root"]
        nprim = len(prim)
    else:
        np_, pp_, root, nprim = None, None, 0, 0
    if scene_arrays.get("pm") is not None:
        _gather_pm(lib(), scene_arrays, v, i, vn, sb, ex, lights, gp, tx, sbs, tree, out, None, gs, threads)
        return out, gs
    ev, keep = _env(scene_arrays)
    ind = _ind(scene_arrays)
    L = lib()
    P, U = ctypes.c_void_p, ctypes.c_uint32
    L.lsgo_gather_env.argtypes = [P, U, P, U, P, P, P, P, P, U, P, P, U, P, P, ctypes.c_int32, P, U, P, P, P, ctypes.c_int]
    L.lsgo_gather_env.restype = None
    L.lsgo_gather_env(_p(v), len(v), _p(i), len(i), _p(vn), _p(sb), _p(ex), _p(ind) if ind is not None else None,
                      _p(lights), len(lights), ctypes.byref(gp), _p(tx), len(tx), _p(sbs), np_, root, pp_, nprim,
                      ctypes.byref(ev) if ev is not None else None, _p(out), ctypes.byref(gs), threads)
    return out, gs


class _PTex(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("format", ctypes.c_int32), ("scaled", ctypes.c_int32),
                ("data", ctypes.c_void_p)]


class _PM(ctypes.Structure):
    _fields_ = [("tri_mat", ctypes.c_void_p), ("tri_uv", ctypes.c_void_p), ("mats", ctypes.c_void_p), ("tex", ctypes.c_void_p),
                ("gamma", ctypes.c_float)]


def _pm(scene_arrays):
    """Point materials of the scene as lsgo_pm (None = none), with the arrays it points to."""
    pm = scene_arrays.get("pm")
    if pm is None:
        return None, None
    tm = np.ascontiguousarray(pm["tri_mat"], np.uint32)
    uv = np.ascontiguousarray(pm["tri_uv"], np.float32)
    mats = np.ascontiguousarray(pm["mats"])
    texs = pm["textures"]
    arr = (_PTex * max(len(texs), 1))()
    datas = []
    for k, t in enumerate(texs):
        d = np.ascontiguousarray(t["data"])
        datas.append(d)
        arr[k] = _PTex(t["width"], t["height"], t["format"], int(t["scaled"]), d.ctypes.data)
    c = _PM(tm.ctypes.data, uv.ctypes.data, mats.ctypes.data, ctypes.cast(arr, ctypes.c_void_p), pm["gamma"])
    return c, (tm, uv, mats, arr, datas)


def texture_at(tex, u, v, gamma=0.0, interpolated=False):
    """lsgo_texture_at: RRBufferInMemory::getElementAtPosition of a texture() dict."""
    L = lib()
    L.lsgo_texture_at.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_void_p]
    L.lsgo_texture_at.restype = None
    d = np.ascontiguousarray(tex["data"])
    t = _PTex(tex["width"], tex["height"], tex["format"], int(tex["scaled"]), d.ctypes.data)
    out = np.zeros(4, np.float32)
    L.lsgo_texture_at(ctypes.byref(t), u, v, gamma, int(interpolated), out.ctypes.data)
    return out


def _gather_pm(L, scene_arrays, v, i, vn, sb, ex, lights, gp, tx, sbs, tree, out, layer_ptrs, gs, threads):
    P, U = ctypes.c_void_p, ctypes.c_uint32
    L.lsgo_gather_pm.argtypes = [P, U, P, U, P, P, P, P, P, P, P, U, P, P, U, P, P, ctypes.c_int32, P, U, P, P, P, P, P,
                                 ctypes.c_int]
    L.lsgo_gather_pm.restype = None
    if tree is not None:
        nodes = np.ascontiguousarray(tree["nodes"], dtype=np.float32)
        prim = np.ascontiguousarray(tree["prim"], dtype=np.uint32)
        np_, pp_, root, nprim = _p(nodes), _p(prim), tree["root"], len(prim)
    else:
        np_, pp_, root, nprim = None, None, 0, 0
    vt, vb = scene_arrays.get("tangents"), scene_arrays.get("bitangents")
    vt = None if vt is None else np.ascontiguousarray(vt, dtype=np.float32)
    vb = None if vb is None else np.ascontiguousarray(vb, dtype=np.float32)
    ev, keep = _env(scene_arrays)
    ind = _ind(scene_arrays)
    pm, keep_pm = _pm(scene_arrays)
    L.lsgo_gather_pm(_p(v), len(v), _p(i), len(i), _p(vn), _p(vt) if vt is not None else None,
                     _p(vb) if vb is not None else None, _p(sb), _p(ex), _p(ind) if ind is not None else None,
                     _p(lights), len(lights), ctypes.byref(gp), _p(tx), len(tx), _p(sbs), np_, root, pp_, nprim,
                     ctypes.byref(ev) if ev is not None else None, ctypes.byref(pm) if pm is not None else None,
                     _p(out) if out is not None else None, layer_ptrs, ctypes.byref(gs), threads)


def _ind(scene_arrays):
    """Per-corner indirect records [nt, 12] (corner k rgb at 4k..4k+2, albedo channel k at
    4k+3), or None when the scene has none."""
    ind = scene_arrays.get("ind")
    return None if ind is None else np.ascontiguousarray(ind, dtype=np.float32).reshape(-1, 12)


def rasterize(tri_uv, tri_ids, width, height, object_id=0, out_base=0):
    """The oracle's texels (bake.TEXEL_DTYPE) and pieces (bake.SUBTEXEL_DTYPE) of one lightmap."""
    from lightsprintsdk_amd import bake as B
    L = lib()
    P, U = ctypes.c_void_p, ctypes.c_uint32
    L.lsgo_rasterize.argtypes = [P, P, U, U, U, U, U, P, U, P, U, ctypes.POINTER(U), ctypes.POINTER(U)]
    L.lsgo_rasterize.restype = ctypes.c_int
    uv = np.ascontiguousarray(tri_uv, np.float32).reshape(-1, 6)
    ids = np.ascontiguousarray(tri_ids, np.uint32)
    nt, ns = U(0), U(0)
    L.lsgo_rasterize(_p(uv), _p(ids), len(ids), width, height, object_id, out_base, None, 0, None, 0,
                     ctypes.byref(nt), ctypes.byref(ns))
    tex = np.zeros(nt.value, B.TEXEL_DTYPE)
    sub = np.zeros(ns.value, B.SUBTEXEL_DTYPE)
    r = L.lsgo_rasterize(_p(uv), _p(ids), len(ids), width, height, object_id, out_base, _p(tex), len(tex), _p(sub), len(sub),
                         ctypes.byref(nt), ctypes.byref(ns))
    assert r == 0
    return tex, sub


def vertex_corners(verts, idx, corner_group, group_pos, library=None):
    """(key, power) of every triangle corner: the group it counts for (0xFFFFFFFF = none)
    and its power = corner angle x triangle area.  `library`: a planted mutation."""
    L = lib() if library is None else library
    if library is not None:
        L.lsgo_vertex_corners.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_uint32] + [ctypes.c_void_p] * 4
    v, i = _mesh(verts, idx)
    g = np.ascontiguousarray(corner_group, np.uint32).reshape(-1)
    gp = np.ascontiguousarray(group_pos, np.float32).reshape(-1, 3)
    key = np.zeros(3 * len(i), np.uint32)
    power = np.zeros(3 * len(i), np.float32)
    L.lsgo_vertex_corners(_p(v), _p(i), len(i), _p(g), _p(gp), _p(key), _p(power))
    return key, power


def smooth_vertices(key, power, tri_data, ngroups, library=None):
    """[ngroups, 4] smoothed values of per-triangle data [ntri, 4]."""
    L = lib() if library is None else library
    if library is not None:
        L.lsgo_smooth_vertices.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                           ctypes.c_uint32, ctypes.c_void_p]
    k = np.ascontiguousarray(key, np.uint32)
    pw = np.ascontiguousarray(power, np.float32)
    d = np.ascontiguousarray(tri_data, np.float32).reshape(-1, 4)
    out = np.zeros((max(ngroups, 1), 4), np.float32)
    L.lsgo_smooth_vertices(_p(k), _p(pw), len(d), _p(d), ngroups, _p(out))
    return out[:ngroups]


STITCH_BIN = os.path.join(ORACLE_DIR, "_ref", "stitch_ref")


def have_stitch_ref() -> bool:
    return os.path.exists(STITCH_BIN)


def reference_stitch(verts, idx, corner_normals, weld, angle):
    """The compiled reference's createOptimizedVertices (oracle/_ref/stitch_ref): (stitched
    vertex of every corner [ntri*3], stitched vertex positions [n, 3])."""
    v, i = _mesh(verts, idx)
    n = np.ascontiguousarray(corner_normals, np.float32).reshape(-1, 9)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.array([0x48435453, len(v), len(i)], np.uint32).tobytes())
            f.write(np.array([weld, angle], np.float32).tobytes())
            f.write(v.tobytes()); f.write(i.tobytes()); f.write(n.tobytes())
        subprocess.run([STITCH_BIN, fin, fout], check=True)
        raw = np.fromfile(fout, np.uint8)
    ns = int(raw[:4].view(np.uint32)[0])
    corners = raw[4:4 + 12 * len(i)].view(np.uint32).copy()
    pos = raw[4 + 12 * len(i):].view(np.float32).reshape(ns, 3).copy()
    return corners, pos


def subtexel_walk(areas, series):
    """The subtexel of every shooter of the given series (list of shooter counts)."""
    a = np.ascontiguousarray(areas, dtype=np.float32)
    n = np.ascontiguousarray(series, dtype=np.uint32)
    out = np.zeros(int(n.sum()), np.uint32)
    lib().lsgo_subtexel_walk(_p(a), len(a), _p(n), len(n), _p(out))
    return out


def gather_layers(scene_arrays, params, texels, subs, layers=(0, 1, 2, 3, 4), tree=None, threads=0):
    """processTexel with the output layers of updateLightmap (lightmap, directions 1-3,
    bent normals).  scene_arrays as gather(), plus optional per-vertex tangents and
    bitangents (None = buildBasisFromNormal of the normals).  Returns {layer: [nout,4]}."""
    L = lib()
    if not getattr(L, "_layers_declared", False):
        P, U = ctypes.c_void_p, ctypes.c_uint32
        L.lsgo_gather_layers.argtypes = [P, U, P, U, P, P, P, P, P, P, P, U, P, P, U, P, P, ctypes.c_int32, P, U, P, P, P,
                                         ctypes.c_int]
        L.lsgo_gather_layers.restype = None
        L.lsgo_default_frames.argtypes = [P, U, P, P]
        L.lsgo_default_frames.restype = None
        L._layers_declared = True
    v, i = _mesh(scene_arrays["verts"], scene_arrays["idx"])
    vn = np.ascontiguousarray(scene_arrays["normals"], dtype=np.float32)
    vt, vb = scene_arrays.get("tangents"), scene_arrays.get("bitangents")
    vt = None if vt is None else np.ascontiguousarray(vt, dtype=np.float32)
    vb = None if vb is None else np.ascontiguousarray(vb, dtype=np.float32)
    sb = np.ascontiguousarray(scene_arrays["side_bits"], dtype=np.uint8)
    ex = np.ascontiguousarray(scene_arrays["exitance"], dtype=np.float32)
    lights = np.ascontiguousarray(scene_arrays["lights"], dtype=LIGHT_DTYPE)
    tx = np.ascontiguousarray(texels, dtype=np.uint32)
    sbs = np.ascontiguousarray(subs).view(np.float32)
    nout = int(tx[:, 0].max()) + 1 if len(tx) else 0
    outs = {k: np.zeros((max(nout, 1), 4), np.float64) for k in layers}
    ptrs = (ctypes.c_void_p * 5)(*[(_p(outs[k]) if k in outs else None) for k in range(5)])
    gp = GParams(params["quality"], params["hemisphere"], params["ray_length_min"], params["locality"],
                 params["rug_distance"], params["light_multiplier"], params["seed"], params.get("light_rounds", 0),
                 params.get("environment_multiplier", 0.0))
    gs = GStats()
    if scene_arrays.get("pm") is not None:
        _gather_pm(L, scene_arrays, v, i, vn, sb, ex, lights, gp, tx, sbs, tree, None, ptrs, gs, threads)
        return outs
    if tree is not None:
        nodes = np.ascontiguousarray(tree["nodes"], dtype=np.float32)
        prim = np.ascontiguousarray(tree["prim"], dtype=np.uint32)
        np_, pp_, root, nprim = _p(nodes), _p(prim), tree["root"], len(prim)
    else:
        np_, pp_, root, nprim = None, None, 0, 0
    ev, keep = _env(scene_arrays)
    ind = _ind(scene_arrays)
    L.lsgo_gather_layers(_p(v), len(v), _p(i), len(i), _p(vn), _p(vt) if vt is not None else None,
                         _p(vb) if vb is not None else None, _p(sb), _p(ex), _p(ind) if ind is not None else None,
                         _p(lights), len(lights), ctypes.byref(gp),
                         _p(tx), len(tx), _p(sbs), np_, root, pp_, nprim, ctypes.byref(ev) if ev is not None else None,
                         ptrs, ctypes.byref(gs), threads)
    return outs


def default_frames(normals):
    """Per-vertex (tangent, bitangent) by TangentBasis::buildBasisFromNormal."""
    L = lib()
    P, U = ctypes.c_void_p, ctypes.c_uint32
    L.lsgo_default_frames.argtypes = [P, U, P, P]
    L.lsgo_default_frames.restype = None
    n = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
    t = np.zeros_like(n)
    b = np.zeros_like(n)
    L.lsgo_default_frames(_p(n), len(n), _p(t), _p(b))
    return t, b


DROP_IN_BIN = os.path.join(ORACLE_DIR, "_ref", "drop_in_test")


def have_drop_in() -> bool:
    return os.path.exists(DROP_IN_BIN)


def drop_in(verts, idx, rays, technique):
    """Runs oracle/_ref/drop_in_test: RRCollider::create(mesh, nullptr, technique)
    through the compiled reference with the IT_HIP_LBVH adapter registered.
    Returns (single-ray hits, batch hits, front-only-handler hits, info dict)."""
    v, i = _mesh(verts, idx)
    r = np.ascontiguousarray(rays, dtype=np.float32)
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.array([0x3152534C, len(v), len(i), len(r)], np.uint32).tobytes())
            f.write(v.tobytes()); f.write(i.tobytes()); f.write(r.tobytes())
        p = subprocess.run([DROP_IN_BIN, str(technique), fin, fout], check=True, capture_output=True, text=True)
        info = json.loads(p.stdout.strip().splitlines()[-1])
        out = np.fromfile(fout, dtype=HIT_DTYPE).reshape(3, len(r))
    return out[0], out[1], out[2], info


def filter_lightmap(rgba, sigma=1.0, spread=1000.0, bg=(0, 0, 0, 0), library=None):
    """RRBuffer filter chain on one RGBA lightmap [h,w,4]; returns (rgba_out, rgbf).
    `library` is another ctypes library exporting lsgo_filter_lightmap (a planted
    mutation of the oracle, in the mutation tests); default: the oracle."""
    L = lib() if library is None else library
    L.lsgo_filter_lightmap.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_void_p, ctypes.c_void_p]
    img = np.ascontiguousarray(rgba, dtype=np.float32).copy()
    h, w = img.shape[:2]
    b = np.asarray(bg, np.float32)
    out = np.zeros((h, w, 3), np.float32)
    L.lsgo_filter_lightmap(_p(img), w, h, sigma, spread, _p(b), _p(out))
    return img, out


def filter_lightmap_ex(rgba, sigma=1.0, spread=1000.0, bg=(0, 0, 0, 0), wrap=False, seams=None, library=None):
    """The filter chain with FilteringParameters::wrap and the unwrap seams of the map's
    object (float32 [n, 8], or None) smoothed after every blur pass."""
    L = lib() if library is None else library
    L.lsgo_filter_lightmap_ex.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32]
    img = np.ascontiguousarray(rgba, dtype=np.float32).copy()
    h, w = img.shape[:2]
    b = np.asarray(bg, np.float32)
    out = np.zeros((h, w, 3), np.float32)
    sm = None if seams is None else np.ascontiguousarray(seams, np.float32).reshape(-1, 8)
    L.lsgo_filter_lightmap_ex(_p(img), w, h, sigma, spread, _p(b), _p(out), 1 if wrap else 0,
                              None if sm is None else _p(sm), 0 if sm is None else len(sm))
    return img, out


ROTATE_BIN = os.path.join(ORACLE_DIR, "_ref", "rotate_ref")


def env_rotate(yaws, dirs, library=None):
    """The oracle's yaw rotation of directions (lsgo_env_rotate): float32 [n, 3]."""
    L = lib() if library is None else library
    L.lsgo_env_rotate.argtypes = [ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]
    L.lsgo_env_rotate.restype = None
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    out = np.zeros_like(d)
    r = np.zeros(3, np.float32)
    for k in range(len(d)):
        dk = np.ascontiguousarray(d[k])
        L.lsgo_env_rotate(float(yaws[k]), _p(dk), _p(r))
        out[k] = r
    return out


def reference_rotate(yaws, dirs):
    """The compiled reference's rotationByYawPitchRoll((yaw,0,0)).transformDirection
    (oracle/_ref/rotate_ref) of each direction: float32 [n, 3]."""
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    rec = np.concatenate([np.asarray(yaws, np.float32).reshape(-1, 1), d], axis=1).astype(np.float32)
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.uint32([len(d)]).tobytes())
            f.write(rec.tobytes())
        subprocess.run([ROTATE_BIN, fin, fout], check=True, capture_output=True)
        return np.fromfile(fout, np.float32).reshape(-1, 3)


def unwrap_seams_from_corners(corners, nverts, tri_uv, library=None):
    """UnwrapSeams's search over a stitched mesh (the stitched vertex of every corner,
    [ntri*3]): float32 [n, 8]."""
    L = lib() if library is None else library
    L.lsgo_unwrap_seams.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_uint32]
    L.lsgo_unwrap_seams.restype = ctypes.c_uint32
    c = np.ascontiguousarray(corners, np.uint32).reshape(-1)
    uv = np.ascontiguousarray(tri_uv, np.float32).reshape(-1, 6)
    nt = len(uv)
    n = L.lsgo_unwrap_seams(_p(c), int(nverts), nt, _p(uv), None, 0)
    out = np.zeros((n, 8), np.float32)
    if n:
        L.lsgo_unwrap_seams(_p(c), int(nverts), nt, _p(uv), _p(out), n)
    return out


def mlp_num_params():
    L = lib()
    L.lsgo_mlp_num_params.restype = ctypes.c_uint32
    return int(L.lsgo_mlp_num_params())


def mlp_loss_grad(params, x, t, lam=0.5, bf16=True, want_grad=True):
    """(loss, y[n,4], grad[nparams] float64) of the surrogate MLP in binary64."""
    L = lib()
    L.lsgo_mlp_loss_grad.restype = ctypes.c_double
    L.lsgo_mlp_loss_grad.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    p = np.ascontiguousarray(params, np.float32)
    xx = np.ascontiguousarray(x, np.float32)
    tt = np.ascontiguousarray(t, np.float32)
    y = np.zeros((len(xx), 4), np.float32)
    g = np.zeros(len(p), np.float64)
    loss = L.lsgo_mlp_loss_grad(_p(p), _p(xx), _p(tt), len(xx), lam, 1 if bf16 else 0, _p(y), _p(g) if want_grad else None)
    return loss, y, g
