"""Seeded procedural inputs for the ray-cast and lightmap-bake workloads.

Every generator is deterministic for a given seed (numpy PCG64, float64
arithmetic, cast to float32 at the end), so the GPU path, the CPU oracle and the
compiled reference collider all see bit-identical meshes and rays.

Workloads (sizes as used by bench.py and the tests):
  * displaced icosphere: icosahedron subdivided at frequency n (20*n^2 triangles;
    n=71 gives 100,820), each shared vertex displaced radially by U(-0.1,0.1)*r.
  * random rays: origins uniform in the 1.2x-scaled mesh box, directions
    uniform on the unit sphere (normalised Gaussian triples), tmin=0, tmax=inf.
  * box room: closed 10x10x4 m room + 20 axis-aligned boxes, faces tessellated
    into grids, one point light and one emissive ceiling quad, per-object
    albedo U(0.2,0.8), per-object lightmap unwrap (one chart per face, shelf
    packed with a gutter).
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

ICOSPHERE_100K_FREQUENCY = 71   # 20 * 71^2 = 100,820 triangles


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v, f


def icosphere(frequency: int, seed: int = 1, displacement: float = 0.1, radius: float = 1.0):
    """Displaced geodesic sphere with 20*frequency^2 triangles.

    Returns (verts float32 [nv,3], indices uint32 [nt,3]); triangles face outward."""
    n = int(frequency)
    if n < 1:
        raise ValueError("frequency must be >= 1")
    base_v, base_f = _icosahedron()
    # barycentric grid per face: point (i,j) with i+j<=n
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    keep = ii + jj <= n
    ii, jj = ii[keep], jj[keep]
    local = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(ii, jj))}
    tris_local = []
    for i in range(n):
        for j in range(n - i):
            tris_local.append((local[(i, j)], local[(i + 1, j)], local[(i, j + 1)]))
            if i + j < n - 1:
                tris_local.append((local[(i + 1, j)], local[(i + 1, j + 1)], local[(i, j + 1)]))
    tris_local = np.array(tris_local, dtype=np.int64)
    w0 = (n - ii - jj) / n
    w1 = ii / n
    w2 = jj / n
    pts = []
    faces = []
    npf = len(ii)
    for fi, (a, b, c) in enumerate(base_f):
        p = w0[:, None] * base_v[a] + w1[:, None] * base_v[b] + w2[:, None] * base_v[c]
        p /= np.linalg.norm(p, axis=1, keepdims=True)
        pts.append(p)
        faces.append(tris_local + fi * npf)
    pts = np.concatenate(pts)
    faces = np.concatenate(faces)
    key = np.round(pts * 1e9).astype(np.int64)
    _, uniq_idx, inverse = np.unique(key, axis=0, return_index=True, return_inverse=True)
    verts = pts[uniq_idx]
    inverse = inverse.reshape(-1)
    faces = inverse[faces]
    rng = np.random.default_rng(seed)
    scale = 1.0 + rng.uniform(-displacement, displacement, size=len(verts))
    verts = verts * (scale * radius)[:, None]
    # orient outward
    v0, v1, v2 = verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]
    nrm = np.cross(v1 - v0, v2 - v0)
    flip = np.einsum("ij,ij->i", nrm, v0 + v1 + v2) < 0
    faces[flip] = faces[flip][:, [0, 2, 1]]
    return verts.astype(np.float32), faces.astype(np.uint32)


def random_rays(verts: np.ndarray, n: int, seed: int = 2, box_scale: float = 1.2,
                tmin: float = 0.0, tmax: float = np.inf):
    """Rays as float32 [n,8] = (ox,oy,oz,tmin, dx,dy,dz,tmax)."""
    rng = np.random.default_rng(seed)
    lo = verts.min(axis=0).astype(np.float64)
    hi = verts.max(axis=0).astype(np.float64)
    c = (lo + hi) / 2
    h = (hi - lo) / 2 * box_scale
    o = c + rng.uniform(-1.0, 1.0, size=(n, 3)) * h
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.empty((n, 8), dtype=np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def mesh_digest(verts: np.ndarray, idx: np.ndarray) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(verts, dtype=np.float32).tobytes())
    h.update(np.ascontiguousarray(idx, dtype=np.uint32).tobytes())
    return h.hexdigest()


# ---------------------------------------------------------------------------------------
# lightmap scenes

@dataclass
class LightmapObject:
    name: str
    tri_begin: int
    tri_count: int
    width: int
    height: int
    albedo: tuple


@dataclass
class BakeScene:
    verts: np.ndarray            # float32 [nv,3]
    idx: np.ndarray              # uint32 [nt,3]
    normals: np.ndarray          # float32 [nv,3] per-vertex normals
    tri_uv: np.ndarray           # float32 [nt,6] lightmap uv per triangle corner
    side_bits: np.ndarray        # uint8 [nt]
    albedo: np.ndarray           # float32 [nt,3]
    emittance: np.ndarray        # float32 [nt,3] diffuse emittance (front side)
    objects: list = field(default_factory=list)   # LightmapObject for lightmapped objects
    lights: list = field(default_factory=list)     # dicts: type, position, direction, color, cast_shadows

    @property
    def ntri(self):
        return int(self.idx.shape[0])


def _grid_face(origin, du, dv, nu, nv_):
    """Quad origin + s*du + t*dv tessellated into nu x nv cells (2 tris each).
    Returns local verts [(nu+1)(nv+1),3], tris [2*nu*nv,3], (s,t) per vertex."""
    s = np.linspace(0.0, 1.0, nu + 1)
    t = np.linspace(0.0, 1.0, nv_ + 1)
    S, T = np.meshgrid(s, t, indexing="ij")
    P = origin[None, None, :] + S[..., None] * du[None, None, :] + T[..., None] * dv[None, None, :]
    verts = P.reshape(-1, 3)
    st = np.stack([S.reshape(-1), T.reshape(-1)], axis=1)
    return verts, _grid_tris(nu, nv_), st


def _grid_tris(nu, nv_):
    """Two triangles per cell of an (nu+1) x (nv+1) vertex grid (vertex i*(nv+1)+j),
    cells in i-major order: (a,b,c), (a,c,d)."""
    i, j = np.meshgrid(np.arange(nu), np.arange(nv_), indexing="ij")
    a = (i * (nv_ + 1) + j).reshape(-1)
    b = a + (nv_ + 1)
    c = b + 1
    d = a + 1
    return np.stack([np.stack([a, b, c], 1), np.stack([a, c, d], 1)], 1).reshape(-1, 3).astype(np.int64)


def _box_faces(lo, hi, inward):
    """Six faces of an axis-aligned box as (origin, du, dv) with du x dv pointing
    outward (or inward)."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    ex = hi - lo
    X, Y, Z = np.eye(3)
    faces = [
        (np.array([lo[0], lo[1], lo[2]]), Y * ex[1], X * ex[0]),  # z- (normal -z)
        (np.array([lo[0], lo[1], hi[2]]), X * ex[0], Y * ex[1]),  # z+
        (np.array([lo[0], lo[1], lo[2]]), X * ex[0], Z * ex[2]),  # y- (normal -y)
        (np.array([lo[0], hi[1], lo[2]]), Z * ex[2], X * ex[0]),  # y+
        (np.array([lo[0], lo[1], lo[2]]), Z * ex[2], Y * ex[1]),  # x- (normal -x)
        (np.array([hi[0], lo[1], lo[2]]), Y * ex[1], Z * ex[2]),  # x+
    ]
    if inward:
        faces = [(o, dv, du) for (o, du, dv) in faces]
    return faces


def _pack_charts(sizes, gutter):
    """Shelf-pack rectangles (w,h) (in relative units) into the unit square.
    Returns list of (x,y,scale) so chart k occupies [x, x+w*scale] x [y, y+h*scale]."""
    total = sum(w * h for w, h in sizes)
    scale = 1.0 / np.sqrt(total) * 0.75
    for _ in range(60):
        x = y = gutter
        row_h = 0.0
        placed = []
        ok = True
        order = sorted(range(len(sizes)), key=lambda k: -sizes[k][1])
        pos = [None] * len(sizes)
        for k in order:
            w, h = sizes[k][0] * scale, sizes[k][1] * scale
            if x + w + gutter > 1.0:
                x = gutter
                y += row_h + gutter
                row_h = 0.0
            if x + w + gutter > 1.0 or y + h + gutter > 1.0:
                ok = False
                break
            pos[k] = (x, y, scale)
            x += w + gutter
            row_h = max(row_h, h)
        if ok:
            return pos
        scale *= 0.95
    raise RuntimeError("chart packing failed")


def box_room(seed: int = 3, n_boxes: int = 20, target_tris: int = 50_000, lightmap_size: int = 256,
             room=(10.0, 10.0, 4.0), box_size=(0.2, 2.0), light_power: float = 40.0,
             emitter_radiance: float = 8.0) -> BakeScene:
    """Closed room + axis-aligned boxes, faces tessellated to about target_tris."""
    rng = np.random.default_rng(seed)
    RX, RY, RZ = room
    objs = [("room", (0.0, 0.0, 0.0), (RX, RY, RZ), True)]
    for b in range(n_boxes):
        sz = rng.uniform(box_size[0], box_size[1], size=3)
        sz[2] = min(sz[2], RZ * 0.8)
        px = rng.uniform(0.2, RX - sz[0] - 0.2)
        py = rng.uniform(0.2, RY - sz[1] - 0.2)
        objs.append((f"box{b}", (px, py, 0.0), (px + sz[0], py + sz[1], sz[2]), False))
    # tessellation density: cells per metre so that total triangles ~ target
    area = 0.0
    for _, lo, hi, _inw in objs:
        e = np.subtract(hi, lo)
        area += 2 * (e[0] * e[1] + e[1] * e[2] + e[0] * e[2])
    cells_per_m2 = target_tris / 2.0 / area
    density = np.sqrt(cells_per_m2)
    V, I, UV, SB, ALB, EM, N = [], [], [], [], [], [], []
    objects = []
    nv = 0
    for name, lo, hi, inward in objs:
        faces = _box_faces(lo, hi, inward)
        albedo = tuple(float(a) for a in rng.uniform(0.2, 0.8, size=3))
        sizes = [(np.linalg.norm(du), np.linalg.norm(dv)) for _, du, dv in faces]
        gutter = 3.0 / lightmap_size
        pos = _pack_charts(sizes, gutter)
        tri_begin = sum(len(x) for x in I)
        for (o, du, dv), (cx, cy, sc), (w, h) in zip(faces, pos, sizes):
            nu = max(1, int(round(w * density)))
            nvv = max(1, int(round(h * density)))
            fv, ft, st = _grid_face(o, du, dv, nu, nvv)
            nrm = np.cross(du, dv)
            nrm /= np.linalg.norm(nrm)
            V.append(fv)
            N.append(np.repeat(nrm[None, :], len(fv), axis=0))
            I.append(ft + nv)
            uv_v = np.stack([cx + st[:, 0] * w * sc, cy + st[:, 1] * h * sc], axis=1)
            UV.append(uv_v[ft].reshape(-1, 6))
            nv += len(fv)
        count = sum(len(x) for x in I) - tri_begin
        SB.append(np.full(count, 1 | 4, np.uint8))
        ALB.append(np.repeat(np.array(albedo, np.float32)[None, :], count, axis=0))
        EM.append(np.zeros((count, 3), np.float32))
        objects.append(LightmapObject(name, tri_begin, count, lightmap_size, lightmap_size, albedo))
    # emissive ceiling quad (not lightmapped), facing down, 1 cm below the ceiling
    qx, qy = RX * 0.5, RY * 0.5
    o = np.array([qx - 0.6, qy - 0.6, RZ - 0.01])
    fv, ft, st = _grid_face(o, np.array([0.0, 1.2, 0.0]), np.array([1.2, 0.0, 0.0]), 2, 2)
    V.append(fv)
    N.append(np.repeat(np.array([[0.0, 0.0, -1.0]]), len(fv), axis=0))
    I.append(ft + nv)
    UV.append(np.zeros((len(ft), 6)))
    SB.append(np.full(len(ft), 1 | 4, np.uint8))
    ALB.append(np.full((len(ft), 3), 0.5, np.float32))
    EM.append(np.full((len(ft), 3), emitter_radiance, np.float32))
    nv += len(fv)
    lights = [dict(type=0, position=(RX * 0.3, RY * 0.6, RZ * 0.75), direction=(0.0, 0.0, -1.0),
                   color=(light_power, light_power * 0.9, light_power * 0.8), cast_shadows=1)]
    return BakeScene(
        verts=np.concatenate(V).astype(np.float32),
        idx=np.concatenate(I).astype(np.uint32),
        normals=np.concatenate(N).astype(np.float32),
        tri_uv=np.concatenate(UV).astype(np.float32),
        side_bits=np.concatenate(SB),
        albedo=np.concatenate(ALB).astype(np.float32),
        emittance=np.concatenate(EM).astype(np.float32),
        objects=objects,
        lights=lights,
    )


def minimal_safe_distance(verts: np.ndarray) -> float:
    """RRSolver.cpp:333: (maxi-mini).avg()*1e-6 in float32, the box as RRMesh::getAABB
    finds it (RRMesh.cpp:164-196): per axis over the finite coordinates only, an axis
    without any (or no vertices at all) spanning 0."""
    v = np.asarray(verts, np.float32).reshape(-1, 3)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    for j in range(3):
        c = v[:, j][np.isfinite(v[:, j])]
        if len(c):
            lo[j], hi[j] = c.min(), c.max()
    e = (hi - lo).astype(np.float32)
    avg = np.float32((np.float32(e[0] + e[1]) + e[2]) * np.float32(0.33333333333333))
    return float(np.float32(avg * np.float32(1e-6)))


# ---------------------------------------------------------------------------------------
# batched-bake scene: many objects of uneven lightmap sizes

def _grid_normals(verts, tris):
    n = np.zeros_like(verts)
    fn = np.cross(verts[tris[:, 1]] - verts[tris[:, 0]], verts[tris[:, 2]] - verts[tris[:, 0]])
    for k in range(3):
        np.add.at(n, tris[:, k], fn)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    ln[ln == 0] = 1
    return n / ln


def _chart_object(charts, lightmap_size, gutter_texels=3.0):
    """charts: list of (verts [n,3], tris [m,3], st [n,2] in [0,1]^2 chart coords, (w,h) chart size).
    Packs the charts into one unwrap; returns verts, tris, normals, per-triangle uv [m,6]."""
    sizes = [c[3] for c in charts]
    pos = _pack_charts(sizes, gutter_texels / lightmap_size)
    V, I, N, UV = [], [], [], []
    nv = 0
    for (v, t, st, (w, h)), (cx, cy, sc) in zip(charts, pos):
        uv_v = np.stack([cx + st[:, 0] * w * sc, cy + st[:, 1] * h * sc], axis=1)
        V.append(v)
        N.append(_grid_normals(v, t))
        I.append(t + nv)
        UV.append(uv_v[t].reshape(-1, 6))
        nv += len(v)
    return np.concatenate(V), np.concatenate(I), np.concatenate(N), np.concatenate(UV)


def _displaced_box(rng, lo, hi, cells_per_m, amp):
    charts = []
    for o, du, dv in _box_faces(lo, hi, inward=False):
        w, h = np.linalg.norm(du), np.linalg.norm(dv)
        nu = max(2, int(round(w * cells_per_m)))
        nvv = max(2, int(round(h * cells_per_m)))
        v, t, st = _grid_face(o, du, dv, nu, nvv)
        nrm = np.cross(du, dv)
        nrm /= np.linalg.norm(nrm)
        interior = (st[:, 0] > 0) & (st[:, 0] < 1) & (st[:, 1] > 0) & (st[:, 1] < 1)
        v = v + nrm[None, :] * (rng.uniform(-amp, amp, size=len(v)) * interior)[:, None]
        charts.append((v, t, st, (w, h)))
    return charts


def _cylinder(rng, center, radius, height, cells_per_m):
    seg = max(8, int(round(2 * np.pi * radius * cells_per_m)))
    rows = max(2, int(round(height * cells_per_m)))
    s = np.linspace(0, 1, seg + 1)
    z = np.linspace(0, 1, rows + 1)
    S, Z = np.meshgrid(s, z, indexing="ij")
    ang = S * 2 * np.pi
    v = np.stack([center[0] + radius * np.cos(ang), center[1] + radius * np.sin(ang), center[2] + Z * height], -1).reshape(-1, 3)
    st = np.stack([S.reshape(-1), Z.reshape(-1)], 1)
    t = []
    for i in range(seg):
        for j in range(rows):
            a = i * (rows + 1) + j
            b = (i + 1) * (rows + 1) + j
            t.append((a, b, b + 1))
            t.append((a, b + 1, a + 1))
    charts = [(v, np.array(t), st, (2 * np.pi * radius, height))]
    # caps as fans (top faces up, bottom faces down)
    for top in (True, False):
        ang2 = np.linspace(0, 2 * np.pi, seg, endpoint=False)
        ring = np.stack([center[0] + radius * np.cos(ang2), center[1] + radius * np.sin(ang2),
                         np.full(seg, center[2] + (height if top else 0.0))], 1)
        cv = np.concatenate([np.array([[center[0], center[1], center[2] + (height if top else 0.0)]]), ring])
        cst = np.concatenate([np.array([[0.5, 0.5]]), 0.5 + 0.5 * np.stack([np.cos(ang2), np.sin(ang2)], 1)])
        ct = np.array([(0, 1 + k, 1 + (k + 1) % seg) if top else (0, 1 + (k + 1) % seg, 1 + k) for k in range(seg)])
        charts.append((cv, ct, cst, (2 * radius, 2 * radius)))
    return charts


def _terrain(rng, origin, size, cells, amp):
    v, t, st = _grid_face(np.array(origin, np.float64), np.array([size, 0.0, 0.0]), np.array([0.0, size, 0.0]), cells, cells)
    k = rng.uniform(0.5, 2.0, size=(3, 2)) * 2 * np.pi / size
    ph = rng.uniform(0, 2 * np.pi, size=3)
    hgt = sum(np.sin(k[i, 0] * v[:, 0] + k[i, 1] * v[:, 1] + ph[i]) for i in range(3)) * amp / 3
    v = v.copy()
    v[:, 2] += hgt
    return [(v, t, st, (size, size))]


def object_field(seed: int = 4, n_objects: int = 200, target_tris: int = 2_000_000, lightmap_range=(32, 1024),
                 area: float = 200.0) -> BakeScene:
    """Batched-bake scene: displaced boxes, cylinders and terrain patches from one seed, each with its
    own packed unwrap and a lightmap side log-uniform in lightmap_range; one directional
    sun and four point lights."""
    rng = np.random.default_rng(seed)
    per_obj = target_tris / n_objects
    V, I, N, UV, SB, ALB, EM = [], [], [], [], [], [], []
    objects = []
    nv = 0
    ntri = 0
    grid = int(np.ceil(np.sqrt(n_objects)))
    cell = area / grid
    for k in range(n_objects):
        kind = k % 3
        gx, gy = k % grid, k // grid
        base = np.array([gx * cell, gy * cell, 0.0])
        lm = int(round(2 ** rng.uniform(np.log2(lightmap_range[0]), np.log2(lightmap_range[1]))))
        if kind == 0:
            sz = rng.uniform(1.0, cell * 0.6, size=3)
            area_o = 2 * (sz[0] * sz[1] + sz[1] * sz[2] + sz[0] * sz[2])
            cpm = np.sqrt(per_obj / 2 / area_o)
            charts = _displaced_box(rng, base + [0.5, 0.5, 0.0], base + [0.5, 0.5, 0.0] + sz, cpm, 0.02 * sz.min())
        elif kind == 1:
            r = rng.uniform(0.5, cell * 0.3)
            hgt = rng.uniform(1.0, cell * 0.8)
            area_o = 2 * np.pi * r * hgt + 2 * np.pi * r * r
            cpm = np.sqrt(per_obj / 2 / area_o)
            charts = _cylinder(rng, base + [cell / 2, cell / 2, 0.0], r, hgt, cpm)
        else:
            cells = max(4, int(np.sqrt(per_obj / 2)))
            charts = _terrain(rng, base + [0.0, 0.0, -0.5], cell * 0.95, cells, cell * 0.05)
        v, t, n, uv = _chart_object(charts, lm)
        albedo = rng.uniform(0.2, 0.8, size=3).astype(np.float32)
        V.append(v); I.append(t + nv); N.append(n); UV.append(uv)
        SB.append(np.full(len(t), 1 | 4, np.uint8))
        ALB.append(np.repeat(albedo[None, :], len(t), 0))
        EM.append(np.zeros((len(t), 3), np.float32))
        objects.append(LightmapObject(f"obj{k}", ntri, len(t), lm, lm, tuple(float(a) for a in albedo)))
        nv += len(v)
        ntri += len(t)
    lights = [dict(type=1, position=(0, 0, 0), direction=tuple(np.array([0.3, 0.2, -1.0]) / np.linalg.norm([0.3, 0.2, -1.0])),
                   color=(3.0, 2.9, 2.6), cast_shadows=1)]
    for j in range(4):
        p = (rng.uniform(0, area), rng.uniform(0, area), rng.uniform(5, 15))
        lights.append(dict(type=0, position=p, direction=(0, 0, -1), color=(400.0, 380.0, 350.0), cast_shadows=1))
    return BakeScene(verts=np.concatenate(V).astype(np.float32), idx=np.concatenate(I).astype(np.uint32),
                     normals=np.concatenate(N).astype(np.float32), tri_uv=np.concatenate(UV).astype(np.float32),
                     side_bits=np.concatenate(SB), albedo=np.concatenate(ALB).astype(np.float32),
                     emittance=np.concatenate(EM).astype(np.float32), objects=objects, lights=lights)


# ---------------------------------------------------------------------------------------
# large re-bake scene: city of extruded boxes on a heightfield terrain

def _terrain_height(rng_params, x, y):
    k, ph, amp = rng_params
    return sum(np.sin(k[i, 0] * x + k[i, 1] * y + ph[i]) for i in range(len(ph))) * amp / len(ph)


def city_terrain(seed: int = 5, n_boxes: int = 5000, grid_vertices: int = 2048, terrain_tiles: int = 4,
                 lightmap_size: int = 2048, city_maps: int = 48, box_tris: int = 1_600_000,
                 cell_m: float = 1.0, sun_direction=(0.4, 0.25, -0.88)) -> BakeScene:
    """City and terrain: a grid_vertices^2 heightfield (cell_m metres per cell) split into
    terrain_tiles^2 lightmapped tiles, and n_boxes extruded boxes (5 faces) spread over
    city_maps lightmaps; all lightmaps lightmap_size^2.  One directional sun."""
    rng = np.random.default_rng(seed)
    cells = grid_vertices - 1
    size = cells * cell_m
    tp = (rng.uniform(0.5, 2.0, size=(4, 2)) * 2 * np.pi / size, rng.uniform(0, 2 * np.pi, size=4), size * 0.004)
    V, I, N, UV, ALB = [], [], [], [], []
    objects = []
    nv = 0
    ntri = 0
    # terrain tiles: one planar chart each
    edges = np.linspace(0, cells, terrain_tiles + 1).round().astype(int)
    gutter = 2.0 / lightmap_size
    for ty in range(terrain_tiles):
        for tx in range(terrain_tiles):
            c0, c1 = edges[tx], edges[tx + 1]
            r0, r1 = edges[ty], edges[ty + 1]
            nu, nvv = c1 - c0, r1 - r0
            xs = np.arange(c0, c1 + 1) * cell_m
            ys = np.arange(r0, r1 + 1) * cell_m
            X, Y = np.meshgrid(xs, ys, indexing="ij")
            Z = _terrain_height(tp, X, Y)
            v = np.stack([X, Y, Z], -1).reshape(-1, 3)
            t = _grid_tris(nu, nvv)
            st = np.stack([(X - xs[0]) / (xs[-1] - xs[0]), (Y - ys[0]) / (ys[-1] - ys[0])], -1).reshape(-1, 2)
            uvv = gutter + st * (1 - 2 * gutter)
            albedo = rng.uniform(0.25, 0.45, size=3).astype(np.float32)
            V.append(v); I.append(t + nv); N.append(_grid_normals(v, t)); UV.append(uvv[t].reshape(-1, 6))
            ALB.append(np.repeat(albedo[None, :], len(t), 0))
            objects.append(LightmapObject(f"terrain{ty}_{tx}", ntri, len(t), lightmap_size, lightmap_size,
                                          tuple(float(a) for a in albedo)))
            nv += len(v); ntri += len(t)
    # city: jittered grid of boxes over the central 85% of the terrain
    side = int(np.ceil(np.sqrt(n_boxes)))
    pitch = size * 0.85 / side
    off = size * 0.075
    fp = rng.uniform(0.3, 0.75, size=(n_boxes, 2)) * pitch
    hgt = np.exp(rng.uniform(np.log(6.0), np.log(60.0), size=n_boxes))
    areas = 2 * fp[:, 0] * fp[:, 1] / 2 + 2 * hgt * (fp[:, 0] + fp[:, 1])
    cpm = np.sqrt(box_tris / 2 / areas.sum())
    per_map = int(np.ceil(n_boxes / city_maps))
    for m in range(city_maps):
        charts = []
        for b in range(m * per_map, min(n_boxes, (m + 1) * per_map)):
            gx, gy = b % side, b // side
            cx = off + (gx + 0.5) * pitch + rng.uniform(-0.1, 0.1) * pitch
            cy = off + (gy + 0.5) * pitch + rng.uniform(-0.1, 0.1) * pitch
            lo = np.array([cx - fp[b, 0] / 2, cy - fp[b, 1] / 2, 0.0])
            hi = np.array([cx + fp[b, 0] / 2, cy + fp[b, 1] / 2, 0.0])
            corners = _terrain_height(tp, np.array([lo[0], hi[0], lo[0], hi[0]]), np.array([lo[1], lo[1], hi[1], hi[1]]))
            lo[2] = corners.min() - 0.5
            hi[2] = corners.max() + hgt[b]
            for k, (o, du, dv) in enumerate(_box_faces(lo, hi, inward=False)):
                if k == 0:
                    continue                       # no floor face
                w, h = np.linalg.norm(du), np.linalg.norm(dv)
                v, t, st = _grid_face(o, du, dv, max(1, int(round(w * cpm))), max(1, int(round(h * cpm))))
                charts.append((v, t, st, (w, h)))
        if not charts:
            continue
        v, t, n, uv = _chart_object(charts, lightmap_size, gutter_texels=2.0)
        albedo = rng.uniform(0.3, 0.8, size=3).astype(np.float32)
        V.append(v); I.append(t + nv); N.append(n); UV.append(uv)
        ALB.append(np.repeat(albedo[None, :], len(t), 0))
        objects.append(LightmapObject(f"city{m}", ntri, len(t), lightmap_size, lightmap_size,
                                      tuple(float(a) for a in albedo)))
        nv += len(v); ntri += len(t)
    d = np.asarray(sun_direction, np.float64)
    d /= np.linalg.norm(d)
    lights = [dict(type=1, position=(0, 0, 0), direction=tuple(d), color=(3.0, 2.9, 2.6), cast_shadows=1)]
    nt = sum(len(x) for x in I)
    return BakeScene(verts=np.concatenate(V).astype(np.float32), idx=np.concatenate(I).astype(np.uint32),
                     normals=np.concatenate(N).astype(np.float32), tri_uv=np.concatenate(UV).astype(np.float32),
                     side_bits=np.full(nt, 1 | 4, np.uint8), albedo=np.concatenate(ALB).astype(np.float32),
                     emittance=np.zeros((nt, 3), np.float32), objects=objects, lights=lights)


def sun_sweep(step: int, steps: int = 60):
    """Sun direction of step `step` of a sweep: azimuth once around, elevation 15..75 deg."""
    f = step / max(steps - 1, 1)
    az = 2 * np.pi * f
    el = np.radians(15 + 60 * np.sin(np.pi * f))
    d = np.array([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), -np.sin(el)])
    return tuple(d / np.linalg.norm(d))
