"""Hit-record comparison helpers shared by the CPU and GPU parity tests."""
import numpy as np


def classify_vs_reference(ref, got, rays, tris, delta):
    """Classifies rays where `got` (closest hit, lowest index on ties) differs from
    the reference IT_BSP_FASTER record.

    tie: both are real hits of the SRLNP test at distances within `delta`
         (DELTA_BSP) of each other -- shared edges/vertices and the reference's
         DELTA_BSP interval overlap (IntersectBspFast.cpp:398-400,491).
    surface_start: the closest hit is at distance <= delta (the ray starts on
         the mesh, e.g. at a vertex) and the reference's triangle is a real hit
         further along; the reference's BSP plane test classifies the start
         point as in front of/behind the plane (IntersectBspFast.cpp:443-445)
         and can step past the triangle the ray starts on.
    Returns dict of counts and the indices of unexplained mismatches."""
    import oracle_py as O
    diff = np.where(ref["tri"] != got["tri"])[0]
    ties, starts, bad = [], [], []
    for k in diff:
        r = rays[k]
        o, d = r[0:3], r[4:7]
        rt, gt = int(ref["tri"][k]), int(got["tri"][k])
        if rt < 0 or gt < 0:
            bad.append(k)
            continue
        mn, mx = tris["box"]
        ok, lo, hi = O.box_intersect(mn, mx, o, d, r[3], r[7])
        T = tris["rec"][rt]
        rec = np.array([T], dtype=O.TRI_DTYPE)
        th = np.zeros(1, np.float32); u = np.zeros(1, np.float32); v = np.zeros(1, np.float32)
        f = np.zeros(1, np.int32)
        import ctypes
        hit = O.lib().lsgo_srlnp_hit(rec.ctypes.data_as(ctypes.c_void_p), np.ascontiguousarray(o, np.float32).ctypes.data_as(ctypes.c_void_p),
                                     np.ascontiguousarray(d, np.float32).ctypes.data_as(ctypes.c_void_p), -np.inf, np.inf,
                                     th.ctypes.data_as(ctypes.c_void_p), u.ctypes.data_as(ctypes.c_void_p),
                                     v.ctypes.data_as(ctypes.c_void_p), f.ctypes.data_as(ctypes.c_void_p))
        if hit and abs(float(th[0]) - float(got["t"][k])) <= delta:
            ties.append(k)
        elif hit and float(got["t"][k]) <= delta and float(th[0]) >= float(got["t"][k]):
            starts.append(k)
        else:
            bad.append(k)
    return dict(n=len(ref), mismatched=len(diff), ties=len(ties), surface_start=len(starts),
                unexplained=np.array(bad, dtype=np.int64))


def srlnp_table(verts, idx):
    import oracle_py as O
    rec = np.zeros(len(idx), O.TRI_DTYPE)
    for t in range(len(idx)):
        rec[t] = O.srlnp_setup(verts[idx[t, 0]], verts[idx[t, 1]], verts[idx[t, 2]])
    mn, mx = O.mesh_box(verts)
    return dict(rec=rec, box=(mn, mx))


def plane_node_explained(ref, got, rays, verts, idx, k):
    """True when the reference's distance for ray k equals, bit for bit, the plane
    distance of SOME triangle of the mesh: the reference's plane-node rule
    (IntersectBspFast.cpp:436-442) computes the distance of every triangle in a
    BSP plane node from the node's first triangle's plane."""
    import oracle_py as O
    r = rays[k].astype(np.float32)
    o, d = r[0:3], r[4:7]
    v0 = verts[idx[:, 0]]; v1 = verts[idx[:, 1]]; v2 = verts[idx[:, 2]]
    s1 = (v1 - v0).astype(np.float32); s2 = (v2 - v0).astype(np.float32)
    n = np.stack([s1[:, 1] * s2[:, 2] - s1[:, 2] * s2[:, 1],
                  -s1[:, 0] * s2[:, 2] + s1[:, 2] * s2[:, 0],
                  s1[:, 0] * s2[:, 1] - s1[:, 1] * s2[:, 0]], 1).astype(np.float32)
    ln = np.sqrt((n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1] + n[:, 2] * n[:, 2]).astype(np.float32)).astype(np.float32)
    n = (n / ln[:, None]).astype(np.float32)
    w = -((v0[:, 0] * n[:, 0] + v0[:, 1] * n[:, 1]).astype(np.float32) + (v0[:, 2] * n[:, 2]).astype(np.float32))
    ndd = ((d[0] * n[:, 0] + d[1] * n[:, 1]).astype(np.float32) + (d[2] * n[:, 2]).astype(np.float32)).astype(np.float32)
    ndo = (((o[0] * n[:, 0] + o[1] * n[:, 1]).astype(np.float32) + (o[2] * n[:, 2]).astype(np.float32)).astype(np.float32) + w).astype(np.float32)
    with np.errstate(all="ignore"):
        t = (-ndo / ndd).astype(np.float32)
    return bool(np.any(t == ref["t"][k]))
