"""GPU parity of the collider: liblsgpu's LBVH build and closest-hit traversal on
gfx950 against the CPU oracle (bit-exact) and the compiled reference IT_BSP_FASTER
goldens (triangle bit-exact up to counted ties)."""
import ctypes
import os

import numpy as np
import pytest

from lightsprintsdk_amd import scenes
import parity

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def _col(V, I, leaf=4, lbvh_only=False):
    from lightsprintsdk_amd.collider import RRCollider
    return RRCollider.create(V, I, leaf_size=leaf, lbvh_only=lbvh_only)


def _same_hits(g, gf, o, h):
    """GPU (tri,t,u,v,front) vs oracle hit records: all fields bit-exact."""
    np.testing.assert_array_equal(g["tri"], o["tri"])
    m = o["tri"] >= 0
    for f in ("t", "u", "v"):
        np.testing.assert_array_equal(g[f][m], o[f][m], err_msg=f)
    np.testing.assert_array_equal(gf[m].astype(np.int32), o["front"][m])


@pytest.mark.parametrize("freq,leaf", [(30, 1), (30, 4), (30, 8), (100, 1), (100, 2)])
def test_gpu_tree_equals_oracle_tree(gpu, oracle, freq, leaf):
    """The GPU LBVH equals the oracle's node for node; the default build's tree equals
    the host SAH rebuild of that LBVH (lsg_bvh_rebuild_sah, checked on the CPU).  At
    frequency 100 (200,000 triangles) the top tasks have more than 2^15 leaves and are
    rebuilt across the chip."""
    from lightsprintsdk_amd.collider import rebuild_sah
    V, I = scenes.icosphere(freq, seed=11)
    c = _col(V, I, leaf, lbvh_only=True)
    t_gpu = c.download_tree()
    t_cpu = oracle.lbvh_build(V, I, leaf_size=leaf)
    assert t_gpu["nvalid"] == t_cpu["nvalid"] and t_gpu["root"] == t_cpu["root"]
    assert t_gpu["depth"] == t_cpu["depth"]
    np.testing.assert_array_equal(t_gpu["prim"], t_cpu["prim"])
    np.testing.assert_array_equal(t_gpu["nodes"].view(np.uint32), t_cpu["nodes"].view(np.uint32))
    t_sah = _col(V, I, leaf).download_tree()
    want = rebuild_sah(t_cpu)
    assert t_sah["root"] == want["root"] and t_sah["depth"] == want["depth"] <= t_cpu["depth"]
    np.testing.assert_array_equal(t_sah["prim"], t_cpu["prim"])
    np.testing.assert_array_equal(t_sah["nodes"].view(np.uint32), want["nodes"].view(np.uint32))


def test_gpu_vs_oracle_small(gpu, oracle):
    V, I = scenes.icosphere(10, seed=12)
    R = scenes.random_rays(V, 20000, seed=13)
    c = _col(V, I)
    g, gf = c.intersect_batch(R)
    o = oracle.intersect_linear_srlnp(V, I, R, threads=16)
    _same_hits(g, gf, o, None)


def test_gpu_vs_oracle_icosphere_full(gpu, oracle):
    """Full size: 100,820-triangle displaced icosphere, 1M rays,
    bit-exact against the oracle's traversal of the same tree (itself equal to
    brute force, pinned on CPU)."""
    V, I = scenes.icosphere(scenes.ICOSPHERE_100K_FREQUENCY, seed=1)
    R = scenes.random_rays(V, 1_000_000, seed=2)
    c = _col(V, I)
    print(f"icosphere: tris={len(I)} rays={len(R)}", flush=True)
    g, gf = c.intersect_batch(R)
    tree = c.download_tree()
    o, st = oracle.bvh_traverse(tree, V, I, R, threads=16)
    _same_hits(g, gf, o, None)
    sub = slice(0, 3000)
    lin = oracle.intersect_linear_srlnp(V, I, R[sub], threads=16)
    np.testing.assert_array_equal(lin["tri"], o["tri"][sub])
    hit_rate = float((g["tri"] >= 0).mean())
    assert 0.2 < hit_rate < 0.9


def test_gpu_vs_reference_bsp_faster_golden(gpu, oracle):
    """Against the compiled reference IT_BSP_FASTER on the 100,820-triangle icosphere: the
    same triangle for every golden ray, distance bit-exact except the counted
    plane-node cases."""
    z = _load("collider_ico100k")
    V, I = scenes.icosphere(int(z["frequency"]), seed=int(z["mesh_seed"]))
    assert scenes.mesh_digest(V, I) == str(z["mesh_sha256"])
    R = z["rays"]
    ref = z["bsp_faster"]
    g, gf = _col(V, I).intersect_batch(R)
    assert np.array_equal(g["tri"], ref["tri"])
    m = ref["tri"] >= 0
    tdiff = np.where(m & (g["t"] != ref["t"]))[0]
    for k in tdiff:
        assert parity.plane_node_explained(ref, g, R, V, I, k)
    assert len(tdiff) <= 4
    same = m & (g["t"] == ref["t"])
    for f in ("u", "v"):
        np.testing.assert_array_equal(g[f][same], ref[f][same])
    np.testing.assert_array_equal(gf[m].astype(np.int32), ref["front"][m])


def test_gpu_vs_reference_bsp_faster_100k_rays(gpu, oracle):
    """100,000 rays of the 100,820-triangle icosphere against the compiled IT_BSP_FASTER:
    the GPU's hits equal the oracle's bit for bit, and against the reference every ray is
    identical or in a counted class: shared-edge tie, surface start, plane-node distance
    (the reference's t is another triangle's plane distance), plane-node miss (the
    reference returned a farther triangle past a nearly coplanar one)."""
    V, I, R, ref = parity.wide_golden(GOLDEN)
    g, gf = _col(V, I).intersect_batch(R)
    o, _ = oracle.bvh_traverse(oracle.lbvh_build(V, I, leaf_size=2), V, I, R, threads=16)
    np.testing.assert_array_equal(g["tri"], o["tri"])
    h = g["tri"] >= 0
    np.testing.assert_array_equal(g["t"][h], o["t"][h])
    tab = parity.srlnp_table(V, I)
    delta = float(oracle.lib().lsgo_delta_bsp(*[np.ascontiguousarray(b).ctypes.data_as(ctypes.c_void_p) for b in tab["box"]]))
    c = parity.count_classes(ref, g, R, V, I, 10 * delta, tab)
    print("GPU vs IT_BSP_FASTER, 100k rays:", c, flush=True)
    assert c["unexplained"] == 0
    assert c["exact"] + c["shared_edge_ties"] + c["surface_start"] + c["plane_node_distance"] + c["plane_node_miss"] == c["rays"]
    assert c["exact"] >= 0.999 * c["rays"]


@pytest.mark.parametrize("name", ["collider_edge", "collider_planes"])
def test_gpu_edge_cases(gpu, oracle, name):
    """Axis-aligned rays (zero direction components), rays from vertices and
    edges, tmin>0, short segments, coplanar groups: GPU == oracle bit-exact, and
    vs the reference only counted ties / surface starts."""
    z = _load(name)
    V, I, R = z["verts"], z["idx"], z["rays"]
    g, gf = _col(V, I).intersect_batch(R)
    o = oracle.intersect_linear_srlnp(V, I, R)
    _same_hits(g, gf, o, None)
    tab = parity.srlnp_table(V, I)
    mn, mx = tab["box"]
    delta = float(oracle.lib().lsgo_delta_bsp(mn.ctypes.data_as(ctypes.c_void_p), mx.ctypes.data_as(ctypes.c_void_p)))
    cls = parity.classify_vs_reference(z["bsp_faster"], g, R, tab, 10 * delta)
    print(name, {k: (v if k != "unexplained" else len(v)) for k, v in cls.items()}, flush=True)
    assert len(cls["unexplained"]) == 0


def test_gpu_rrray_batch(gpu, oracle):
    """intersectBatch over the reference's 128-byte RRRay records: every output
    field the flags request, equal to the reference IT_BSP_FASTER golden."""
    from lightsprintsdk_amd.collider import make_rrrays
    z = _load("collider_planes")
    V, I, R = z["verts"], z["idx"], z["rays"]
    ref = z["bsp_faster"]
    rr = make_rrrays(R, flags=0x3F)
    mask = _col(V, I).intersect_rrrays(rr)
    np.testing.assert_array_equal(mask.astype(bool), ref["tri"] >= 0)
    m = ref["tri"] >= 0
    np.testing.assert_array_equal(rr["hitTriangle"][m], ref["tri"][m].astype(np.uint32))
    np.testing.assert_array_equal(rr["hitDistance"][m], ref["t"][m])
    np.testing.assert_array_equal(rr["hitPoint2d"][m], np.stack([ref["u"], ref["v"]], 1)[m])
    np.testing.assert_array_equal(rr["hitPoint3d"][m], ref["p"][m])
    np.testing.assert_array_equal(rr["hitPlane"][m], ref["plane"][m])
    np.testing.assert_array_equal(rr["hitFrontSide"][m].astype(np.int32), ref["front"][m])
    # flags limit what is written: FILL_TRIANGLE only leaves distance untouched
    rr2 = make_rrrays(R, flags=1 << 4)
    rr2["hitDistance"] = -7.0
    _col(V, I).intersect_rrrays(rr2)
    assert np.all(rr2["hitDistance"] == -7.0)
    np.testing.assert_array_equal(rr2["hitTriangle"][m], ref["tri"][m].astype(np.uint32))


def test_gpu_degenerate_and_tiny_meshes(gpu, oracle):
    rng = np.random.default_rng(3)
    # single triangle (root is a leaf), and a mesh with degenerate triangles mixed in
    V = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [2, 2, 2], [3, 3, 3], [4, 4, 4]])
    I1 = np.uint32([[0, 1, 2]])
    I2 = np.uint32([[0, 1, 2], [3, 4, 5], [0, 0, 1], [2, 1, 0]])
    for I in (I1, I2):
        R = np.zeros((500, 8), np.float32)
        R[:, 0:3] = rng.uniform(-0.2, 1.2, (500, 3)); R[:, 2] = 1.0
        d = rng.normal(size=(500, 3)); d[:, 2] = -np.abs(d[:, 2]) - 0.5
        R[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
        R[:, 7] = np.inf
        c = _col(V, I)
        g, gf = c.intersect_batch(R)
        o = oracle.intersect_linear_srlnp(V, I, R)
        _same_hits(g, gf, o, None)
        assert c.info().num_valid == (1 if len(I) == 1 else 2)
    # zero rays is a no-op
    c = _col(V, I2)
    g, gf = c.intersect_batch(np.zeros((0, 8), np.float32))
    assert len(g) == 0


def test_gpu_deterministic(gpu):
    V, I = scenes.icosphere(25, seed=21)
    R = scenes.random_rays(V, 200000, seed=22)
    c = _col(V, I)
    a = c.intersect_batch(R)
    b = c.intersect_batch(R)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    c2 = _col(V, I)
    np.testing.assert_array_equal(c.download_tree()["nodes"], c2.download_tree()["nodes"])


@pytest.mark.parametrize("scene", ["ico100k", "room", "deep"])
def test_gpu_bvh4_stack_bound(gpu, oracle, scene):
    """The GPU traverses a 4-wide tree collapsed from the LBVH under a stack budget: the
    worst-case stack need stays below the 32 (or 64) LDS entries, and hits equal the
    oracle's."""
    from lightsprintsdk_amd.collider import RRCollider
    if scene == "ico100k":
        V, I = scenes.icosphere(scenes.ICOSPHERE_100K_FREQUENCY, seed=1)
    elif scene == "room":
        sc = scenes.box_room(seed=3, n_boxes=20, target_tris=50_000, lightmap_size=256)
        V, I = sc.verts, sc.idx
    else:
        # a fan of slivers crowded towards one end: Morton codes collide, the tree is
        # deep (26 levels at one triangle per leaf)
        n = 3000
        a = np.linspace(0, 1, n + 1, dtype=np.float64) ** 6
        V = np.zeros((n + 2, 3), np.float32)
        V[1:, 0] = a
        V[1:, 1] = a ** 2
        V[1:, 2] = 0.01 * np.sin(np.arange(n + 1))
        I = np.stack([np.zeros(n, np.uint32), np.arange(1, n + 1, dtype=np.uint32),
                      np.arange(2, n + 2, dtype=np.uint32)], 1)
    c = RRCollider.create(V, I, leaf_size=1 if scene == "deep" else 0)
    inf = c.info()
    stack = 32 if inf.depth + 1 <= 32 else 64
    print(f"{scene}: depth {inf.depth}, BVH4 nodes {inf.num_nodes4}, stack need {inf.stack_need} of {stack}", flush=True)
    assert inf.stack_need <= stack - 1
    assert inf.num_nodes4 >= 1
    R = scenes.random_rays(V, 20000, seed=5)
    g, gf = c.intersect_batch(R)
    o = oracle.intersect_linear_srlnp(V, I, R)
    np.testing.assert_array_equal(g["tri"], o["tri"])
