"""Per-vertex outputs on the GPU (lsg_vertex_plan_*) against the CPU oracle:

- corner groups and powers (angle x area) bit-exact against lsgo_vertex_corners, on a box
  room (shared vertices), a triangle soup stitched by position (with dropped and
  stitch-degenerate triangles) and the displaced icosphere;
- smoothing of per-triangle data bit-exact against lsgo_smooth_vertices (same binary32
  sums in the same triangle order);
- the per-triangle final gather (one texel per triangle, the whole triangle its subtexel)
  against the oracle gather on the same texels (relative error <= 1e-4 per channel, bent
  normals 1e-4 absolute on the unit vector), and
  the smoothed vertex values against the oracle's smoothing of the oracle's gather, for the
  lightmap alone and with the directional and bent-normal layers;
- an object's vertex buffer: each vertex takes its group's value."""
import numpy as np
import pytest

from lightsprintsdk_amd import scenes as S

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def _room():
    from lightsprintsdk_amd import bake as B
    sc = S.box_room(seed=3, n_boxes=4, target_tris=4000, lightmap_size=32)
    sb = B.SceneBake(sc, B.GatherParams(quality=64, ray_length_min=S.minimal_safe_distance(sc.verts), seed=11))
    sb.first_gather()
    return sc, sb


def _soup_scene():
    """The box room as a triangle soup (own vertices per triangle), plus a collinear triangle
    and one whose two corners share a position: stitching joins the soup back."""
    from lightsprintsdk_amd import bake as B
    sc = S.box_room(seed=5, n_boxes=3, target_tris=3000, lightmap_size=16)
    v = sc.verts[sc.idx].reshape(-1, 3)
    n = sc.normals[sc.idx].reshape(-1, 3)
    extra_v = np.float32([[1, 1, 1], [2, 1, 1], [3, 1, 1], [1, 2, 1], [1, 2, 1], [1, 3, 2]])
    extra_n = np.float32([[0, 0, 1]] * 6)
    v = np.concatenate([v, extra_v]).astype(np.float32)
    n = np.concatenate([n, extra_n]).astype(np.float32)
    nt = len(sc.idx) + 2
    idx = np.arange(3 * nt, dtype=np.uint32).reshape(-1, 3)
    side = np.concatenate([sc.side_bits, np.uint8([5, 5])])
    alb = np.concatenate([sc.albedo, np.float32([[0.5] * 3] * 2)])
    em = np.concatenate([sc.emittance, np.zeros((2, 3), np.float32)])
    uv = np.concatenate([sc.tri_uv, np.zeros((2, 6), np.float32)])
    return S.BakeScene(verts=v, idx=idx, normals=n, tri_uv=uv, side_bits=side, albedo=alb, emittance=em,
                       objects=list(sc.objects), lights=list(sc.lights))


def _gs(sc):
    from lightsprintsdk_amd import bake as B
    ex = np.zeros((sc.ntri, 8), np.float32)
    ex[:, 0:3] = sc.emittance
    return B.GatherScene(sc.verts, sc.idx, sc.normals, sc.side_bits, ex, sc.lights)


@pytest.mark.parametrize("which", ["room", "soup", "icosphere"])
def test_gpu_corners_and_smoothing_equal_oracle(gpu, oracle, which):
    from lightsprintsdk_amd import bake as B
    if which == "room":
        sc = _room()[0]
    elif which == "soup":
        sc = _soup_scene()
    else:
        v, i = S.icosphere(30, seed=2)
        nrm = v / np.linalg.norm(v, axis=1, keepdims=True)
        nt = len(i)
        sc = S.BakeScene(verts=v.astype(np.float32), idx=i.astype(np.uint32), normals=nrm.astype(np.float32),
                         tri_uv=np.zeros((nt, 6), np.float32), side_bits=np.full(nt, 5, np.uint8),
                         albedo=np.full((nt, 3), 0.5, np.float32), emittance=np.zeros((nt, 3), np.float32))
    gs = _gs(sc)
    vp = B.VertexPlan(gs, sc)
    key, power = vp.corners()
    okey, opower = oracle.vertex_corners(sc.verts, sc.idx, vp.corner_group, vp.group_pos)
    np.testing.assert_array_equal(key, okey)
    np.testing.assert_array_equal(power.view(np.uint32), opower.view(np.uint32))
    dropped = int((key == 0xFFFFFFFF).sum())
    print(f"{which}: {sc.ntri} triangles, {vp.ngroups} groups, {dropped} corners not counted", flush=True)
    if which == "soup":
        assert vp.ngroups < len(sc.verts) and dropped == 6
    data = np.random.default_rng(1).uniform(0, 2, (sc.ntri, 4)).astype(np.float32)
    got = vp.smooth(data)
    want = oracle.smooth_vertices(okey, opower, data, vp.ngroups)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    vp.close()
    gs.close()


BENT_ATOL = 1e-4


def _cmp(g, r, what, layer=0):
    """irradiance layers: relative error per channel; bent normals: absolute error on the
    unit vector (its small components are sums that cancel)"""
    g = g.astype(np.float64)
    if layer == 4:
        err = np.abs((2 * g[:, :3] - 1) - (2 * r[:, :3] - 1))
        assert err.max() <= BENT_ATOL, (what, err.max())
        return err.max()
    err = np.abs(g[:, :3] - r[:, :3]) / np.maximum(np.abs(r[:, :3]), 1e-6)
    assert err.max() <= RTOL, (what, err.max())
    return err.max()


@pytest.mark.parametrize("layers", [(0,), (0, 1, 2, 3, 4)])
def test_gpu_vertex_gather_vs_oracle(gpu, oracle, layers):
    from lightsprintsdk_amd import bake as B
    sc, sb = _room()
    vp = B.VertexPlan(sb.gs, sc)
    p = B.GatherParams(quality=96, ray_length_min=S.minimal_safe_distance(sc.verts), seed=21)
    got = vp.gather(p, layers=layers)
    ttex, tsub = B.triangle_texels(sc.ntri)
    arrays = sb.gs.oracle_arrays(sb.exitance)
    tree = sb.gs.col.download_tree()
    if layers == (0,):
        ref = {0: oracle.gather(arrays, p.as_dict(), ttex.view(np.uint32).reshape(-1, 4), tsub, tree=tree, threads=16)[0]}
    else:
        ref = oracle.gather_layers(arrays, p.as_dict(), ttex.view(np.uint32).reshape(-1, 4), tsub, tree=tree, threads=16)
    key, power = vp.corners()
    for k in layers:
        tri = vp.triangle_data(k)
        e1 = _cmp(tri, ref[k], f"triangle layer {k}", k)
        want = oracle.smooth_vertices(key, power, ref[k].astype(np.float32), vp.ngroups)
        counted = np.zeros(vp.ngroups, bool)
        counted[key[key != 0xFFFFFFFF]] = True
        e2 = _cmp(got[k][counted], want[counted].astype(np.float64), f"vertex layer {k}", k)
        assert np.all(got[k][~counted] == 0)
        print(f"layer {k}: triangles max rel err {e1:.1e}, vertices {e2:.1e}", flush=True)
    # an object's vertex buffer
    ids, vals = vp.object_vertices(1, got[0])
    ob = sc.objects[1]
    assert set(ids) == set(np.unique(sc.idx[ob.tri_begin:ob.tri_begin + ob.tri_count]))
    np.testing.assert_array_equal(vals, got[0][vp.vertex_group[ids], :3])
    assert vals.max() > 0
    vp.close()
    sb.gs.close()
