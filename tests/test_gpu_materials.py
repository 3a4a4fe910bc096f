"""Point materials on the GPU (lsg_scene_set_point_materials) against the CPU oracle.

- The closed-form scenes of test_materials_cpu.py (split-coloured ceiling, emissive
  texture, keyed checkerboard shadows with and without interpolation, hemisphere rays
  through a 1/4-transparent keyed occluder): the GPU gather equals the oracle within float
  rounding (shadow answers exactly).
- A box room whose objects carry every supported kind of material (scaled and unscaled byte
  and float textures, diffuse and emissive, keyed transmittance in the diffuse map's alpha
  and in a texture of its own, inverted, nearest and interpolated shadow lookups, a material
  below its minimal quality): every kernel form (grouped low-quality lanes, the per-texel
  wave, the layered kernel, the persistent plan with and without lane refill) against the
  oracle, max relative error 1e-4; the plan equals the plain gather bit for bit; a graph
  captured before the materials change is not replayed after.
- Refusals: unkeyed transmittance, a keyed material that emits, a texture index out of
  range, and a gather whose quality does not reach a keyed material's minimum."""
import numpy as np
import pytest

from lightsprintsdk_amd import scenes as S
import test_materials_cpu as M

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def _cmp(gpu, ref, sel=None):
    g = gpu[:, :3].astype(np.float64)
    r = ref[:, :3]
    if sel is not None:
        g, r = g[sel], r[sel]
    err = np.abs(g - r) / np.maximum(np.abs(r), 1e-3)
    assert err.max() <= RTOL, (err.max(), np.unravel_index(err.argmax(), err.shape))
    return err.max()


def gpu_gather(arr, quality=4096, hemisphere=1, light_multiplier=1.0, seed=5):
    """The oracle's run() of test_materials_cpu on the GPU: same scene, texel and params."""
    from lightsprintsdk_amd import bake as B
    gs = B.GatherScene(arr["verts"], arr["idx"], arr["normals"], arr["side_bits"], arr["exitance"],
                       [dict(type=int(l["type"]), position=tuple(l["pos"]), direction=tuple(l["dir"]), color=tuple(l["color"]),
                             cast_shadows=int(l["cast_shadows"])) for l in arr["lights"]])
    pm = arr.get("pm")
    if pm is not None:
        gs.set_point_materials(pm["tri_mat"], pm["tri_uv"], pm["mats"], pm["textures"], pm["gamma"])
    p = B.GatherParams(quality=quality, hemisphere=hemisphere, ray_length_min=1e-6, locality=1e5, rug_distance=1e-3,
                       light_multiplier=light_multiplier, seed=seed)
    tex, sub = M._single_texel(0, uv=(0.4, 0.4, 0.6, 0.4, 0.4, 0.6))
    out = gs.gather(tex, sub, p, nout=1)
    gs.close()
    return out[0, :3]


def _gpu_run(arr, p, tex, sub):
    """Adapter with the oracle.gather signature (used by the closed-form checks)."""
    out = gpu_gather(arr, quality=p["quality"], hemisphere=p["hemisphere"], light_multiplier=p["light_multiplier"],
                     seed=p["seed"])
    return out[None].astype(np.float64), None


CLOSED = {
    "diffuse split": lambda: M.split_ceiling("diffuse"),
    "emissive split": lambda: M.split_ceiling("emissive"),
    "below min quality": lambda: M.split_ceiling("diffuse", min_quality=5000),
    "back side emits": lambda: M.split_ceiling("emissive", emit_sides=2),
}


@pytest.mark.parametrize("name", list(CLOSED))
def test_gpu_closed_form_ceilings_equal_oracle(gpu, oracle, name):
    arr = CLOSED[name]()
    got = gpu_gather(arr)
    want = M.run(oracle.gather, arr)
    np.testing.assert_allclose(got, want, rtol=RTOL, atol=1e-7)
    print(f"{name}: gpu {got}, oracle {want}", flush=True)


def test_gpu_closed_forms_hold(gpu, oracle):
    """The same closed-form checks the oracle passes, run through the GPU gather."""
    M.check_split(_gpu_run)
    M.check_shadows(_gpu_run)
    M.check_quarter(_gpu_run)
    g = np.load(M.GOLDEN)
    arr = M.with_pm(M.scene([(M.H, M.W, True, 1.0, 0, 1.0)]), [M._mat(diffuse_texture=0)],
                    [M._tex(0, np.uint8([[[200, 100, 50]]]), True)], gamma=float(g["gamma"]))
    plain = M.with_pm(M.scene([(M.H, M.W, True, 1.0, 0, 1.0)]), [M._mat(albedo=[1, 1, 1])], [])
    ratio = gpu_gather(arr, quality=1024).astype(np.float64) / gpu_gather(plain, quality=1024)
    np.testing.assert_allclose(ratio, g["byte_linear"][[200, 100, 50]], rtol=1e-6)


# ---- a box room with every kind of material ------------------------------------------------

def room_materials(sc, seed=7, quality=64):
    """Materials for the objects of a box room; returns (tri_mat, tri_uv, mats, textures)."""
    from lightsprintsdk_amd import bake as B
    rng = np.random.default_rng(seed)
    T = []
    T.append(B.texture(B.FORMAT_RGB, rng.integers(0, 256, (16, 16, 3)), True))                       # 0 diffuse, scaled bytes
    a = rng.integers(0, 256, (8, 8, 4)).astype(np.uint8)
    a[:, :, 3] = np.where(rng.random((8, 8)) < 0.4, 0, 255)
    a[0, :, 3] = 128                                                                                  # on the threshold
    T.append(B.texture(B.FORMAT_RGBA, a, True))                                                       # 1 alpha-keyed
    T.append(B.texture(B.FORMAT_BGR, rng.integers(0, 256, (4, 8, 3)), False))                         # 2 unscaled BGR
    T.append(B.texture(B.FORMAT_RGBF, rng.random((4, 4, 3)).astype(np.float32), False))               # 3 transmittance
    T.append(B.texture(B.FORMAT_RGBF, (2 * rng.random((8, 8, 3))).astype(np.float32), True))          # 4 emissive
    T.append(B.texture(B.FORMAT_LUMINANCEF, rng.random((4, 4)).astype(np.float32), True))             # 5 (unused below min)
    t6 = rng.random((4, 4, 4)).astype(np.float32)
    t6[:, :, 3] = np.where(rng.random((4, 4)) < 0.5, 0.0, 1.0)
    T.append(B.texture(B.FORMAT_RGBAF, t6, False))                                                    # 6 alpha of its own
    T.append(B.texture(B.FORMAT_LUMINANCE, rng.integers(0, 256, (8, 8)), True))                       # 7 diffuse luminance
    mats = np.zeros(6, B.POINT_MATERIAL_DTYPE)
    mats["diffuse_texture"] = -1
    mats["emissive_texture"] = -1
    mats["transmittance_texture"] = -1
    mats["threshold"] = 0.5
    mats["emittance_multiplier"] = 1.0
    mats["albedo"] = rng.uniform(0.2, 0.9, (6, 3))
    mats[0]["diffuse_texture"] = 0
    mats[1]["diffuse_texture"] = 1
    mats[1]["transmittance_texture"] = 1
    mats[1]["flags"] = B.PM_TRANSMITTANCE_IN_ALPHA | B.PM_KEYED
    mats[1]["min_quality"] = 4                      # quality / 10 may or may not exceed it: both lookups
    mats[2]["diffuse_texture"] = 2
    mats[2]["transmittance_texture"] = 3
    mats[2]["flags"] = B.PM_KEYED | B.PM_INVERTED
    mats[2]["threshold"] = 0.45
    mats[3]["emissive_texture"] = 4
    mats[3]["emit_sides"] = 1
    mats[3]["emittance_multiplier"] = 3.0
    mats[3]["emittance"] = 0.2
    mats[4]["diffuse_texture"] = 5
    mats[4]["min_quality"] = 100000                 # per-triangle material: albedo and flat emittance
    mats[4]["emit_sides"] = 3
    mats[4]["emittance"] = 0.1
    mats[5]["diffuse_texture"] = 7
    mats[5]["transmittance_texture"] = 6
    mats[5]["flags"] = B.PM_TRANSMITTANCE_IN_ALPHA | B.PM_KEYED
    tri_mat = np.full(sc.ntri, B.NO_MATERIAL, np.uint32)
    for k, ob in enumerate(sc.objects):
        m = k % 7
        if m < 6:
            tri_mat[ob.tri_begin:ob.tri_begin + ob.tri_count] = m
    uv = (sc.tri_uv * np.float32(3) + np.float32(0.37)).astype(np.float32)
    return tri_mat, uv, mats, T


def _room(seed=3, quality=64, rounds=0):
    from lightsprintsdk_amd import bake as B
    sc = S.box_room(seed=seed, n_boxes=8, target_tris=5000, lightmap_size=24)
    sc.side_bits = (sc.side_bits | 4).astype(np.uint8)      # every object casts shadows
    params = B.GatherParams(quality=quality, ray_length_min=S.minimal_safe_distance(sc.verts), seed=11)
    params.light_rounds = rounds
    return sc, params


def _table(sc, seed=2):
    rng = np.random.default_rng(seed)
    ex = np.zeros((sc.ntri, 8), np.float32)
    ex[:, 0:3] = rng.uniform(0, 2, (sc.ntri, 3))
    ex[:, 4:7] = rng.uniform(0, 1, (sc.ntri, 3))
    return ex


@pytest.mark.parametrize("quality,rounds", [(16, 0), (64, 0), (64, 40)])
def test_gpu_room_gather_vs_oracle(gpu, oracle, quality, rounds):
    from lightsprintsdk_amd import bake as B
    sc, params = _room(quality=quality, rounds=rounds)
    sb = B.SceneBake(sc, params)
    tm, uv, mats, T = room_materials(sc, quality=quality)
    sb.gs.set_point_materials(tm, uv, mats, T, 0.45)
    ex = _table(sc)
    sb.gs.set_exitance(ex)
    sb.final_gather()
    flat = np.concatenate([m.reshape(-1, 4) for m in sb.lightmaps_rgba()])
    tx = sb.texels
    ref, st = oracle.gather(sb.gs.oracle_arrays(ex), params.as_dict(), tx.view(np.uint32).reshape(-1, 4), sb.subs,
                            tree=sb.gs.col.download_tree(), threads=16)
    print(f"quality {quality}, rounds {rounds}: {len(tx)} texels, {st.hemi_rays:.0f} hemisphere + {st.shadow_rays:.0f} "
          f"shadow rays, max rel err {_cmp(flat, ref, tx['out_index']):.2e}", flush=True)
    # the materials matter: the same gather without them differs
    sb.gs.set_point_materials(None)
    sb.clear()
    sb.final_gather()
    plain = np.concatenate([m.reshape(-1, 4) for m in sb.lightmaps_rgba()])
    assert np.abs(plain[:, :3] - flat[:, :3]).max() > 0.1


def test_gpu_room_layers_vs_oracle(gpu, oracle):
    from lightsprintsdk_amd import bake as B
    sc, params = _room(quality=48)
    sb = B.SceneBake(sc, params)
    tm, uv, mats, T = room_materials(sc)
    sb.gs.set_point_materials(tm, uv, mats, T, 0.45)
    ex = _table(sc)
    sb.gs.set_exitance(ex)
    got = sb.gs.gather_layers(sb.texels, sb.subs, params, layers=(0, 1, 4), nout=sb.nout)
    ref = oracle.gather_layers(sb.gs.oracle_arrays(ex), params.as_dict(), sb.texels.view(np.uint32).reshape(-1, 4), sb.subs,
                               layers=(0, 1, 4), tree=sb.gs.col.download_tree(), threads=16)
    sel = sb.texels["out_index"]
    for k in (0, 1):
        print(f"layer {k}: max rel err {_cmp(got[k], ref[k], sel):.2e}", flush=True)
    err = np.abs((2 * got[4][sel, :3] - 1) - (2 * ref[4][sel, :3] - 1)).max()
    assert err <= 1e-4, err


def test_gpu_room_plan_equals_plain_and_oracle(gpu, oracle):
    from lightsprintsdk_amd import bake as B
    from lightsprintsdk_amd import native as N
    sc, params = _room(quality=96, rounds=40)
    tm, uv, mats, T = room_materials(sc)
    plan = B.BakePlan(sc, params)
    plan.gs.set_point_materials(tm, uv, mats, T, 0.45)
    plan.set_tuning(lane_refill=1)
    plan.run(first=True, graph=False)
    got = plan.rgba()
    ref = B.SceneBake(sc, params)
    ref.gs.set_point_materials(tm, uv, mats, T, 0.45)
    ref.first_gather()
    ex = ref.exitance
    # a triangle with a material keeps its direct irradiance: no albedo, no emittance
    np.testing.assert_array_equal(plan.exitance().view(np.uint32), ex.view(np.uint32))
    ref.bake()
    want = ref.d_out.download(np.float32, (max(ref.nout, 1), 4))
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    # the final gather of the plan's table against the oracle (before the filters)
    ref.clear()
    ref.final_gather()
    flat = np.concatenate([m.reshape(-1, 4) for m in ref.lightmaps_rgba()])
    tx = ref.texels
    oref, _ = oracle.gather(ref.gs.oracle_arrays(ex), params.as_dict(), tx.view(np.uint32).reshape(-1, 4), ref.subs,
                            tree=ref.gs.col.download_tree(), threads=16)
    print(f"plan table: final gather max rel err {_cmp(flat, oref, tx['out_index']):.2e}", flush=True)
    # lanes refilled as they free up: same rays, sums in another order
    plan.set_tuning(lane_refill=2)
    plan.run(first=True, graph=False)
    np.testing.assert_allclose(plan.rgba()[:, :3], want[:, :3], rtol=2e-5, atol=1e-7)
    # graph: captured with these materials, replayed; new materials force a new capture
    plan.run(first=True, graph=True)
    a = plan.rgba().copy()
    N.check(N.lib().lsg_synchronize(), "sync")
    mats2 = mats.copy()
    mats2["albedo"] = mats2["albedo"][::-1]
    mats2[0]["diffuse_texture"] = 7
    plan.gs.set_point_materials(tm, uv, mats2, T, 0.45)
    plan.run(first=True, graph=True)
    b = plan.rgba().copy()
    plan.run(first=True, graph=False)
    np.testing.assert_array_equal(plan.rgba().view(np.uint32), b.view(np.uint32))
    assert not np.array_equal(a, b)


def test_gpu_point_material_refusals(gpu):
    from lightsprintsdk_amd import bake as B
    from lightsprintsdk_amd import native as N
    sc, params = _room(quality=32)
    sb = B.SceneBake(sc, params)
    tm, uv, mats, T = room_materials(sc)
    bad = mats.copy()
    bad[2]["flags"] = 0                               # transmittance, not keyed
    with pytest.raises(N.LsgError):
        sb.gs.set_point_materials(tm, uv, bad, T, 0.45)
    bad = mats.copy()
    bad[1]["emit_sides"] = 1
    bad[1]["emittance"] = 0.5                          # keyed and emitting
    with pytest.raises(N.LsgError):
        sb.gs.set_point_materials(tm, uv, bad, T, 0.45)
    bad = mats.copy()
    bad[0]["diffuse_texture"] = len(T)                 # no such texture
    with pytest.raises(N.LsgError):
        sb.gs.set_point_materials(tm, uv, bad, T, 0.45)
    tm_bad = tm.copy()
    tm_bad[0] = len(mats)
    with pytest.raises(N.LsgError):
        sb.gs.set_point_materials(tm_bad, uv, mats, T, 0.45)
    # a keyed material whose minimal quality the gather does not exceed
    hi = mats.copy()
    hi[1]["min_quality"] = 32
    sb.gs.set_point_materials(tm, uv, hi, T, 0.45)
    with pytest.raises(N.LsgError):
        sb.final_gather()
    sb.gs.set_point_materials(tm, uv, mats, T, 0.45)
    sb.final_gather()
    N.check(N.lib().lsg_synchronize(), "sync")
