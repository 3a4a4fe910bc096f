"""GPU parity of environment lighting: hemisphere rays that leave the scene sample a
cube map or a panorama (PathtracerWorker::getIncidentRadiance, pathtracer.cpp:156-173;
RRBufferInMemory::getElementAtDirection, RRBufferInMemory.cpp:410-455).  An open city
block under random environments, gathered by liblsgpu's environment kernels and by
the CPU oracle on the same rays: relative error <= 1e-4, alphas bit-exact; the
persistent bake plan gives the same texels as the one-wave-per-texel gather."""
import numpy as np
import pytest

from lightsprintsdk_amd import scenes as S

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def _scene():
    return S.city_terrain(seed=7, n_boxes=30, grid_vertices=40, terrain_tiles=2, lightmap_size=48, city_maps=2,
                          box_tris=3000, cell_m=1.0)


def _env(kind, seed):
    rng = np.random.default_rng(seed)
    if kind == 1:
        w = h = 16
        n = 6 * w * w
    else:
        w, h = 64, 32
        n = w * h
    return dict(kind=kind, width=w, height=h, rgb=rng.uniform(0.0, 2.0, (n, 3)).astype(np.float32))


def _cmp(gpu, ref, sel, what):
    g = gpu[sel].astype(np.float64)
    r = ref[sel]
    err = np.abs(g[:, :3] - r[:, :3]) / np.maximum(np.abs(r[:, :3]), 1e-6)
    assert err.max() <= RTOL, (what, err.max())
    np.testing.assert_array_equal(g[:, 3].astype(np.float32), r[:, 3].astype(np.float32))
    return err.max()


@pytest.mark.parametrize("kind,quality", [(1, 64), (2, 64), (2, 16), (1, 24)])
def test_gpu_environment_vs_oracle(gpu, oracle, kind, quality):
    from lightsprintsdk_amd import bake as B
    sc = _scene()
    sb = B.SceneBake(sc, B.GatherParams(quality=quality, ray_length_min=S.minimal_safe_distance(sc.verts), seed=9,
                                        environment_multiplier=1.3))
    e = _env(kind, seed=kind * 10 + quality)
    sb.gs.set_environment(e["kind"], e["width"], e["height"], e["rgb"])
    sb.first_gather()
    sb.final_gather()
    flat = sb.d_out.download(np.float32, (sb.nout, 4))
    tx = sb.texels
    ref, st = oracle.gather(sb.gs.oracle_arrays(sb.exitance), sb.params.as_dict(), tx.view(np.uint32).reshape(-1, 4),
                            sb.subs, tree=sb.gs.col.download_tree(), threads=16)
    err = _cmp(flat, ref, tx["out_index"], "final")
    # the environment reaches the texels: the sky-facing terrain is brighter than the sun alone
    sb.gs.set_environment(0)
    sb.clear()
    sb.final_gather()
    dark = sb.d_out.download(np.float32, (sb.nout, 4))
    assert flat[tx["out_index"], :3].mean() > dark[tx["out_index"], :3].mean() + 0.1
    print(f"env kind {kind} quality {quality}: {len(tx)} texels, max rel err {err:.2e}", flush=True)


def test_gpu_environment_layers_vs_oracle(gpu, oracle):
    from lightsprintsdk_amd import bake as B
    sc = _scene()
    sb = B.SceneBake(sc, B.GatherParams(quality=64, ray_length_min=S.minimal_safe_distance(sc.verts), seed=2,
                                        environment_multiplier=0.8))
    e = _env(2, seed=5)
    sb.gs.set_environment(e["kind"], e["width"], e["height"], e["rgb"])
    sb.first_gather()
    tx = sb.texels
    layers = (0, 1, 2, 3, 4)
    got = sb.gs.gather_layers(tx, sb.subs, sb.params, layers=layers, nout=sb.nout)
    ref = oracle.gather_layers(sb.gs.oracle_arrays(sb.exitance), sb.params.as_dict(), tx.view(np.uint32).reshape(-1, 4),
                               sb.subs, layers=layers, tree=sb.gs.col.download_tree(), threads=16)
    sel = tx["out_index"]
    for k in (0, 1, 2, 3):
        _cmp(got[k], ref[k], sel, f"layer {k}")
    bent = np.abs((2 * got[4][sel, :3] - 1) - (2 * ref[4][sel, :3] - 1))
    assert bent.max() <= 1e-4, bent.max()


def test_gpu_environment_plan_matches_scene_bake(gpu):
    """The persistent kernel's environment variant (bake plan, graph replay included)
    gives the texels of the one-wave-per-texel environment gather."""
    from lightsprintsdk_amd import bake as B
    sc = _scene()
    params = B.GatherParams(quality=64, ray_length_min=S.minimal_safe_distance(sc.verts), seed=3,
                            environment_multiplier=1.0)
    e = _env(1, seed=3)
    ref = B.SceneBake(sc, params)
    ref.gs.set_environment(e["kind"], e["width"], e["height"], e["rgb"])
    ref.first_gather()
    ref.bake()
    want = ref.d_out.download(np.float32, (max(ref.nout, 1), 4))
    plan = B.BakePlan(sc, params)
    plan.set_tuning(lane_refill=1)       # 64 shooters at a time, as the plain gather
    plan.gs.set_environment(e["kind"], e["width"], e["height"], e["rgb"])
    plan.run(first=True, graph=False)
    np.testing.assert_array_equal(plan.rgba().view(np.uint32), want.view(np.uint32))
    for rep in range(2):
        plan.run(first=True, graph=True)
        np.testing.assert_array_equal(plan.rgba().view(np.uint32), want.view(np.uint32), err_msg=f"replay {rep}")
    # a new environment after capture is seen by the next replay
    e2 = _env(2, seed=4)
    plan.gs.set_environment(e2["kind"], e2["width"], e2["height"], e2["rgb"])
    plan.run(first=True, graph=True)
    assert not np.array_equal(plan.rgba(), want)
    plan.close()


@pytest.mark.parametrize("kind,quality,rounds", [(1, 16, 0), (2, 20, 8)])
def test_gpu_environment_lanes_vs_oracle(gpu, oracle, kind, quality, rounds):
    """The same open block with 5.3M texels: low-quality environment gathers run one texel
    per lane; the oracle re-gathers a spread and a random sample of texels."""
    from lightsprintsdk_amd import bake as B
    sc = S.city_terrain(seed=7, n_boxes=30, grid_vertices=40, terrain_tiles=2, lightmap_size=1024, city_maps=2,
                        box_tris=3000, cell_m=1.0)
    sb = B.SceneBake(sc, B.GatherParams(quality=quality, ray_length_min=S.minimal_safe_distance(sc.verts), seed=13,
                                        environment_multiplier=0.7, light_rounds=rounds))
    n = len(sb.texels)
    assert n >= 16 * 64 * 16 * 256, n      # the size from which the one-texel-per-lane kernel runs
    e = _env(kind, seed=kind + quality)
    sb.gs.set_environment(e["kind"], e["width"], e["height"], e["rgb"])
    sb.first_gather()
    sb.final_gather()
    flat = sb.d_out.download(np.float32, (sb.nout, 4))
    rng = np.random.default_rng(kind)
    sel = np.unique(np.concatenate([np.linspace(0, n - 1, 2000).astype(np.int64), rng.integers(0, n, 2000)]))
    tx = sb.texels[sel]
    ref, _ = oracle.gather(sb.gs.oracle_arrays(sb.exitance), sb.params.as_dict(), tx.view(np.uint32).reshape(-1, 4),
                           sb.subs, tree=sb.gs.col.download_tree(), threads=16)
    err = _cmp(flat, ref, tx["out_index"], "lanes")
    print(f"env lanes kind {kind} quality {quality} rounds {rounds}: {n} texels, {len(sel)} sampled, "
          f"max rel err {err:.2e}", flush=True)
    sb.gs.close()


@pytest.mark.parametrize("kinds,angles,blend", [((1, 2), (0.4, -1.3), 0.3), ((2, 0), (0.9, 0.0), 0.6), ((0, 1), (0.0, 2.1), 0.5)])
def test_gpu_environment_blend_vs_oracle(gpu, oracle, kinds, angles, blend):
    """Rotated and blended environments (lsg_scene_set_environment_blend,
    RRBufferBlend.cpp:49-66): a cube and a panorama blended, a rotated panorama alone (a
    blend with nothing), and a rotated cube as the second environment with none first
    (hemisphere gathered because the gather asks for it)."""
    from lightsprintsdk_amd import bake as B
    sc = _scene()
    sb = B.SceneBake(sc, B.GatherParams(quality=48, ray_length_min=S.minimal_safe_distance(sc.verts), seed=4,
                                        environment_multiplier=1.1, hemisphere=1))
    k0, k1 = kinds
    if k0:
        e0 = _env(k0, seed=3)
        sb.gs.set_environment(e0["kind"], e0["width"], e0["height"], e0["rgb"])
    e1 = _env(k1, seed=5) if k1 else None
    sb.gs.set_environment_blend(k1, e1["width"] if e1 else 0, e1["height"] if e1 else 0, e1["rgb"] if e1 else None,
                                angles[0], angles[1], blend)
    sb.first_gather()
    sb.final_gather()
    flat = sb.d_out.download(np.float32, (sb.nout, 4))
    tx = sb.texels
    ref, _ = oracle.gather(sb.gs.oracle_arrays(sb.exitance), sb.params.as_dict(), tx.view(np.uint32).reshape(-1, 4),
                           sb.subs, tree=sb.gs.col.download_tree(), threads=16)
    err = _cmp(flat, ref, tx["out_index"], f"blend {kinds}")
    print(f"env blend {kinds} angles {angles} blend {blend}: max rel err {err:.2e}", flush=True)
