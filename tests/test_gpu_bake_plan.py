"""The whole-scene bake plan on the GPU: persistent XCD-queue final gather, hipGraph
capture/replay and autotuned launch shapes must all produce exactly the texels of
the plain one-wave-per-texel gather + filter chain (which is pinned against the
CPU oracle in test_gpu_gather.py).  With lanes taking shooters as they free up
(lane_refill=2, the default) the rays are the same and only the order of each texel's
sums changes: equal within float rounding, and bit-identical across launch shapes and
graph replays."""
import numpy as np
import pytest

from lightsprintsdk_amd import scenes as S

pytestmark = pytest.mark.gpu


def _scene():
    return S.object_field(seed=5, n_objects=12, target_tris=60_000, lightmap_range=(16, 96), area=40.0)


def test_plan_matches_plain_gather(gpu):
    from lightsprintsdk_amd import bake as B
    sc = _scene()
    params = B.GatherParams(quality=128, ray_length_min=S.minimal_safe_distance(sc.verts), seed=3)
    plan = B.BakePlan(sc, params)
    plan.set_tuning(lane_refill=1)
    plan.run(first=True, graph=False)
    got = plan.rgba()
    ref = B.SceneBake(sc, params)
    ref.first_gather()
    ref.bake()
    want = ref.d_out.download(np.float32, (max(ref.nout, 1), 4))
    assert plan.info().num_texels == len(ref.texels)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    for k in range(len(sc.objects)):
        m = ref.lightmaps_rgbf()[k]
        np.testing.assert_array_equal(plan.lightmap(k).view(np.uint32), m.view(np.uint32))
    # lanes refilled as they free up: same rays, sums in another order
    plan.set_tuning(lane_refill=2)
    plan.run(first=True, graph=False)
    got2 = plan.rgba()
    np.testing.assert_array_equal(got2[:, 3].view(np.uint32), want[:, 3].view(np.uint32))
    np.testing.assert_allclose(got2[:, :3], want[:, :3], rtol=2e-5, atol=1e-7)


def test_plan_tunings_and_graph_replays_identical(gpu):
    from lightsprintsdk_amd import bake as B
    from lightsprintsdk_amd import native as N
    sc = _scene()
    params = B.GatherParams(quality=64, ray_length_min=S.minimal_safe_distance(sc.verts), seed=4)
    plan = B.BakePlan(sc, params)
    plan.run(first=True, graph=False)
    base = plan.rgba().copy()
    for t in (dict(tile=1, waves_per_block=8, blocks_per_cu=2, xcd_queues=1),
              dict(tile=7, waves_per_block=2, blocks_per_cu=8, xcd_queues=8)):
        plan.set_tuning(**t)
        plan.run(first=True, graph=False)
        np.testing.assert_array_equal(plan.rgba().view(np.uint32), base.view(np.uint32))
    # graph: capture once, replay several times with eager work and syncs in between
    for rep in range(3):
        plan.run(first=True, graph=True)
        np.testing.assert_array_equal(plan.rgba().view(np.uint32), base.view(np.uint32), err_msg=f"replay {rep}")
        N.check(N.lib().lsg_synchronize(), "sync")
        plan.run(first=True, graph=False)
    # changing the lights between replays (sun sweep) goes through device memory
    lights = [dict(l) for l in sc.lights]
    lights[0]["direction"] = tuple(np.array([-0.4, 0.1, -1.0]) / np.linalg.norm([-0.4, 0.1, -1.0]))
    plan.gs.set_lights(lights)
    plan.run(first=True, graph=True)
    moved = plan.rgba().copy()
    plan.run(first=True, graph=False)
    np.testing.assert_array_equal(plan.rgba().view(np.uint32), moved.view(np.uint32))
    assert not np.array_equal(moved, base)


def test_autotune_picks_a_candidate(gpu, tmp_path):
    from lightsprintsdk_amd import bake as B
    sc = _scene()
    plan = B.BakePlan(sc, B.GatherParams(quality=64, ray_length_min=S.minimal_safe_distance(sc.verts)))
    plan.run(first=True, graph=False)
    cache = str(tmp_path / "tuning.json")
    entry = B.autotune(plan, sample_texels=4000, cache_path=cache)
    shape = {k: v for k, v in entry["tuning"].items() if k != "lane_refill"}
    assert shape in B.TUNING_CANDIDATES and entry["tuning"].get("lane_refill", 2) in (1, 2)
    assert entry["sample_ms"] <= entry["worst_ms"]
    again = B.autotune(plan, cache_path=cache)
    assert again["tuning"] == entry["tuning"]
