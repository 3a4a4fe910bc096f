"""The two large bake workloads at their full size on the GPU.

Batched bake (200 objects, about 1.8M triangles, lightmaps 32..1024, 1024 rays per texel,
sun + 4 point lights): the persistent plan's gathered texels equal the plain gather's bit
for bit, a spread sample of 300 texels equals the CPU oracle (relative error <= 1e-4 per
channel, alpha bit-exact), and after the autotuner picks a launch shape the hipGraph
replays equal the eager run bit for bit; every lightmap is finite and non-negative.

Large re-bake (city of 5000 boxes on a 2048^2-vertex terrain, 10.0M triangles, 64
lightmaps of 2048^2): one sun step of the learned+fallback gather stays within 2% RMSE of
the plain 1024-ray re-bake of the same sun, and a spread sample of 200 texels of the plain
gather equals the oracle."""
import numpy as np
import pytest

from lightsprintsdk_amd import scenes as S

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def _local(tex, sub, sel):
    """The selected texels with their subtexels renumbered into a compact list."""
    t = tex[sel].copy()
    n = (t["sub_count_flags"] & 0xFFFF).astype(np.int64)
    idx = np.concatenate([np.arange(b, b + c) for b, c in zip(t["sub_begin"].astype(np.int64), n)])
    t["sub_begin"] = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.uint32)
    orig = t["out_index"].copy()
    t["out_index"] = np.arange(len(t), dtype=np.uint32)
    return t, sub[idx], orig


def _vs_oracle(oracle, gs, exitance, params, tex, sub, got_rgba, n, label):
    sel = np.linspace(0, len(tex) - 1, n).astype(np.int64)
    t, s, orig = _local(tex, sub, sel)
    ref, _ = oracle.gather(gs.oracle_arrays(exitance), params.as_dict(), t.view(np.uint32).reshape(-1, 4), s,
                           tree=gs.col.download_tree(), threads=16)
    g = got_rgba[orig].astype(np.float64)
    err = np.abs(g[:, :3] - ref[:, :3]) / np.maximum(np.abs(ref[:, :3]), 1e-6)
    print(f"{label}: {n} texels vs oracle, max rel err {err.max():.2e}", flush=True)
    assert err.max() <= RTOL, (label, err.max())
    np.testing.assert_array_equal(g[:, 3].astype(np.float32), ref[:, 3].astype(np.float32))


def test_gpu_batched_bake_full_size(gpu, oracle, tmp_path):
    from lightsprintsdk_amd import bake as B
    sc = S.object_field(seed=4, n_objects=200, target_tris=2_000_000)
    params = B.GatherParams(quality=1024, ray_length_min=S.minimal_safe_distance(sc.verts), seed=1)
    print(f"batched bake: {sc.ntri} triangles, {len(sc.objects)} objects, {len(sc.lights)} lights", flush=True)
    assert len(sc.objects) == 200 and 1_500_000 <= sc.ntri <= 2_200_000 and len(sc.lights) == 5
    sizes = [ob.width for ob in sc.objects]
    assert min(sizes) >= 32 and max(sizes) <= 1024
    plan = B.BakePlan(sc, params)
    plan.set_tuning(lane_refill=1)                      # the plain gather's shooter order
    plan.run(first=True, graph=False, filtering=False)
    raw = plan.rgba()
    ex = plan.exitance()
    inf = plan.info()
    print(f"  {inf.num_texels} texels, {inf.num_subtexels} subtexels", flush=True)
    # the plain one-wave-per-texel gather of the same texels over the same exitance
    sb = B.SceneBake(sc, params)
    assert len(sb.texels) == inf.num_texels
    sb.gs.set_exitance(ex)
    sb.final_gather()
    plain = sb.d_out.download(np.float32, (max(sb.nout, 1), 4))
    np.testing.assert_array_equal(raw[sb.texels["out_index"]].view(np.uint32), plain[sb.texels["out_index"]].view(np.uint32))
    _vs_oracle(oracle, plan.gs, ex, params, sb.texels, sb.subs, raw, 300, "batched bake")
    del sb, plain
    # autotuned shape, eager run, then graph replays with eager work in between
    entry = B.autotune(plan, sample_texels=150_000, cache_path=str(tmp_path / "tuning.json"))
    print(f"  autotune: {entry['tuning']}", flush=True)
    plan.run(first=True, graph=False)
    eager = plan.rgba().copy()
    for rep in range(2):
        plan.run(first=True, graph=True)
        np.testing.assert_array_equal(plan.rgba().view(np.uint32), eager.view(np.uint32), err_msg=f"replay {rep}")
    cover = eager[raw[:, 3] > 0]
    assert np.isfinite(eager).all() and (cover[:, :3] >= 0).all()
    for k in range(0, len(sc.objects), 37):
        m = plan.lightmap(k)
        assert np.isfinite(m).all() and (m >= 0).all() and m.max() > 0
    plan.close()


def test_gpu_large_rebake_full_size_sun_step(gpu, oracle):
    from lightsprintsdk_amd import learned as Lr
    got = Lr.load_weights(Lr.weights_path("city"))
    assert got is not None, "the trained city surrogate ships with the package"
    net, bound, _ = got
    sc = S.city_terrain(seed=5, sun_direction=S.sun_sweep(0, 60))
    print(f"large re-bake: {sc.ntri} triangles, {len(sc.objects)} lightmaps", flush=True)
    assert sc.ntri >= 9_900_000 and len(sc.objects) == 64
    assert all(ob.width == 2048 and ob.height == 2048 for ob in sc.objects)
    ls = Lr.LearnedScene(sc, seed=1)
    sb = ls.sb
    print(f"  {ls.n} texels", flush=True)
    # plain re-bake of this sun: gathered texels (oracle sample), then the filters
    sb.first_gather()
    ex = sb.exitance
    ls.plain(filtering=False)
    raw = sb.d_out.download(np.float32, (max(sb.nout, 1), 4))
    _vs_oracle(oracle, sb.gs, ex, sb.params, sb.texels, sb.subs, raw, 200, "large re-bake, plain gather")
    del raw
    sb.filter()
    ref = ls.final_rgb()
    # the learned step for the same sun
    sb.first_gather()
    nfb = ls.learned(net, target=bound["target"])
    out = ls.final_rgb()
    rmse = Lr.rmse_rel(out, ref)
    print(f"  learned+fallback: rmse {rmse:.4f}, fallback {nfb / ls.n:.3f}", flush=True)
    assert np.isfinite(out).all()
    assert rmse <= 0.02
