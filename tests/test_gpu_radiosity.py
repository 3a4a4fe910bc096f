"""Radiosity on the GPU (lsg_bake_set_bounces): Jacobi iterations of the per-triangle
gather spread light between triangles before the final gather.

- Closed form: inside a closed box whose walls all emit E and reflect albedo rho, every
  ray sees the same exitance, so each bounce gives B_k = E + rho * B_(k-1) exactly:
  E * (1 + rho + ... + rho^k), and the final gather returns B_K.
- Parity: the plan's exitance table after two bounces on a box room equals the same
  sequence driven through the CPU oracle (first gather -> light part; each bounce a
  hemisphere-only gather over the previous table) to 1e-4.
"""
import numpy as np
import pytest

from lightsprintsdk_amd import scenes as S

pytestmark = pytest.mark.gpu


def test_gpu_radiosity_closed_box_geometric_series(gpu):
    from lightsprintsdk_amd import bake as B
    sc = S.box_room(seed=0, n_boxes=0, target_tris=400, lightmap_size=16, room=(2.0, 2.0, 2.0))
    sc.emittance[:] = 1.0
    sc.albedo[:] = 0.5
    sc.lights = []
    plan = B.BakePlan(sc, B.GatherParams(quality=64, ray_length_min=S.minimal_safe_distance(sc.verts), seed=2))
    for k, want in [(0, 1.0), (1, 1.5), (3, 1.875)]:
        plan.set_bounces(k, 128)
        plan.run(first=True, graph=False)
        ex = plan.exitance()
        np.testing.assert_array_equal(ex[:, 0:3], np.float32(want), err_msg=f"{k} bounces")
    # the final gather sees the converged walls: every covered texel gathers B_3
    rgba = plan.rgba()
    cov = rgba[:, 3] > 0
    assert cov.sum() > 100
    np.testing.assert_allclose(rgba[cov, :3], 1.875, rtol=1e-6)
    # graph replays start from emittance each time: the same table again
    for _ in range(2):
        plan.run(first=True, graph=True)
        np.testing.assert_array_equal(plan.exitance()[:, 0:3], np.float32(1.875))
    plan.close()


def test_gpu_radiosity_vs_oracle(gpu, oracle):
    from lightsprintsdk_amd import bake as B
    sc = S.box_room(seed=5, n_boxes=3, target_tris=3000, lightmap_size=16)
    params = B.GatherParams(quality=64, ray_length_min=S.minimal_safe_distance(sc.verts), seed=4)
    plan = B.BakePlan(sc, params)
    bounces, bq = 2, 48
    plan.set_bounces(bounces, bq)
    plan.run(first=True, graph=False, filtering=False)
    got = plan.exitance()
    # the same sequence through the oracle
    nt = sc.ntri
    tex, sub = B.triangle_texels(nt)
    t4 = tex.view(np.uint32).reshape(-1, 4)
    tree = plan.gs.col.download_tree()
    alb = np.asarray(sc.albedo, np.float32)
    emi = np.asarray(sc.emittance, np.float32)
    ex = np.zeros((nt, 8), np.float32)
    ex[:, 0:3] = emi
    fp = plan.first_params.as_dict()
    light, _ = oracle.gather(plan.gs.oracle_arrays(ex), fp, t4, sub, tree=tree, threads=16)
    pl = light[:nt, 0:3].astype(np.float32)
    ex[:, 0:3] = (np.float32(0) + pl) * alb + emi
    for b in range(bounces):
        bp = dict(fp, hemisphere=1, light_multiplier=0.0, light_rounds=0, quality=bq,
                  seed=(fp["seed"] ^ ((0x9E3779B9 * (b + 1)) & 0xFFFFFFFF)) & 0xFFFFFFFF)
        hemi, _ = oracle.gather(plan.gs.oracle_arrays(ex), bp, t4, sub, tree=tree, threads=16)
        ex = ex.copy()
        ex[:, 0:3] = (hemi[:nt, 0:3].astype(np.float32) + pl) * alb + emi
    err = np.abs(got[:, 0:3] - ex[:, 0:3]) / np.maximum(np.abs(ex[:, 0:3]), 1e-3)
    print(f"radiosity {bounces} bounces: {nt} triangles, max rel err {err.max():.2e}", flush=True)
    assert err.max() <= 1e-4
    # bounces add light: indirect exitance beyond the direct table
    assert (got[:, 0:3] - ((pl * alb) + emi)).mean() > 1e-3
    plan.close()
