"""GPU parity of the output layers of updateLightmap: the lightmap, the three
directional lightmaps and the bent normal (RRSolver/gather.cpp:284-306, 453-482,
809-830), gathered by liblsgpu's layered kernels and by the CPU oracle on the same
seeded rays.  Irradiance layers: relative error <= 1e-4 per channel; bent normals:
<= 1e-4 absolute on the unit vector (its small components are sums that cancel);
alphas bit-exact."""
import numpy as np
import pytest

from lightsprintsdk_amd import scenes as S

pytestmark = pytest.mark.gpu

RTOL = 1e-4
BENT_ATOL = 1e-4


def _cmp_layer(k, gpu, ref, sel):
    g = gpu[sel].astype(np.float64)
    r = ref[sel]
    np.testing.assert_array_equal(g[:, 3].astype(np.float32), r[:, 3].astype(np.float32))
    if k == 4:
        err = np.abs((2 * g[:, :3] - 1) - (2 * r[:, :3] - 1))
        assert err.max() <= BENT_ATOL, (k, err.max())
        return err.max()
    err = np.abs(g[:, :3] - r[:, :3]) / np.maximum(np.abs(r[:, :3]), 1e-6)
    assert err.max() <= RTOL, (k, err.max(), np.unravel_index(err.argmax(), err.shape))
    return err.max()


def _bake(seed=3, quality=64, rounds=0, tris=3000, size=24):
    from lightsprintsdk_amd import bake as B
    sc = S.box_room(seed=seed, n_boxes=4, target_tris=tris, lightmap_size=size)
    sb = B.SceneBake(sc, B.GatherParams(quality=quality, ray_length_min=S.minimal_safe_distance(sc.verts), seed=11,
                                        light_rounds=rounds))
    sb.first_gather()
    return sc, sb


def _random_frames(normals, seed):
    """Per-vertex orthonormal tangent frames, randomly turned about each normal."""
    rng = np.random.default_rng(seed)
    n = normals.astype(np.float64)
    a = np.where(np.abs(n[:, :1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(n, a)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(n, t)
    phi = rng.uniform(0, 2 * np.pi, (len(n), 1))
    t2 = np.cos(phi) * t + np.sin(phi) * b
    b2 = np.cross(n, t2)
    return t2.astype(np.float32), b2.astype(np.float32)


@pytest.mark.parametrize("quality,rounds,layers,frames", [
    (64, 0, (0, 1, 2, 3, 4), False),     # whole-wave kernel, directions on
    (64, 0, (0, 4), False),              # directions off: lightmap + bent normal
    (16, 0, (1, 2, 3, 4), False),        # 16-lane groups
    (24, 6, (0, 1, 2, 3, 4), True),      # 32-lane groups, caller tangent frames
])
def test_gpu_layers_vs_oracle(gpu, oracle, quality, rounds, layers, frames):
    sc, sb = _bake(quality=quality, rounds=rounds)
    if frames:
        t, b = _random_frames(sc.normals, seed=quality)
        sb.gs.set_tangents(t, b)
    tx = sb.texels
    gpu_out = sb.gs.gather_layers(tx, sb.subs, sb.params, layers=layers, nout=sb.nout)
    ref = oracle.gather_layers(sb.gs.oracle_arrays(sb.exitance), sb.params.as_dict(), tx.view(np.uint32).reshape(-1, 4),
                               sb.subs, layers=layers, tree=sb.gs.col.download_tree(), threads=16)
    errs = {k: _cmp_layer(k, gpu_out[k], ref[k], tx["out_index"]) for k in layers}
    print(f"quality {quality} rounds {rounds} layers {layers} frames {frames}: {len(tx)} texels, max err "
          + ", ".join(f"{k}:{e:.1e}" for k, e in errs.items()), flush=True)
    # directions on: each directional layer differs from the lightmap somewhere
    if 1 in layers and 0 in layers:
        assert not np.allclose(gpu_out[1][tx["out_index"], :3], gpu_out[0][tx["out_index"], :3], rtol=1e-3)


def test_gpu_layers_lightmap_equals_plain_gather(gpu):
    """Without directional layers the layered kernel's lightmap plane is bit-identical
    to the plain gather's (the same sums in the same order)."""
    sc, sb = _bake(quality=64)
    tx = sb.texels
    plain = sb.gs.gather(tx, sb.subs, sb.params, nout=sb.nout)
    lay = sb.gs.gather_layers(tx, sb.subs, sb.params, layers=(0, 4), nout=sb.nout)
    np.testing.assert_array_equal(lay[0], plain)


def test_gpu_bake_layers_plan_equals_scene_bake(gpu, oracle):
    """The bake plan's layered run (first gather, layered final gather, filter chain
    per layer) gives the same RGBF layers as the SceneBake steps, and each layer's
    RGBF equals the oracle's filter chain applied to that layer's gathered RGBA."""
    from lightsprintsdk_amd import bake as B
    sc = S.box_room(seed=6, n_boxes=3, target_tris=2000, lightmap_size=16)
    params = B.GatherParams(quality=48, ray_length_min=S.minimal_safe_distance(sc.verts), seed=5)
    sb = B.SceneBake(sc, params)
    sb.first_gather()
    layers = (0, 1, 2, 3, 4)
    got = sb.bake_layers(layers)
    plan = B.BakePlan(sc, params)
    plan.run_layers(layers)
    for k in layers:
        for o in range(len(sc.objects)):
            np.testing.assert_array_equal(plan.lightmap_layer(o, k), got[k][o])
    # filter chain of one layer against the oracle's
    raw = sb.gs.gather_layers(sb.texels, sb.subs, sb.params, layers=(2,), nout=sb.nout)[2]
    base, w, h = sb.slots[0]
    rgba, rgbf = oracle.filter_lightmap(raw[base:base + w * h].reshape(h, w, 4))
    np.testing.assert_array_equal(got[2][0], rgbf.astype(np.float32))
    # the RGBA work buffer the drop-in solver writes back: the oracle's, alpha included
    np.testing.assert_array_equal(plan.lightmap_layer_rgba(0, 2), rgba.astype(np.float32))
    for k in layers:
        for o in range(len(sc.objects)):
            np.testing.assert_array_equal(plan.lightmap_layer_rgba(o, k)[..., :3], got[k][o])
    plan.close()
