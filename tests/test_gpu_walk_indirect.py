"""GPU parity of the reference's subtexel walk (gather.cpp:645-688) and of the per-corner
indirect term at hemisphere hits (pathtracer.cpp:326-348) against the CPU oracle on the
same seeded rays (relative error <= 1e-4 per channel, alpha bit-exact):

- texels of 1 to 14 subtexels at several qualities: the first series from the walk
  records (up to 8 subtexels) or stepped in the kernel (more), in the grouped, chunked
  and persistent kernels;
- texels whose light rays are mostly unreliable, so the reliability loop shoots more
  series and the walk carries on from the first;
- a box room whose triangles carry a non-constant per-corner indirect field, in the
  plain, layered and persistent (bake plan) gathers."""
import numpy as np
import pytest

from lightsprintsdk_amd import scenes as S

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def _cmp(gpu, ref, sel, what):
    g = gpu[sel].astype(np.float64)
    r = ref[sel]
    err = np.abs(g[:, :3] - r[:, :3]) / np.maximum(np.abs(r[:, :3]), 1e-6)
    assert err.max() <= RTOL, (what, err.max(), np.unravel_index(err.argmax(), err.shape))
    np.testing.assert_array_equal(g[:, 3].astype(np.float32), r[:, 3].astype(np.float32))
    return err.max()


def _synthetic_texels(tris, counts, seed):
    """Texel k: counts[k] subtexels on triangle tris[k], small random triangles in
    triangle space with random areas (the walk weighs by area only)."""
    from lightsprintsdk_amd import bake as B
    rng = np.random.default_rng(seed)
    tex = np.zeros(len(counts), B.TEXEL_DTYPE)
    sub = np.zeros(int(np.sum(counts)), B.SUBTEXEL_DTYPE)
    b = 0
    for k, (t, m) in enumerate(zip(tris, counts)):
        tex[k] = (k, b, m | ((16 + k % 240) << 16), 1000 + k)
        for j in range(m):
            c = rng.uniform(0.05, 0.6, 2)
            d = rng.uniform(-0.04, 0.04, (3, 2))
            uv = np.clip(c + d, 0.0, 0.9)
            sub[b + j] = (t, uv.reshape(-1), rng.uniform(0.05, 1.0) * (1e-3 if rng.random() < 0.15 else 1.0) / 1000)
        b += m
    return tex, sub


def _room_scene(seed=3):
    from lightsprintsdk_amd import bake as B
    sc = S.box_room(seed=seed, n_boxes=4, target_tris=4000, lightmap_size=32)
    sb = B.SceneBake(sc, B.GatherParams(quality=64, ray_length_min=S.minimal_safe_distance(sc.verts), seed=11))
    sb.first_gather()
    return sc, sb


@pytest.mark.parametrize("quality,rounds", [(16, 0), (40, 0), (256, 0), (1024, 0), (300, 700)])
def test_gpu_walk_many_subtexels_vs_oracle(gpu, oracle, quality, rounds):
    from lightsprintsdk_amd import bake as B
    sc, sb = _room_scene()
    rng = np.random.default_rng(quality)
    n = 600 if quality < 1024 else 150
    counts = np.concatenate([np.arange(1, 15), rng.integers(1, 15, n - 14)])
    tris = rng.integers(0, sc.ntri, len(counts))
    tex, sub = _synthetic_texels(tris, counts, seed=quality + rounds)
    p = B.GatherParams(quality=quality, ray_length_min=S.minimal_safe_distance(sc.verts), seed=17, light_rounds=rounds)
    print(f"walk case: quality {quality}, rounds {rounds}, {len(tex)} texels, {len(sub)} subtexels", flush=True)
    gpu_out = sb.gs.gather(tex, sub, p)
    ref, _ = oracle.gather(sb.gs.oracle_arrays(sb.exitance), p.as_dict(), tex.view(np.uint32).reshape(-1, 4), sub,
                           tree=sb.gs.col.download_tree(), threads=16)
    print(f"  max rel err {_cmp(gpu_out, ref, np.arange(len(tex)), 'walk'):.2e}", flush=True)
    sb.gs.close()


def test_gpu_walk_carries_over_series_vs_oracle(gpu, oracle):
    """A floor mostly under an occluder 0.5 mm above it (inside rugDistance): most light
    rays are unreliable, the reliability loop shoots further light-only series
    (gather.cpp:652-661), and each series' subtexels continue the walk."""
    from lightsprintsdk_amd import bake as B
    v = np.float32([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                    [-1, -1, 0.0005], [0.8, -1, 0.0005], [0.8, 2, 0.0005], [-1, 2, 0.0005],
                    [-3, -3, 2.0], [4, -3, 2.0], [4, 4, 2.0], [-3, 4, 2.0]])
    i = np.uint32([[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7], [8, 10, 9], [8, 11, 10]])
    nrm = np.float32([[0, 0, 1]] * 8 + [[0, 0, -1]] * 4)
    side = np.uint8([5, 5, 5, 5, 5, 5])
    ex = np.zeros((6, 8), np.float32)
    ex[4:, 0:3] = [0.5, 0.75, 1.0]              # the ceiling emits
    lights = [dict(type=0, position=(0.5, 0.5, 1.5), color=(4.0, 3.0, 2.0))]
    gs = B.GatherScene(v, i, nrm, side, ex, lights)
    tex = np.zeros(40, B.TEXEL_DTYPE)
    sub = np.zeros(40 * 3, B.SUBTEXEL_DTYPE)
    rng = np.random.default_rng(4)
    for k in range(40):
        tex[k] = (k, 3 * k, 3 | (255 << 16), 77 + k)
        # triangle 0 maps (ux, uy) -> (ux + uy, uy): the first two subtexels under the
        # occluder (x < 0.8), the third in the open strip x > 0.8, with a small area
        under = rng.uniform(0.2, 0.35, 2)
        sub[3 * k] = (0, [under[0], 0.1, under[0] + 0.05, 0.1, under[0], 0.15], 0.5)
        sub[3 * k + 1] = (0, [under[1], 0.2, under[1] + 0.05, 0.2, under[1], 0.25], 0.45)
        sub[3 * k + 2] = (0, [0.86, 0.05, 0.9, 0.05, 0.86, 0.08], rng.uniform(0.01, 0.06))
    for quality in (64, 200):
        p = B.GatherParams(quality=quality, ray_length_min=1e-6, seed=23)
        gpu_out = gs.gather(tex, sub, p)
        ref, st = oracle.gather(gs.oracle_arrays(ex), p.as_dict(), tex.view(np.uint32).reshape(-1, 4), sub,
                                tree=gs.col.download_tree(), threads=16)
        assert st.series > 2 * len(tex), st.series       # further series were shot
        print(f"series case: quality {quality}, {st.series / len(tex):.1f} series per texel, "
              f"max rel err {_cmp(gpu_out, ref, np.arange(len(tex)), 'series'):.2e}", flush=True)
    gs.close()


def _indirect_field(sc, seed):
    """A smooth field plus per-corner noise: corners of one vertex differ between
    triangles, as the solver's smoothing groups allow."""
    rng = np.random.default_rng(seed)
    p = sc.verts[sc.idx].astype(np.float64)                      # [nt, 3, 3]
    base = 0.3 + 0.05 * p[..., 0:1] + 0.08 * np.sin(p[..., 1:2]) + 0.1 * p[..., 2:3]
    corners = np.abs(base * np.float64([1.0, 0.8, 0.6]) + rng.normal(0, 0.02, size=p.shape)).astype(np.float32)
    albedo = rng.uniform(0.2, 0.8, size=(sc.ntri, 3)).astype(np.float32)
    sides = rng.integers(0, 4, sc.ntri).astype(np.uint8)
    return corners, albedo, sides


def test_gpu_indirect_vs_oracle(gpu, oracle):
    sc, sb = _room_scene(seed=7)
    corners, albedo, sides = _indirect_field(sc, 1)
    sb.gs.set_indirect(corners, albedo, sides)
    sb.params.quality = 256
    sb.final_gather()
    flat = sb.d_out.download(np.float32, (sb.nout, 4))
    tx = sb.texels
    ref, _ = oracle.gather(sb.gs.oracle_arrays(sb.exitance), sb.params.as_dict(), tx.view(np.uint32).reshape(-1, 4),
                           sb.subs, tree=sb.gs.col.download_tree(), threads=16)
    print(f"indirect: {len(tx)} texels, max rel err {_cmp(flat, ref, tx['out_index'], 'indirect'):.2e}", flush=True)
    # the term matters: without it the same gather differs
    sb.gs.set_indirect(None)
    sb.clear()
    sb.final_gather()
    plain = sb.d_out.download(np.float32, (sb.nout, 4))
    assert np.abs(plain[tx["out_index"], :3] - flat[tx["out_index"], :3]).max() > 1e-3
    sb.gs.close()


def test_gpu_indirect_layers_vs_oracle(gpu, oracle):
    sc, sb = _room_scene(seed=8)
    corners, albedo, sides = _indirect_field(sc, 2)
    sb.gs.set_indirect(corners, albedo, sides)
    from lightsprintsdk_amd import bake as B
    p = B.GatherParams(quality=128, ray_length_min=S.minimal_safe_distance(sc.verts), seed=31)
    sel = np.linspace(0, len(sb.texels) - 1, 400).astype(np.int64)
    tx = sb.texels[sel]
    got = sb.gs.gather_layers(tx, sb.subs, p, nout=sb.nout)
    ref = oracle.gather_layers(sb.gs.oracle_arrays(sb.exitance), p.as_dict(), tx.view(np.uint32).reshape(-1, 4), sb.subs,
                               tree=sb.gs.col.download_tree(), threads=16)
    for k in range(4):
        _cmp(got[k], ref[k], tx["out_index"], f"layer {k}")
    sb.gs.close()


def test_gpu_indirect_bake_plan_equals_plain(gpu):
    """The persistent bake kernels (lanes refilled as they free up) read the same term:
    a plan's final gather equals the plain gather of the same texels within float
    summation order."""
    from lightsprintsdk_amd import bake as B
    sc = S.object_field(seed=5, n_objects=8, target_tris=30_000, lightmap_range=(16, 64), area=30.0)
    params = B.GatherParams(quality=256, ray_length_min=S.minimal_safe_distance(sc.verts), seed=13)
    corners, albedo, sides = _indirect_field(sc, 3)
    plan = B.BakePlan(sc, params)
    plan.gs.set_indirect(corners, albedo, sides)
    plan.run(first=True, graph=False, filtering=False)
    got = plan.rgba()
    ref = B.SceneBake(sc, params)
    ref.gs.set_indirect(corners, albedo, sides)
    ref.first_gather()
    ref.final_gather()
    want = ref.d_out.download(np.float32, (max(ref.nout, 1), 4))
    sel = ref.texels["out_index"]
    err = np.abs(got[sel, :3].astype(np.float64) - want[sel, :3]) / np.maximum(np.abs(want[sel, :3]), 1e-6)
    print(f"plan vs plain with indirect: {len(sel)} texels, max rel err {err.max():.2e}", flush=True)
    assert err.max() < 2e-5
    np.testing.assert_array_equal(got[sel, 3].view(np.uint32), want[sel, 3].view(np.uint32))
    plan.close()
