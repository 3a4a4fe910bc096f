"""Wrapped filtering and unwrap-seam smoothing on the GPU, bit for bit against the oracle.

The filter plan (lsg_filter_plan_*) runs FilteringParameters::wrap on maps whose texel
counts do and do not divide 2^32 (lightmapGrow's wrapped neighbours differ there), and the
seam pass on random seams: long and short edges, edges partly outside [0,1)^2, chains
sharing texels (a seam reads what the seams before it wrote) and edges that cross their
own partner's texels (taken one point at a time on the GPU).  The bake plan smooths the
seams set per object (lsg_bake_set_seams), also through a replayed graph."""
import ctypes

import numpy as np
import pytest

from lightsprintsdk_amd import native as N
from lightsprintsdk_amd.collider import DeviceBuffer

pytestmark = pytest.mark.gpu

SHAPES = [(3, 3), (5, 7), (8, 8), (1, 9), (9, 1), (64, 48), (33, 17), (16, 16)]


class _FilterPlan:
    def __init__(self, shapes):
        self.shapes = shapes
        self.maps = (N.Lightmap * len(shapes))()
        base = 0
        for k, (w, h) in enumerate(shapes):
            self.maps[k] = N.Lightmap(base, w, h, base)
            base += w * h
        self.n = base
        h = ctypes.c_void_p()
        N.check(N.lib().lsg_filter_plan_create(self.maps, len(shapes), ctypes.byref(h)), "lsg_filter_plan_create")
        self.h = h.value
        self.d_rgba = DeviceBuffer(16 * self.n)
        self.d_rgbf = DeviceBuffer(12 * self.n)

    def set_seams(self, seams):
        if seams is None:
            N.check(N.lib().lsg_filter_plan_set_seams(self.h, None, None), "set_seams")
            return
        arrs = [np.ascontiguousarray(s, np.float32).reshape(-1, 8) for s in seams]
        begin = np.cumsum([0] + [len(a) for a in arrs]).astype(np.uint32)
        allv = np.ascontiguousarray(np.concatenate(arrs))
        N.check(N.lib().lsg_filter_plan_set_seams(self.h, allv.ctypes.data_as(ctypes.c_void_p),
                                                  begin.ctypes.data_as(ctypes.c_void_p)), "set_seams")

    def run(self, imgs, sigma, spread, wrap, bg=(0.0, 0.0, 0.0, 0.0)):
        self.d_rgba.upload(np.concatenate([m.reshape(-1, 4) for m in imgs]).astype(np.float32))
        fp = N.FilterParams(sigma, spread, (ctypes.c_float * 4)(*bg), 1 if wrap else 0)
        N.check(N.lib().lsg_filter_plan_run(self.h, self.d_rgba.ptr, self.d_rgbf.ptr, ctypes.byref(fp), None), "plan run")
        rgba = self.d_rgba.download(np.float32, (self.n, 4))
        rgbf = self.d_rgbf.download(np.float32, (self.n, 3))
        out, base = [], 0
        for w, h in self.shapes:
            out.append((rgba[base:base + w * h].reshape(h, w, 4), rgbf[base:base + w * h].reshape(h, w, 3)))
            base += w * h
        return out

    def close(self):
        N.lib().lsg_filter_plan_destroy(self.h)


def _random_map(rng, w, h, fill=0.7):
    m = np.zeros((h, w, 4), np.float32)
    used = rng.random((h, w)) < fill
    m[..., :3] = rng.random((h, w, 3), dtype=np.float32)
    m[..., 3] = np.where(used, rng.integers(1, 256, (h, w)) / np.float32(255), 0).astype(np.float32)
    m[~used, :3] = 0
    return m


def _random_seams(rng, w, h, n=24):
    """Seams of every kind the walk meets, in texel-centred uv of a w x h map."""
    out = []
    for k in range(n):
        kind = k % 6
        if kind == 0:          # anywhere, partly outside [0,1)^2
            e = rng.uniform(-0.2, 1.2, 8)
        elif kind == 1:        # inside one texel each
            tx, ty = rng.integers(0, w, 2), rng.integers(0, h, 2)
            e = np.array([(tx[0] + 0.2) / w, (ty[0] + 0.5) / h, (tx[0] + 0.8) / w, (ty[0] + 0.5) / h,
                          (tx[1] + 0.2) / w, (ty[1] + 0.5) / h, (tx[1] + 0.8) / w, (ty[1] + 0.5) / h])
        elif kind == 2:        # the partner edge is the same edge reversed (runs share texels)
            a = rng.uniform(0, 1, 4)
            e = np.concatenate([a, [a[2], a[3], a[0], a[1]]])
        elif kind == 3:        # the partner edge half a texel aside (crossing texels)
            a = rng.uniform(0.05, 0.9, 4)
            e = np.concatenate([a, a + np.array([0.5 / w, 0.0, 0.5 / w, 0.0])])
        elif kind == 4 and out:  # continues the previous seam (shares its end texel)
            p = out[-1]
            e = np.concatenate([[p[2], p[3]], rng.uniform(0, 1, 2), [p[6], p[7]], rng.uniform(0, 1, 2)])
        else:                  # long parallel pair
            x1, x2 = rng.uniform(0, 1, 2)
            e = np.array([x1, 0.01, x1, 0.99, x2, 0.01, x2, 0.99])
        out.append(np.float32(e))
    return np.stack(out)


def _check(plan, imgs, oracle, sigma, spread, wrap, seams=None, tag=""):
    got = plan.run(imgs, sigma, spread, wrap)
    for k, m in enumerate(imgs):
        ref_rgba, ref_rgbf = oracle.filter_lightmap_ex(m, sigma=sigma, spread=spread, wrap=wrap,
                                                       seams=None if seams is None else seams[k])
        np.testing.assert_array_equal(got[k][0].view(np.uint32), ref_rgba.view(np.uint32), err_msg=f"{tag} map {k} rgba")
        np.testing.assert_array_equal(got[k][1].view(np.uint32), ref_rgbf.view(np.uint32), err_msg=f"{tag} map {k} rgbf")


@pytest.mark.parametrize("sigma,spread", [(0.0, 1.0), (0.0, 5.0), (1.0, 1000.0), (2.5, 3.0)])
def test_gpu_wrap_filters_vs_oracle(gpu, oracle, sigma, spread):
    rng = np.random.default_rng(int(sigma * 10 + spread))
    plan = _FilterPlan(SHAPES)
    try:
        imgs = [_random_map(rng, w, h, fill=0.3) for w, h in SHAPES]
        print(f"wrap sigma {sigma} spread {spread}", flush=True)
        _check(plan, imgs, oracle, sigma, spread, True, tag="wrap")
        _check(plan, imgs, oracle, sigma, spread, False, tag="clamp")
    finally:
        plan.close()


@pytest.mark.parametrize("wrap", [False, True])
def test_gpu_seams_vs_oracle(gpu, oracle, wrap):
    rng = np.random.default_rng(7 + wrap)
    plan = _FilterPlan(SHAPES)
    try:
        imgs = [_random_map(rng, w, h, fill=0.9) for w, h in SHAPES]
        seams = [_random_seams(rng, w, h, n=6 + 6 * k) for k, (w, h) in enumerate(SHAPES)]
        seams[3] = np.zeros((0, 8), np.float32)                 # a map without seams among maps with them
        plan.set_seams(seams)
        print(f"seams wrap {wrap}: {sum(len(s) for s in seams)} seams", flush=True)
        _check(plan, imgs, oracle, 1.0, 1000.0, wrap, seams, tag="seams")
        _check(plan, imgs, oracle, 2.0, 2.0, wrap, seams, tag="seams s2")
        plan.set_seams(None)
        _check(plan, imgs, oracle, 1.0, 1000.0, wrap, None, tag="seams removed")
    finally:
        plan.close()


def test_gpu_seams_many_per_map(gpu, oracle):
    """A 256 x 256 map with 2,000 seams (the seams of one map run one after another)."""
    rng = np.random.default_rng(3)
    plan = _FilterPlan([(256, 256)])
    try:
        imgs = [_random_map(rng, 256, 256, fill=0.95)]
        seams = [_random_seams(rng, 256, 256, n=2000)]
        plan.set_seams(seams)
        _check(plan, imgs, oracle, 1.0, 10.0, False, seams, tag="many")
    finally:
        plan.close()


def test_gpu_bake_plan_seams(gpu, oracle):
    """lsg_bake_set_seams: the plan's filtered lightmaps equal the oracle's filter chain
    with each object's seams over the plan's unfiltered gather, eagerly and through a
    graph replayed before and after the seams change."""
    from lightsprintsdk_amd import bake as B
    from lightsprintsdk_amd import scenes as S
    sc = S.box_room(seed=3, n_boxes=3, target_tris=3000, lightmap_size=48)
    plan = B.BakePlan(sc, B.GatherParams(quality=32, ray_length_min=S.minimal_safe_distance(sc.verts)))
    rng = np.random.default_rng(11)
    try:
        plan.run(filtering=False)
        raw = [plan.lightmap_layer_rgba(k) for k in range(len(sc.objects))]
        seams = [_random_seams(rng, ob.width, ob.height, n=12) if k % 2 == 0 else np.zeros((0, 8), np.float32)
                 for k, ob in enumerate(sc.objects)]
        for k, s in enumerate(seams):
            plan.set_seams(k, s)

        def check(tag):
            for k in range(len(sc.objects)):
                _, ref = oracle.filter_lightmap_ex(raw[k], sigma=1.0, spread=1000.0, seams=seams[k])
                np.testing.assert_array_equal(plan.lightmap(k).view(np.uint32), ref.view(np.uint32), err_msg=f"{tag} object {k}")
        plan.run()
        check("eager")
        plan.run(graph=True)
        plan.run(graph=True)
        check("graph")
        seams[0] = _random_seams(rng, sc.objects[0].width, sc.objects[0].height, n=5)
        plan.set_seams(0, seams[0])
        plan.run(graph=True)
        check("graph after a change")
    finally:
        plan.close()
