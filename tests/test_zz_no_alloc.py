"""The allocation-free filter chain (lsg_filter_lightmaps_noalloc) through the C ABI:
bit-identical to lsg_filter_lightmaps on two calls in a row, the second with a smaller
batch (the static scratch keeps the first call's contents), with and without wrap (maps
whose texel counts do and do not divide 2^32), and a clean LSG_ERR_LIMIT beyond its
stated maximum."""
import ctypes

import numpy as np
import pytest

from lightsprintsdk_amd import native as N

pytestmark = pytest.mark.gpu


def _maps(shapes, seed):
    """Synthetic RGBA temp lightmaps as a gather leaves them: covered texels carry random
    texelFlags/255 in alpha, empty ones 0; one sparse map makes lightmapGrow run for
    hundreds of steps.  rgba slots leave gaps between maps, rgbf slots run in reverse."""
    rng = np.random.default_rng(seed)
    imgs, base, lm = [], 0, []
    out_bases = np.cumsum([0] + [w * h for w, h in shapes[::-1]])[:-1][::-1]
    for k, (w, h) in enumerate(shapes):
        img = np.zeros((h, w, 4), np.float32)
        if k == 0:
            cov = np.zeros((h, w), bool)
            cov[3:9, 4:12] = True
        else:
            cov = rng.random((h, w)) < 0.7
        img[..., :3] = np.where(cov[..., None], rng.uniform(0.0, 5.0, size=(h, w, 3)), 0)
        img[..., 3] = np.where(cov, rng.integers(1, 256, size=(h, w)) / 255.0, 0)
        imgs.append(img)
        lm.append((base, w, h, int(out_bases[k])))
        base += w * h + 37
    rgba = np.zeros((base, 4), np.float32)
    for (b, w, h, _), img in zip(lm, imgs):
        rgba[b:b + w * h] = img.reshape(-1, 4)
    arr = (N.Lightmap * len(lm))(*[N.Lightmap(*m) for m in lm])
    return rgba, arr, sum(w * h for w, h in shapes)


def _run(fn, rgba, maps, nout, params):
    L = N.lib()
    d_rgba, d_rgbf = ctypes.c_void_p(), ctypes.c_void_p()
    N.check(L.lsg_malloc(ctypes.byref(d_rgba), rgba.nbytes))
    N.check(L.lsg_malloc(ctypes.byref(d_rgbf), 12 * nout))
    try:
        N.check(L.lsg_memcpy_h2d(d_rgba, rgba.ctypes.data_as(ctypes.c_void_p), rgba.nbytes))
        N.check(L.lsg_memset(d_rgbf, 0, 12 * nout))
        N.check(getattr(L, fn)(maps, len(maps), d_rgba, d_rgbf, ctypes.byref(params), None), fn)
        out_rgba = np.empty_like(rgba)
        out_rgbf = np.empty((nout, 3), np.float32)
        N.check(L.lsg_memcpy_d2h(out_rgba.ctypes.data_as(ctypes.c_void_p), d_rgba, rgba.nbytes))
        N.check(L.lsg_memcpy_d2h(out_rgbf.ctypes.data_as(ctypes.c_void_p), d_rgbf, 12 * nout))
    finally:
        N.check(L.lsg_free(d_rgba))
        N.check(L.lsg_free(d_rgbf))
    return out_rgba, out_rgbf


@pytest.mark.parametrize("wrap", [0, 1])
def test_noalloc_filters_equal_the_allocating_entry_point(gpu, wrap):
    params = N.FilterParams(1.0, 1000.0, (ctypes.c_float * 4)(0.25, 0.5, 0.125, 0.0), wrap)
    cases = [([(256, 256), (64, 64), (128, 32), (100, 37), (17, 9)], 11),   # 84,609 texels
             ([(48, 40), (33, 21)], 12)]                                      # smaller second call
    print()
    for shapes, seed in cases:
        rgba, maps, nout = _maps(shapes, seed)
        print(f"noalloc filter case: {len(shapes)} maps, {nout} texels, seed {seed}, wrap {wrap}", flush=True)
        ref = _run("lsg_filter_lightmaps", rgba, maps, nout, params)
        got = _run("lsg_filter_lightmaps_noalloc", rgba, maps, nout, params)
        np.testing.assert_array_equal(got[0].view(np.uint32), ref[0].view(np.uint32))
        np.testing.assert_array_equal(got[1].view(np.uint32), ref[1].view(np.uint32))
        # the filters did something: the sparse map grew into its empty texels
        assert np.count_nonzero(ref[0][:, 3] == np.float32(0.001)) > 1000 or seed == 12


def test_noalloc_filters_reject_sizes_beyond_the_limit(gpu):
    L = N.lib()
    params = N.FilterParams(1.0, 1000.0, (ctypes.c_float * 4)(0, 0, 0, 0), 0)
    big = (N.Lightmap * 1)(N.Lightmap(0, 1025, 1024, 0))          # 1 texel row beyond 2^20
    # the limit is checked on the host before any device work: no buffers are touched
    code = L.lsg_filter_lightmaps_noalloc(big, 1, ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.byref(params), None)
    assert code == -6 and "at most" in N.last_error()
