"""Closed-form pins of the filter oracle's weights (oracle/lsg_oracle_filters.c),
stated from what the reference's filters compute, not from the oracle's code:

- lightmapSmooth (RRBuffer/RRBuffer.cpp:716-732): passes of variance
  min(remaining sigma^2, 0.34), each a 3x3 Gaussian exp(-|d|^2 / (2 s2)) over the
  neighbours the blur flags admit, normalised by the admitted weights;
- lightmapGrowForBilinearInterpolation (:804-863): an empty texel becomes the
  weighted mean of the neighbours whose covered quadrant touches it, weight 1.6
  for a side neighbour and 1 for a corner neighbour;
- lightmapGrow (:912-947, no wrap): one step gives an empty texel the mean of its
  covered neighbours, weight 1.6 for the four side neighbours and 1 for the four
  corners.

Each check is then run against copies of the oracle with one weight planted
wrong, compiled here on the CPU, and must fail for each of them."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FILTERS_C = os.path.join(ROOT, "oracle", "lsg_oracle_filters.c")

ALL_FLAGS = 255
EDGE_ONLY = 16 | 32 | 64 | 128     # EDGE_X0..EDGE_Y1, no quadrant bits


def _gauss_pass(img, alpha, s2):
    """One smoothing pass in binary64 over a map whose covered texels all admit all
    in-map neighbours (every edge flag set)."""
    h, w = alpha.shape
    num = np.zeros_like(img)
    den = np.zeros((h, w))
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            g = np.exp(-(dx * dx + dy * dy) / (2.0 * s2))
            sh = np.zeros_like(img)
            sa = np.zeros((h, w))
            ys, yd = slice(max(dy, 0), h + min(dy, 0)), slice(max(-dy, 0), h + min(-dy, 0))
            xs, xd = slice(max(dx, 0), w + min(dx, 0)), slice(max(-dx, 0), w + min(-dx, 0))
            sh[yd, xd] = img[ys, xs]
            sa[yd, xd] = alpha[ys, xs]
            num += g * sh * sa[..., None]
            den += g * sa
    out = img.copy()
    m = alpha > 0
    out[m] = num[m] / den[m][:, None]
    return out


def check_smooth_impulse(filt):
    """Impulse response of lightmapSmooth at sigma 1 and 0.5 on a fully flagged map;
    returns the largest relative error against the binary64 Gaussian passes."""
    worst = 0.0
    for sigma, passes in [(1.0, [0.34, 0.34, 0.32]), (0.5, [0.25])]:
        h = w = 15
        img = np.zeros((h, w, 4), np.float32)
        img[..., 3] = ALL_FLAGS / 255.0
        img[7, 7, :3] = [1.0, 2.0, 3.0]
        img[3, 9, :3] = [0.5, 0.25, 4.0]
        out, rgb = filt(img, sigma=sigma, spread=0.0)
        want = img[..., :3].astype(np.float64)
        alpha = np.ones((h, w))
        for s2 in passes:
            want = _gauss_pass(want, alpha, s2)
        err = np.abs(rgb - want) / np.maximum(np.abs(want), 1e-3)
        worst = max(worst, float(err.max()))
        # alpha is preserved by the smoothing pass
        assert np.array_equal(out[..., 3], img[..., 3])
    return worst


def _touches(dx, dy, flags):
    """Does a neighbour at offset (dx, dy) cover a quadrant that touches the centre
    texel?  Quadrant (a, b) has bit 1 << (a + 2b); a = 1 is the high-x half."""
    for a in (0, 1):
        for b in (0, 1):
            if flags & (1 << (a + 2 * b)):
                if (dx == 0 or a == (dx < 0)) and (dy == 0 or b == (dy < 0)):
                    return True
    return False


def check_grow_bilinear(filt, seeds=range(6)):
    """Random maps of isolated-or-not empty texels beside foreground texels with random
    flag bytes; returns the largest relative error of the filled texels."""
    worst = 0.0
    for seed in seeds:
        rng = np.random.default_rng(seed)
        h, w = 7, 9
        img = np.zeros((h, w, 4), np.float32)
        fg = rng.random((h, w)) < 0.6
        flags = rng.integers(1, 256, size=(h, w))
        img[..., :3] = rng.uniform(0.1, 4.0, size=(h, w, 3)).astype(np.float32)
        img[..., 3] = np.where(fg, flags / 255.0, 0.0)
        img[~fg, :3] = 0
        out, rgb = filt(img, sigma=0.0, spread=0.0)
        for y in range(h):
            for x in range(w):
                if fg[y, x]:
                    continue
                num = np.zeros(3)
                den = 0.0
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        xx, yy = x + dx, y + dy
                        if (dx or dy) and 0 <= xx < w and 0 <= yy < h and fg[yy, xx] and _touches(dx, dy, flags[yy, xx]):
                            wt = 1.0 if (dx and dy) else 1.6
                            num += wt * img[yy, xx, :3]
                            den += wt
                if den == 0:
                    assert out[y, x, 3] == 0
                    continue
                assert out[y, x, 3] == np.float32(0.001)
                worst = max(worst, float(np.max(np.abs(rgb[y, x] - num / den) / (num / den))))
    return worst


def check_grow_step(filt, seeds=range(6)):
    """One lightmapGrow step on maps whose foreground has no quadrant bits (so the
    bilinear grow leaves every empty texel alone); interior texels only."""
    worst = 0.0
    for seed in seeds:
        rng = np.random.default_rng(100 + seed)
        h, w = 8, 10
        fg = rng.random((h, w)) < 0.35
        img = np.zeros((h, w, 4), np.float32)
        img[..., :3] = np.where(fg[..., None], rng.uniform(0.1, 4.0, size=(h, w, 3)), 0).astype(np.float32)
        img[..., 3] = np.where(fg, EDGE_ONLY / 255.0, 0.0)
        out, rgb = filt(img, sigma=0.0, spread=1.0)
        for y in range(1, h - 1):
            for x in range(1, w - 1):
                if fg[y, x]:
                    continue
                num = np.zeros(3)
                den = 0.0
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        if (dx or dy) and fg[y + dy, x + dx]:
                            wt = 1.0 if (dx and dy) else 1.6
                            num += wt * img[y + dy, x + dx, :3]
                            den += wt
                if den == 0:
                    assert out[y, x, 3] == 0
                    continue
                assert out[y, x, 3] == np.float32(0.001)
                worst = max(worst, float(np.max(np.abs(rgb[y, x] - num / den) / (num / den))))
    return worst


TOL = 2e-6


def test_smooth_weights_closed_form(oracle):
    assert check_smooth_impulse(oracle.filter_lightmap) < TOL


def test_grow_bilinear_weights_closed_form(oracle):
    assert check_grow_bilinear(oracle.filter_lightmap) < TOL


def test_grow_step_weights_closed_form(oracle):
    assert check_grow_step(oracle.filter_lightmap) < TOL


# (substring of the oracle, planted replacement, check that must fail)
MUTATIONS = [
    ("k1 = expf(-0.5f / s2), k2 = expf(-1 / s2)", "k1 = expf(-1 / s2), k2 = expf(-2 / s2)", check_smooth_impulse),
    ("float wt = (ii && jj) ? 1 : 1.6f;", "float wt = (ii && jj) ? 1 : 1.0f;", check_grow_bilinear),
    ("v4 d = add4(mul4(e, 1.6f), dg);", "v4 d = add4(mul4(e, 1.0f), dg);", check_grow_step),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to build the planted copies")
@pytest.mark.parametrize("k", range(len(MUTATIONS)))
def test_planted_weight_mistakes_fail_the_pins(oracle, tmp_path, k):
    old, new, check = MUTATIONS[k]
    src = open(FILTERS_C).read()
    assert src.count(old) == 1, "the planted site moved: " + old
    c = tmp_path / "mutant.c"
    c.write_text(src.replace(old, new))
    so = tmp_path / "mutant.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=c11", "-o", str(so), str(c), "-lm"],
                   check=True)
    lib = ctypes.CDLL(str(so))

    def filt(img, **kw):
        return oracle.filter_lightmap(img, library=lib, **kw)
    try:
        err = check(filt)
    except AssertionError:
        return
    assert err > 1e-3, f"planted mistake survived: {new} (err {err:.2e})"
