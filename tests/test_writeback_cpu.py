"""The drop-in solver's writeback of a baked layer into the caller's buffer
(integration/lsg_writeback.h), run with the reference's own sRGB colour space
(RRColorSpace::create_sRGB, RRColorSpace.cpp:22-91, compiled from its sources into
oracle/_ref/writeback_test), against what RRSolver::updateLightmap does to each pixel
buffer (lightmap.cpp:866-876): lightmapFillBackground with the background made linear
for scaled targets, then copyElementsTo (RRBuffer.cpp:105-131) into the target's colour
space, alpha unchanged.

The expectations are closed forms of the gamma curve (x^0.45 out, x^(1/0.45) in): an
sRGB target's background is the caller's colour (not encoded twice), covered texels are
x^0.45 with their own alpha (not forced to 1), bent normals are never converted but
their background is still made linear, as the reference does."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "writeback_test")

pytestmark = pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/writeback_test is built only where the reference sources exist")

WORK = np.array([[0.25, 0.5, 1.0, 1.0], [0, 0, 0, 0], [0.3, 0.1, 0.05, 0.001], [0.5, 0.5, 0.5, -1.0], [2.0, 0.0, 0.125, 0.5]])
BG = np.array([0.2, 0.5, 0.8, 0.7])
BACKGROUND = np.array([False, True, False, True, False])


def _run():
    out = {}
    for line in subprocess.run([BIN], check=True, capture_output=True, text=True).stdout.splitlines():
        f = line.split()
        key = tuple(int(x) for x in f[1:4])
        out.setdefault(key, np.zeros((5, 4)))[int(f[4])] = [float(x) for x in f[5:9]]
    return out


def _gamma(x, p):
    return np.sign(x) * np.abs(x) ** p


def test_writeback_cases():
    got = _run()
    assert len(got) == 8
    for (scaled, space, bent), v in got.items():
        # alpha is copied, never forced to 1; the background brings its own alpha
        np.testing.assert_array_equal(v[:, 3].astype(np.float32), np.where(BACKGROUND, BG[3], WORK[:, 3]).astype(np.float32))
        convert = scaled and space
        fg = WORK[~BACKGROUND, :3]
        want_fg = _gamma(fg, 0.45) if convert and not bent else fg
        np.testing.assert_allclose(v[~BACKGROUND, :3], want_fg, rtol=2e-6, atol=1e-7)
        if convert and bent:
            want_bg = _gamma(BG[:3], 1 / 0.45)         # made linear, then copied without a colour space
        else:
            want_bg = BG[:3]                           # unconverted, or linear and back
        for row in v[BACKGROUND, :3]:
            np.testing.assert_allclose(row, want_bg, rtol=2e-6)


def test_writeback_srgb_background_not_double_encoded():
    """The bug the reference guards against (lightmap.cpp:870-872): filling an sRGB target
    with the raw background and encoding it again would give bg^0.45."""
    v = _run()[(1, 1, 0)]
    assert np.all(np.abs(v[1, :3] - _gamma(BG[:3], 0.45)) > 0.05)
    np.testing.assert_allclose(v[1, :3], BG[:3], rtol=2e-6)
