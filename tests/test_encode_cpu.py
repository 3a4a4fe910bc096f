"""The consumer-format fixture (tests/golden/encode_formats.npz, written by the reference's
buffer interface: oracle/_ref/encode_ref) against closed forms computed here in numpy:
a covered texel keeps its colour, a texel with alpha <= 0 takes the background (made linear
with |x|^(1/0.45) first for a scaled target); a scaled target's rgb is sign(x)|x|^0.45 in
binary64 rounded to binary32 unless the layer is bent normals; bytes are int(x * 256)
truncated and clamped to 0..255, with a product outside the int range or NaN giving 0 (the
x86-64 conversion); luminance is (r + g + b) * 0.33333333333333 in binary32.  Where the
reference harness is built the fixture is also regenerated and compared."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "encode_formats.npz")
FORMATS = (0, 1, 2, 3, 4, 9, 10)


def gpow(x, g):
    x = np.asarray(x, np.float32).astype(np.float64)
    with np.errstate(invalid="ignore"):
        return np.where(x >= 0, np.abs(x) ** g, -(np.abs(x) ** g)).astype(np.float32)


def to_byte(x):
    p = np.asarray(x, np.float32) * np.float32(256)
    ok = (p >= -2147483648.0) & (p < 2147483648.0)
    with np.errstate(invalid="ignore"):
        v = np.where(ok, np.trunc(np.where(ok, p, 0)), -2147483648.0)
    return np.clip(v, 0, 255).astype(np.uint8)


def model(rgba, bg, fmt, scaled, bent):
    bg = np.asarray(bg, np.float32).copy()
    if scaled:
        bg[:3] = gpow(bg[:3], np.float64(np.float32(1) / np.float32(0.45)))
    c = np.where((rgba[:, 3] <= 0)[:, None], bg[None, :], rgba).astype(np.float32)
    if scaled and not bent:
        c[:, :3] = gpow(c[:, :3], np.float64(np.float32(0.45)))
    lum = ((c[:, 0] + c[:, 1]).astype(np.float32) + c[:, 2]).astype(np.float32) * np.float32(0.33333333333333)
    if fmt == 0:
        return to_byte(c[:, :3]).reshape(-1)
    if fmt == 1:
        return to_byte(c[:, 2::-1]).reshape(-1)
    if fmt == 2:
        return to_byte(c).reshape(-1)
    if fmt == 3:
        return np.ascontiguousarray(c[:, :3]).view(np.uint8).reshape(-1)
    if fmt == 4:
        return np.ascontiguousarray(c).view(np.uint8).reshape(-1)
    if fmt == 9:
        return to_byte(lum)
    return lum.astype(np.float32).view(np.uint8)


def test_fixture_closed_forms():
    z = np.load(GOLDEN)
    rgba, bg = z["rgba"], z["background"]
    for fmt in FORMATS:
        for scaled in (0, 1):
            for bent in (0, 1):
                got = z[f"f{fmt}_s{scaled}_b{bent}"]
                want = model(rgba, bg, fmt, scaled, bent)
                if fmt in (3, 4, 10):      # floats: NaN payloads compare as values
                    g, w = got.view(np.float32), want.view(np.float32)
                    assert np.array_equal(g, w, equal_nan=True), (fmt, scaled, bent)
                else:
                    np.testing.assert_array_equal(got, want, err_msg=f"{fmt} {scaled} {bent}")


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(HERE), "oracle", "_ref", "encode_ref")),
                    reason="oracle/_ref/encode_ref is built only where the reference sources exist")
def test_fixture_matches_reference_harness():
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_encode_golden as M
    z = np.load(GOLDEN)
    np.testing.assert_array_equal(M.inputs().view(np.uint32), z["rgba"].view(np.uint32))
    with tempfile.TemporaryDirectory() as tmp:
        for fmt in FORMATS:
            for scaled in (0, 1):
                for bent in (0, 1):
                    np.testing.assert_array_equal(M.encode(z["rgba"], fmt, scaled, bent, tmp), z[f"f{fmt}_s{scaled}_b{bent}"])
