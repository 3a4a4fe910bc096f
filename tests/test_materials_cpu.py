"""Point materials in the gather oracle (RRObject::getPointMaterial, RRObject.cpp:196-303, read
by the final-gathering handler, pathtracer.h:194-216), pinned on the CPU.

Texel lookups (lsgo_texture_at, RRBufferInMemory::getElementAtPosition): hand-placed texels
at the nearest texel, after wrap-around of negative and large uv, at texel centres and at
known fractions between them for bilinear interpolation; byte formats decoded with
RR_BYTE2FLOAT; the colour space of scaled textures against the reference's own
RRColorSpace (golden vectors written by oracle/_ref/colorspace_ref).

Whole gathers, one small receiver facing up (closed forms by quadrature of the form factor
h^2 / (pi r^4) over the seen area):
- a ceiling whose diffuse texture is red for x < 0 and blue for x >= 0 gives
  F(x < 0) red + F(x >= 0) blue times its irradiance; its emissive texture likewise, times
  the emittance multiplier, and only on the sides that emit;
- a scaled byte texture scales every hit by the colour space's byte table value, exactly;
- below the material's minimal quality the per-triangle albedo applies;
- a sun behind a keyed checkerboard occluder lights a receiver under a transparent cell and
  not under an opaque one; near a cell edge the interpolated lookup of shadow rays
  (quality / 10 above the minimum) decides differently from the nearest texel;
- hemisphere rays through a keyed occluder of 1/4 transparent cells see 1/4 of the ceiling
  behind it, and its opaque cells with their texel colour.
Copies of the oracle with planted mistakes (u and v swapped in the material uv, keying
inverted, bilinear weights swapped, no wrap-around, texel colour or emittance multiplier
dropped, bytes decoded before the table) are built here and must fail."""
import ctypes
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from test_bake_cpu import _single_texel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden", "colorspace_srgb.npz")

X0, Y0 = 1.5, -0.7          # receiver position
H, W = 1.0, 6.0             # ceiling height and half-width
RED = np.float32([0.9, 0.2, 0.1])
BLUE = np.float32([0.1, 0.3, 0.8])
NOMAT = 0xFFFFFFFF


def _tex(fmt, data, scaled=True):
    from lightsprintsdk_amd import bake as B
    return B.texture(fmt, data, scaled)


def _mat(**kw):
    from lightsprintsdk_amd import bake as B
    m = np.zeros(1, B.POINT_MATERIAL_DTYPE)
    m["diffuse_texture"] = -1
    m["emissive_texture"] = -1
    m["transmittance_texture"] = -1
    m["threshold"] = 0.5
    m["albedo"] = 0.5
    m["emittance_multiplier"] = 1.0
    for k, v in kw.items():
        m[k] = v
    return m


def _quad(z, half, facing_down, x0=0.0, y0=0.0):
    """Two triangles of a square at height z; material uv (x + half) / (2 half), likewise y."""
    v = np.float32([[x0 - half, y0 - half, z], [x0 + half, y0 - half, z], [x0 + half, y0 + half, z], [x0 - half, y0 + half, z]])
    i = [[0, 2, 1], [0, 3, 2]] if facing_down else [[0, 1, 2], [0, 2, 3]]
    uv = (v[:, :2] - np.float32([x0 - half, y0 - half])) / np.float32(2 * half)
    return v, np.uint32(i), uv.astype(np.float32)


def scene(parts, receiver=(X0, Y0), lights=()):
    """parts: list of (z, half, facing_down, table_rgb, material index or NOMAT, uv scale).
    The receiver is triangles 0 and 1."""
    import oracle_py as O
    e = 0.001
    rv, ri, ruv = _quad(0.0, e, False, *receiver)
    verts, idx, uvs, tabs, mats, nrm = [rv], [ri], [ruv[ri].reshape(-1, 6)], [np.zeros((2, 8), np.float32)], [NOMAT] * 2, \
        [np.float32([[0, 0, 1]] * 4)]
    base = 4
    for z, half, down, tab, mat, scale in parts:
        v, i, uv = _quad(z, half, down)
        verts.append(v)
        idx.append(i + base)
        uvs.append((uv[i] * np.float32(scale)).reshape(-1, 6))
        t = np.zeros((2, 8), np.float32)
        t[:, 0:3] = tab
        tabs.append(t)
        mats += [mat] * 2
        nrm.append(np.float32([[0, 0, -1 if down else 1]] * 4))
        base += 4
    lt = np.zeros(len(lights), O.LIGHT_DTYPE)
    for k, l in enumerate(lights):
        lt[k] = l
    nt = 2 * (len(parts) + 1)
    return dict(verts=np.concatenate(verts), idx=np.concatenate(idx), normals=np.concatenate(nrm),
                side_bits=np.full(nt, 5, np.uint8), exitance=np.concatenate(tabs), lights=lt,
                tri_uv=np.concatenate(uvs).astype(np.float32), tri_mat=np.uint32(mats))


def with_pm(arr, mats, textures, gamma=0.0):
    a = dict(arr)
    a["pm"] = dict(tri_mat=arr["tri_mat"], tri_uv=arr["tri_uv"], mats=np.concatenate(mats) if len(mats) else mats,
                   textures=textures, gamma=gamma)
    return a


def run(gather, arr, quality=4096, hemisphere=1, light_multiplier=1.0, seed=5):
    p = dict(quality=quality, hemisphere=hemisphere, ray_length_min=1e-6, locality=1e5, rug_distance=1e-3,
             light_multiplier=light_multiplier, seed=seed)
    tex, sub = _single_texel(0, uv=(0.4, 0.4, 0.6, 0.4, 0.4, 0.6))
    out, _ = gather(arr, p, tex.view(np.uint32).reshape(-1, 4), sub)
    return out[0, :3]


def form_factor(x0, x1, y0, y1, h, px=X0, py=Y0):
    from scipy import integrate
    g = lambda y, x: h * h / (math.pi * ((x - px) ** 2 + (y - py) ** 2 + h * h) ** 2)
    return integrate.dblquad(g, x0, x1, y0, y1, epsabs=1e-12, epsrel=1e-11)[0]


# ---- texel lookups -------------------------------------------------------------------

def _grid_rgbaf(w, h):
    d = np.zeros((h, w, 4), np.float32)
    for y in range(h):
        for x in range(w):
            d[y, x] = [x, y, 10 * x + y, 1 + x + 100 * y]
    return d


def check_lookups(O):
    t = _tex(4, _grid_rgbaf(4, 2))
    for (u, v), (x, y) in {(0.1, 0.2): (0, 0), (0.3, 0.7): (1, 1), (-0.1, 0.2): (3, 0), (2.6, -0.3): (2, 1),
                           (0.99, 0.49): (3, 0), (7.25, 3.5): (1, 1)}.items():
        np.testing.assert_array_equal(O.texture_at(t, u, v), _grid_rgbaf(4, 2)[y, x], err_msg=f"nearest at {u, v}")
    g = _grid_rgbaf(4, 2)
    # texel centres: exactly the texel; fractions between centres: the bilinear blend
    np.testing.assert_array_equal(O.texture_at(t, 1.5 / 4, 0.5 / 2, interpolated=True), g[0, 1])
    got = O.texture_at(t, (1.5 + 0.25) / 4, (0.5 + 0.75) / 2, interpolated=True)
    want = g[0, 1] * 0.75 * 0.25 + g[1, 1] * 0.75 * 0.75 + g[0, 2] * 0.25 * 0.25 + g[1, 2] * 0.25 * 0.75
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5)
    # wrap-around between the last and the first column
    got = O.texture_at(t, 3.75 / 4, 0.5 / 2, interpolated=True)
    np.testing.assert_allclose(got, g[0, 3] * 0.75 + g[0, 0] * 0.25, rtol=1e-5, atol=1e-5)
    # byte formats: RR_BYTE2FLOAT, BGR swapped, luminance copied, alpha 1 where absent
    b = np.uint8([[[10, 20, 30], [40, 50, 60]]])
    k = np.float32(0.003921568627450980392156862745098)
    np.testing.assert_array_equal(O.texture_at(_tex(0, b, False), 0.2, 0.5), np.float32([10, 20, 30, 255]) * k * [1, 1, 1, 0] + [0, 0, 0, 1])
    np.testing.assert_array_equal(O.texture_at(_tex(1, b, False), 0.7, 0.5), np.float32([60, 50, 40, 0]) * k + [0, 0, 0, 1])
    np.testing.assert_array_equal(O.texture_at(_tex(9, np.uint8([[7, 9]]), False), 0.7, 0.5), np.float32([9, 9, 9, 0]) * k + [0, 0, 0, 1])


def check_colour_space(O):
    g = np.load(GOLDEN)
    gamma = float(g["gamma"])
    allb = np.arange(256, dtype=np.uint8)
    rgb = _tex(0, np.stack([allb, allb, allb], axis=1)[None], True)
    rgba = _tex(2, np.stack([allb, allb, allb, allb[::-1]], axis=1)[None], True)
    got = np.array([O.texture_at(rgb, (i + 0.5) / 256, 0.5, gamma)[0] for i in range(256)], np.float32)
    np.testing.assert_array_equal(got, g["byte_linear"])
    got4 = np.array([O.texture_at(rgba, (i + 0.5) / 256, 0.5, gamma) for i in range(256)], np.float32)
    np.testing.assert_array_equal(got4[:, 2], g["byte_linear"])
    np.testing.assert_array_equal(got4[:, 3], allb[::-1].astype(np.float32) * np.float32(0.003921568627450980392156862745098))
    # float formats: toLinear of the stored value
    x = g["x"]
    ft = _tex(3, np.stack([x, x, x], axis=1)[None], True)
    got = np.array([O.texture_at(ft, (i + 0.5) / len(x), 0.5, gamma)[1] for i in range(len(x))], np.float32)
    np.testing.assert_array_equal(got, g["to_linear"])
    # an unscaled texture is read as stored
    np.testing.assert_array_equal(O.texture_at(_tex(3, np.stack([x, x, x], axis=1)[None], False), 0.5 / len(x), 0.5, gamma)[:3],
                                  np.full(3, x[0]))


def test_texel_lookups(oracle):
    check_lookups(oracle)


def test_colour_space_equals_reference(oracle):
    check_colour_space(oracle)


# ---- whole gathers ----------------------------------------------------------------------

def split_ceiling(kind, emit_sides=1, multiplier=2.0, min_quality=0):
    """The ceiling's texture is red for x < 0 and blue for x >= 0 (a 2 x 1 RGBF texture)."""
    t = _tex(3, np.stack([RED, BLUE])[None], True)
    if kind == "diffuse":
        arr = scene([(H, W, True, 1.0, 0, 1.0)])
        m = _mat(diffuse_texture=0, min_quality=min_quality, albedo=[0.25, 0.5, 0.75])
    else:
        arr = scene([(H, W, True, 0.0, 0, 1.0)])
        m = _mat(emissive_texture=0, emit_sides=emit_sides, emittance_multiplier=multiplier, min_quality=min_quality,
                 emittance=[0.3, 0.3, 0.3])
    return with_pm(arr, [m], [t])


def split_expected():
    fl = form_factor(-W, 0, -W, W, H)
    fr = form_factor(0, W, -W, W, H)
    return fl * RED.astype(np.float64) + fr * BLUE.astype(np.float64)


def check_split(gather):
    got = run(gather, split_ceiling("diffuse"))
    np.testing.assert_allclose(got, split_expected(), rtol=1e-2)
    got = run(gather, split_ceiling("emissive"))
    np.testing.assert_allclose(got, 2.0 * split_expected(), rtol=1e-2)


def test_textured_ceiling_closed_forms(oracle):
    check_split(oracle.gather)
    # an emitting back side is not seen from below; below the minimal quality the
    # triangle's albedo and flat emittance apply
    assert np.all(run(oracle.gather, split_ceiling("emissive", emit_sides=2)) == 0)
    f = form_factor(-W, W, -W, W, H)
    np.testing.assert_allclose(run(oracle.gather, split_ceiling("diffuse", min_quality=4096)), f * np.float64([0.25, 0.5, 0.75]), rtol=1e-2)
    np.testing.assert_allclose(run(oracle.gather, split_ceiling("emissive", min_quality=5000)), f * 0.3, rtol=1e-2)


def test_scaled_byte_texture_scales_every_hit(oracle):
    g = np.load(GOLDEN)
    plain = run(oracle.gather, with_pm(scene([(H, W, True, 1.0, 0, 1.0)]), [_mat(albedo=[1, 1, 1])], []), quality=1024)
    arr = with_pm(scene([(H, W, True, 1.0, 0, 1.0)]), [_mat(diffuse_texture=0)],
                  [_tex(0, np.uint8([[[200, 100, 50]]]), True)], gamma=float(g["gamma"]))
    tex = run(oracle.gather, arr, quality=1024)
    np.testing.assert_allclose(tex / plain, g["byte_linear"][[200, 100, 50]].astype(np.float64), rtol=1e-12)


def checker_occluder(tex_alpha, threshold=0.5, min_quality=0, half=2.0, z=1.0, scale=1.0, tab=0.0, rgb=(255, 128, 0),
                     parts_before=(), receiver=(X0, Y0), lights=()):
    h, w = tex_alpha.shape
    d = np.zeros((h, w, 4), np.uint8)
    d[:, :, 0:3] = rgb
    d[:, :, 3] = tex_alpha
    from lightsprintsdk_amd import bake as B
    m = _mat(diffuse_texture=0, transmittance_texture=0, flags=B.PM_TRANSMITTANCE_IN_ALPHA | B.PM_KEYED, threshold=threshold,
             min_quality=min_quality)
    arr = scene(list(parts_before) + [(z, half, True, tab, 0, scale)], receiver=receiver, lights=lights)
    return with_pm(arr, [m], [_tex(2, d, False)])


def _checker8():
    a = np.zeros((8, 8), np.uint8)
    for j in range(8):
        for i in range(8):
            a[j, i] = 255 if (i + j) % 2 == 0 else 0
    return a


SUN = [(1, (0, 0, 0), (0, 0, -1), (1, 1, 1), 1)]


def check_shadows(gather):
    a = _checker8()
    # cell (i, j) centre: x = -2 + 0.5 (i + 0.5); (1, 0) is transparent, (0, 0) opaque
    lit = run(gather, checker_occluder(a, receiver=(-1.25, -1.75), lights=SUN), quality=100, hemisphere=0)
    dark = run(gather, checker_occluder(a, receiver=(-1.75, -1.75), lights=SUN), quality=100, hemisphere=0)
    np.testing.assert_array_equal(lit, [1, 1, 1])
    np.testing.assert_array_equal(dark, [0, 0, 0])
    # u w = 4.1 in row 0: the nearest texel (column 4) is opaque; the interpolated alpha 0.6
    # keys transparent with threshold 0.25.  quality / 10 above min_quality 20 interpolates.
    edge = dict(threshold=0.25, min_quality=20, receiver=(-2 + 4 * 4.1 / 8, -1.75), lights=SUN)
    np.testing.assert_array_equal(run(gather, checker_occluder(a, **edge), quality=300, hemisphere=0), [1, 1, 1])
    np.testing.assert_array_equal(run(gather, checker_occluder(a, **edge), quality=150, hemisphere=0), [0, 0, 0])


def test_keyed_occluder_shadows(oracle):
    check_shadows(oracle.gather)


def quarter_pattern():
    return np.uint8([[0, 255], [255, 255]])


def check_quarter(gather):
    """A ceiling of irradiance 1 behind a large keyed occluder (1/4 of its cells transparent,
    tiled 40 times, cells 0.25 wide) whose opaque cells carry irradiance 0.5 and byte colour
    (255, 128, 0): E = F_ceiling / 4 + 3/4 F_occluder 0.5 colour."""
    arr = checker_occluder(quarter_pattern(), half=10.0, z=0.5, scale=40.0, tab=0.5, parts_before=[(H, W, True, 1.0, NOMAT, 1.0)])
    got = run(gather, arr, quality=16384)
    k = np.float32(0.003921568627450980392156862745098)
    colour = np.float32([255, 128, 0]) * k
    want = 0.25 * form_factor(-W, W, -W, W, H) + 0.75 * form_factor(-10, 10, -10, 10, 0.5) * 0.5 * colour.astype(np.float64)
    np.testing.assert_allclose(got, want, rtol=0.03, atol=2e-3)


def test_keyed_occluder_hemisphere_fraction(oracle):
    check_quarter(oracle.gather)


# ---- planted mistakes ---------------------------------------------------------------------

MUTATIONS = [
    ("lsg_oracle_gather.c", "const float mu = c[0] * w + c[2] * u + c[4] * v, mv = c[1] * w + c[3] * u + c[5] * v;",
     "const float mu = c[0] * w + c[2] * v + c[4] * u, mv = c[1] * w + c[3] * v + c[5] * u;", "uv swapped", "split"),
    ("lsg_oracle_gather.c", "if (M->flags & 2) t = t >= M->threshold ? 1.f : 0.f;",
     "if (M->flags & 2) t = t < M->threshold ? 1.f : 0.f;", "keying inverted", "shadows"),
    ("lsg_oracle_gather.c", "if (M->flags & 2) t = t >= M->threshold ? 1.f : 0.f;",
     "if (M->flags & 2) t = 0.f;", "keying ignored", "quarter"),
    ("lsg_oracle_gather.c", "const float w00 = (1 - rx) * (1 - ry), w01 = (1 - rx) * ry, w10 = rx * (1 - ry), w11 = rx * ry;",
     "const float w00 = (1 - rx) * (1 - ry), w10 = (1 - rx) * ry, w01 = rx * (1 - ry), w11 = rx * ry;", "bilinear weights", "lookups"),
    ("lsg_oracle_gather.c", "const float px = fmodf(u, 1) + 2, py = fmodf(v, 1) + 2;",
     "const float px = u, py = v;", "no wrap-around", "lookups"),
    ("lsg_oracle_gather.c", "for (int k = 0; k < 3; k++) { dif[k] = P.dif[k]; emi[k] = P.emi[k]; }",
     "for (int k = 0; k < 3; k++) { emi[k] = P.emi[k]; }", "texel colour dropped", "split"),
    ("lsg_oracle_gather.c", "for (int k = 0; k < 3; k++) m->emi[k] = rgba[k] * M->emittance_multiplier;",
     "for (int k = 0; k < 3; k++) m->emi[k] = rgba[k];", "multiplier dropped", "split"),
    ("lsg_oracle_gather.c", "r[c] = (float)pow((double)((float)b[k + c] / 255), (double)(1 / gamma));",
     "r[c] = cs_to_linear(byte_to_float(b[k + c]), gamma);", "bytes decoded before the table", "colour"),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to build the planted copies")
@pytest.mark.parametrize("k", range(len(MUTATIONS)))
def test_planted_material_mistakes_fail(oracle, tmp_path, k):
    fname, old, new, label, check = MUTATIONS[k]
    srcs = list(oracle.SOURCES)
    for f in srcs:
        txt = open(os.path.join(ORACLE, f)).read()
        if f == fname:
            assert txt.count(old) == 1, f"the planted site moved: {label}"
            txt = txt.replace(old, new)
        (tmp_path / f).write_text(txt)
    shutil.copy(os.path.join(ORACLE, "lsg_oracle.h"), tmp_path / "lsg_oracle.h")
    so = tmp_path / "mutant.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=c11", "-fopenmp", "-o", str(so)]
                   + [str(tmp_path / f) for f in srcs] + ["-lm"], check=True)
    saved, orig = oracle._lib, oracle.ORACLE_LIB
    try:
        oracle._lib = None
        oracle.ORACLE_LIB = str(so)
        with pytest.raises(AssertionError):
            {"split": lambda: check_split(oracle.gather), "shadows": lambda: check_shadows(oracle.gather),
             "quarter": lambda: check_quarter(oracle.gather), "lookups": lambda: check_lookups(oracle),
             "colour": lambda: check_colour_space(oracle)}[check]()
    finally:
        oracle.ORACLE_LIB = orig
        oracle._lib = saved
