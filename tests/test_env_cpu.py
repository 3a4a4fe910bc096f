"""Environment lighting in the CPU oracle, pinned independently of its own code.

A hemisphere ray that leaves the scene returns environment(dir) * multiplier
(PathtracerWorker::getIncidentRadiance, RRStaticSolver/pathtracer.cpp:156-173), looked up
nearest-texel as RRBufferInMemory::getElementAtDirection (RRBufferInMemory.cpp:410-455).
The lookup is checked against the OpenGL cube-map face table and against an atan2/asin
statement of the panorama; the gather against closed forms (uniform sky, half sky, and
a sky seen past a roof, E = L * (1 - F) with F the analytic form factor of the roof).
"""
import math

import numpy as np
import pytest

from test_bake_cpu import _single_texel, _arrays


def _gl_cube_index(d, w):
    """OpenGL cube-map selection (major axis -> face, sc, tc, ma), faces +x,-x,+y,-y,+z,-z."""
    x, y, z = (float(c) for c in d)
    ax, ay, az = abs(x), abs(y), abs(z)
    if ax >= ay and ax >= az:
        face, sc, tc, ma = (0, -z, -y, ax) if x >= 0 else (1, z, -y, ax)
    elif ay >= az:
        face, sc, tc, ma = (2, x, z, ay) if y >= 0 else (3, x, -z, ay)
    else:
        face, sc, tc, ma = (4, x, -y, az) if z >= 0 else (5, -x, -y, az)
    s = (sc / ma + 1) / 2 * w
    t = (tc / ma + 1) / 2 * w
    return face, s, t


def _panorama_uv(d):
    """Yaw measured from +z towards +x maps to u = 0.75 - yaw/(2 pi) (mod 1); elevation
    maps to v = 0.5 + asin(y)/pi."""
    x, y, z = (float(c) for c in d)
    yaw = math.atan2(x, z)
    u = (0.75 - yaw / (2 * math.pi)) % 1.0
    v = 0.5 + math.asin(y / math.sqrt(x * x + y * y + z * z)) / math.pi
    return u, v


def _dirs(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


def test_cube_lookup_follows_gl_face_table(oracle):
    w = 16
    env = dict(type=oracle.ENV_CUBE, width=w, height=w, rgb=np.zeros((6 * w * w, 3), np.float32))
    # the six axis directions hit the centre of their faces
    for k, d in enumerate([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]):
        assert oracle.env_index(env, d) == w // 2 + (w // 2) * w + k * w * w
    checked = 0
    for d in _dirs(3000, 1):
        face, s, t = _gl_cube_index(d, w)
        if min(abs(s - round(s)), abs(t - round(t))) < 1e-3:
            continue                                   # too close to a texel edge to compare
        a = sorted(np.abs(d))
        if a[2] - a[1] < 1e-4:
            continue                                   # major-axis tie
        want = min(int(s), w - 1) + min(int(t), w - 1) * w + face * w * w
        assert oracle.env_index(env, d) == want, d
        checked += 1
    assert checked > 2500


def test_panorama_lookup_yaw_and_elevation(oracle):
    w, h = 32, 16
    env = dict(type=oracle.ENV_PANORAMA, width=w, height=h, rgb=np.zeros((w * h, 3), np.float32))
    # +z looks at u = 0.75, +x at 0.5, -z at 0.25; the horizon is the middle row
    assert oracle.env_index(env, (0, 0, 1)) == 24 + 8 * w
    assert oracle.env_index(env, (0, 0, -1)) == 8 + 8 * w
    assert oracle.env_index(env, (0.6, 0.0, 0.8)) == int(((0.75 - math.atan2(0.6, 0.8) / (2 * math.pi)) % 1) * w) + 8 * w
    checked = 0
    for d in _dirs(3000, 2):
        u, v = _panorama_uv(d)
        if min(abs(u * w - round(u * w)), abs(v * h - round(v * h))) < 1e-3:
            continue
        want = int(u * w) % w + (int(v * h) % h) * w
        assert oracle.env_index(env, d) == want, d
        checked += 1
    assert checked > 2500


def _floor(normal):
    """Unit quad through the origin with the given (axis-aligned) vertex normal."""
    n = np.asarray(normal, np.float32)
    if abs(n[2]) > 0.5:
        v = np.float32([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]])
    elif abs(n[1]) > 0.5:
        v = np.float32([[0, 0, 0], [0, 0, 1], [1, 0, 1], [1, 0, 0]])
    else:
        v = np.float32([[0, 0, 0], [0, 1, 0], [0, 1, 1], [0, 0, 1]])
    i = np.uint32([[0, 1, 2], [0, 2, 3]])
    return v, i, np.tile(n, (4, 1))


def _params(q=1024, mult=2.0, hemisphere=0, seed=5):
    return dict(quality=q, hemisphere=hemisphere, ray_length_min=1e-6, locality=1e5, rug_distance=1e-3,
                light_multiplier=1.0, seed=seed, environment_multiplier=mult)


def _panorama_half_sky(L, w=64, h=32):
    rgb = np.zeros((h, w, 3), np.float32)
    rgb[h // 2:] = L                     # elevation >= 0 (rows from the horizon up)
    return dict(type=2, width=w, height=h, rgb=rgb.reshape(-1, 3))


def test_uniform_sky_gives_its_radiance(oracle):
    """Nothing in the way: every hemisphere ray leaves and returns L * multiplier, and the
    environment alone turns the hemisphere gather on (gather.cpp:197)."""
    L = np.float32([0.2, 0.5, 1.0])
    v, i, n = _floor((0, 0, 1))
    arr = _arrays(v, i, n, np.zeros((2, 8), np.float32), [])
    arr["env"] = dict(type=1, width=4, height=4, rgb=np.tile(L, (6 * 16, 1)))
    tex, sub = _single_texel(0, uv=(0.2, 0.2, 0.6, 0.2, 0.2, 0.6), flags=200)
    out, st = oracle.gather(arr, _params(q=256), tex.view(np.uint32).reshape(-1, 4), sub)
    np.testing.assert_allclose(out[0, :3], 2.0 * L, rtol=1e-7)
    assert out[0, 3] == pytest.approx(200 / 255) and st.hemi_rays == 256
    # a zero multiplier means no environment and, here, nothing to gather
    out, st = oracle.gather(arr, _params(q=256, mult=0.0), tex.view(np.uint32).reshape(-1, 4), sub)
    assert np.all(out[0] == 0) and st.hemi_rays == 0


def test_half_sky_floor_and_wall(oracle):
    """Bright upper half-space: an upward floor sees only sky (E = L exactly); a vertical
    wall sees it over half of its cosine-weighted hemisphere (E = L/2)."""
    L = np.float32([1.0, 0.8, 0.6])
    env = _panorama_half_sky(L)
    for normal, frac, rtol in [((0, 1, 0), 1.0, 1e-7), ((1, 0, 0), 0.5, 0.02), ((0, 0, -1), 0.5, 0.02)]:
        v, i, n = _floor(normal)
        arr = _arrays(v, i, n, np.zeros((2, 8), np.float32), [])
        arr["env"] = env
        tex, sub = _single_texel(0, uv=(0.2, 0.2, 0.6, 0.2, 0.2, 0.6))
        out, _ = oracle.gather(arr, _params(q=4096, mult=1.0), tex.view(np.uint32).reshape(-1, 4), sub)
        np.testing.assert_allclose(out[0, :3], frac * L, rtol=rtol, err_msg=str(normal))


def test_sky_past_a_roof_form_factor(oracle):
    """A small floor patch under a black square roof (half-size a at height h) in a
    uniform sky: E = L * (1 - F), F the analytic differential-area-to-parallel-rectangle
    form factor."""
    h, a = 1.0, 0.7
    v = np.float32([[-0.001, -0.001, 0], [0.001, -0.001, 0], [0.001, 0.001, 0], [-0.001, 0.001, 0],
                    [-a, -a, h], [a, -a, h], [a, a, h], [-a, a, h]])
    i = np.uint32([[0, 1, 2], [0, 2, 3], [4, 6, 5], [4, 7, 6]])
    nrm = np.float32([[0, 0, 1]] * 4 + [[0, 0, -1]] * 4)
    L = 1.5
    arr = _arrays(v, i, nrm, np.zeros((4, 8), np.float32), [], side=7)
    arr["env"] = dict(type=1, width=2, height=2, rgb=np.full((24, 3), L, np.float32))
    tex, sub = _single_texel(0, uv=(0.4, 0.4, 0.6, 0.4, 0.4, 0.6))
    out, _ = oracle.gather(arr, _params(q=4096, mult=1.0), tex.view(np.uint32).reshape(-1, 4), sub)

    def corner(A, B):
        return (A / math.sqrt(1 + A * A) * math.atan(B / math.sqrt(1 + A * A))
                + B / math.sqrt(1 + B * B) * math.atan(A / math.sqrt(1 + B * B))) / (2 * math.pi)
    F = 4 * corner(a / h, a / h)
    np.testing.assert_allclose(out[0, :3], L * (1 - F), rtol=1e-2)


def test_environment_in_the_layers(oracle):
    """The layered gather sees the same environment: its lightmap plane equals the
    plain gather's, and under a half sky the bent normal of a wall tilts upwards."""
    L = np.float32([1.0, 1.0, 1.0])
    v, i, n = _floor((1, 0, 0))
    arr = _arrays(v, i, n, np.zeros((2, 8), np.float32), [])
    arr["env"] = _panorama_half_sky(L)
    tex, sub = _single_texel(0, uv=(0.2, 0.2, 0.6, 0.2, 0.2, 0.6))
    t4 = tex.view(np.uint32).reshape(-1, 4)
    p = _params(q=1024, mult=1.0)
    plain, _ = oracle.gather(arr, p, t4, sub)
    lay = oracle.gather_layers(arr, p, t4, sub, layers=(0, 4))
    np.testing.assert_array_equal(lay[0], plain)
    bent = lay[4][0, :3] * 2 - 1
    # mean of cosine-weighted directions over the upper half of the x hemisphere:
    # E[x] = 2/3 * ... both components positive, tilted 45-60 degrees towards +y
    assert bent[0] > 0.3 and bent[1] > 0.3 and abs(bent[2]) < 0.05
