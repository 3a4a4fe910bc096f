"""Rotated and blended environments in the CPU oracle, pinned independently of its code.

RRSolver::setEnvironment(env0, env1, angle0, angle1) with setEnvironmentBlendFactor(f):
the gather's rays that leave the scene read RRBuffer::createEnvironmentBlend of them
(pathtracer.cpp:29-34, RRBufferBlend.cpp:21-66), env0(R(angle0) d) * (1 - f) +
env1(R(angle1) d) * f with R(a) the yaw rotation rotationByYawPitchRoll((a, 0, 0))
(RRMath.cpp:457-470), except that with no env1 and angle0 == 0 no blend is made.  Only
env0 turns the hemisphere gather on (gather.cpp:149, 197).

Checked against closed forms: uniform skies mix linearly; the no-blend rule; a cube sky
lit on its +x face only, seen by an upward floor, gives the floor the cosine-weighted
share of the face the rotation brings overhead (a quadrature of the cube face), and
nothing when the rotation turns the face below the floor; the rotation itself equals the
reference's, compiled from RRMath.cpp, bit for bit.  Copies of the oracle with a mistake
planted in the blend must fail."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from test_bake_cpu import _single_texel, _arrays
from test_env_cpu import _floor, _params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GATHER_C = os.path.join(ROOT, "oracle", "lsg_oracle_gather.c")


def _cube(faces, w=4):
    rgb = np.zeros((6, w * w, 3), np.float32)
    for k, L in faces.items():
        rgb[k] = L
    return dict(type=1, width=w, height=w, rgb=rgb.reshape(-1, 3))


def _panorama(L, w=16, h=8):
    return dict(type=2, width=w, height=h, rgb=np.tile(np.float32(L), (w * h, 1)))


def _gather(oracle, env, q=256, mult=1.5, hemisphere=0):
    v, i, n = _floor((0, 0, 1))
    arr = _arrays(v, i, n, np.zeros((2, 8), np.float32), [])
    arr["env"] = env
    tex, sub = _single_texel(0, uv=(0.2, 0.2, 0.6, 0.2, 0.2, 0.6))
    out, st = oracle.gather(arr, _params(q=q, mult=mult, hemisphere=hemisphere), tex.view(np.uint32).reshape(-1, 4), sub)
    return out[0, :3], st


def _with_blend(env0, env1=None, angle0=0.0, angle1=0.0, blend=0.0):
    e = dict(env0) if env0 else dict(type=0, width=0, height=0, rgb=np.zeros((0, 3), np.float32))
    e["blend"] = dict(env1=env1, angle0=angle0, angle1=angle1, blend=blend)
    return e


def _face_shares():
    """Cosine-weighted shares, for an upward receiver, of the cube face overhead (+z) and
    of each side face: F_top = (1/pi) * integral over [-1,1]^2 of du dv / (1+u^2+v^2)^2."""
    n = 2000
    u = (np.arange(n) + 0.5) / n * 2 - 1
    U, V = np.meshgrid(u, u)
    top = float(np.sum(1.0 / (1 + U * U + V * V) ** 2) * (2.0 / n) ** 2 / np.pi)
    return top, (1 - top) / 4


def check_uniform_mix(gather):
    L0, L1 = np.float32([0.2, 0.5, 1.0]), np.float32([1.0, 0.25, 0.5])
    got, _ = gather(_with_blend(_cube({k: L0 for k in range(6)}), _panorama(L1), 0.4, -1.3, 0.25))
    want = 1.5 * (0.75 * L0.astype(np.float64) + 0.25 * L1)
    return float(np.max(np.abs(got - want) / want))


def check_no_blend_rule(gather):
    L0 = np.float32([0.5, 0.5, 1.0])
    sky = _cube({k: L0 for k in range(6)})
    err = 0.0
    got, _ = gather(_with_blend(sky, None, 0.0, 0.7, 0.5))          # no env1, no rotation: env0 as it is
    err = max(err, float(np.max(np.abs(got - 1.5 * L0) / (1.5 * L0))))
    got, _ = gather(_with_blend(sky, None, 0.3, 0.0, 0.5))          # rotated: a blend with nothing
    err = max(err, float(np.max(np.abs(got - 0.75 * L0) / (0.75 * L0))))
    # env1 alone does not turn the hemisphere on; a hemisphere gathered anyway sees it
    got, st = gather(_with_blend(None, _panorama(L0), 0.0, 0.0, 0.25))
    err = max(err, float(np.max(np.abs(got))) + st.hemi_rays)
    got, _ = gather(_with_blend(None, _panorama(L0), 0.0, 0.0, 0.25), hemisphere=1)
    err = max(err, float(np.max(np.abs(got - 0.375 * L0) / (0.375 * L0))))
    return err


def check_rotation(gather):
    """+x face lit: unrotated the floor sees a side face's share; yaw +pi/2 looks the
    world's +z up on the +x face (the top share); yaw -pi/2 puts that face below."""
    L = np.float32([1.0, 1.0, 1.0])
    top, side = _face_shares()
    sky = _cube({0: L}, w=8)
    err = 0.0
    for angle, share in ((0.0, side), (np.pi / 2, top), (-np.pi / 2, 0.0)):
        got, _ = gather(_with_blend(sky, None, float(angle), 0.0, 0.0), q=4096, mult=1.0)
        err = max(err, float(np.max(np.abs(got - share))) / top)
    return err


CHECKS = [(check_uniform_mix, 1e-6), (check_no_blend_rule, 1e-6), (check_rotation, 0.02)]


@pytest.mark.parametrize("check,tol", CHECKS, ids=[c.__name__ for c, _ in CHECKS])
def test_blend_closed_forms(oracle, check, tol):
    assert check(lambda env, **kw: _gather(oracle, env, **kw)) < tol


def test_rotation_equals_the_reference_bit_for_bit(oracle):
    """The oracle's environment rotation equals the reference's own
    rotationByYawPitchRoll((yaw, 0, 0)).transformDirection, compiled from RRMath.cpp
    (oracle/_ref/rotate_ref), bit for bit on 3,000 random yaws and directions, including
    multiples of pi/2 where the matrix entries round to signed zeros."""
    if not os.path.exists(oracle.ROTATE_BIN):
        pytest.skip("oracle/_ref/rotate_ref is built only where the reference sources exist")
    rng = np.random.default_rng(21)
    n = 3000
    yaws = rng.uniform(-7.0, 7.0, n).astype(np.float32)
    yaws[:400] = (np.pi / 2 * rng.integers(-4, 5, 400)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d[:200, rng.integers(0, 3)] = 0.0
    got = oracle.env_rotate(yaws, d)
    ref = oracle.reference_rotate(yaws, d)
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


MUTATIONS = [
    ("const float k0 = 1 - env->blend, k1 = env->blend;", "const float k1 = 1 - env->blend, k0 = env->blend;", check_uniform_mix),
    ("return env->blended && ((env->env1 && env->env1->type) || env->angle0 != 0);", "return env->blended;", check_no_blend_rule),
    ("(p->hemisphere || env0)", "(p->hemisphere || env)", check_no_blend_rule),
    ("const float x = d[0] * m[0] + d[1] * m[1] + d[2] * m[2];", "const float x = d[0] * m[0] + d[1] * m[3] + d[2] * m[6];",
     check_rotation),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to build the planted copies")
@pytest.mark.parametrize("k", range(len(MUTATIONS)))
def test_planted_blend_mistakes_fail(oracle, tmp_path, k):
    old, new, check = MUTATIONS[k]
    src = open(GATHER_C).read()
    assert src.count(old) == 1, "the planted site moved: " + old
    odir = os.path.join(ROOT, "oracle")
    srcs = [os.path.join(odir, f) for f in sorted(os.listdir(odir))
            if f.startswith("lsg_oracle") and f.endswith(".c") and f != "lsg_oracle_gather.c"]
    c = tmp_path / "lsg_oracle_gather.c"
    c.write_text(src.replace(old, new))
    so = tmp_path / "mutant.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-fopenmp", "-shared", "-std=c11", "-I", odir, "-o", str(so),
                    str(c), *srcs, "-lm"], check=True)
    lib = ctypes.CDLL(str(so))
    import oracle_py
    err = check(lambda env, **kw: _gather_with(oracle_py, lib, env, **kw))
    assert err > 1e-3, f"planted mistake survived: {new!r}"


def _gather_with(oracle_py, lib, env, **kw):
    """_gather through another build of the oracle library."""
    orig = oracle_py.lib
    oracle_py.lib = lambda: lib
    try:
        return _gather(oracle_py, env, **kw)
    finally:
        oracle_py.lib = orig
