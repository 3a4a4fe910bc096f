"""Closed form of the per-corner indirect term in the gather oracle (a hemisphere hit adds
albedo * (I0 (1-u-v) + I1 u + I2 v), pathtracer.cpp:326-348, rrcore.cpp:262-273).

A floor element at (x0, y0, 0) facing up sees only a large ceiling at height h made of
two triangles whose corners carry the indirect irradiance of a linear field
I(x, y) = c + a x + b y.  Linear interpolation over a triangle reproduces a linear field
exactly, so every hit returns albedo * I(hit), and the gathered irradiance is
albedo * (c M0 + a Mx + b My) with the form-factor moments
  M_f = integral over the ceiling of f(x, y) h^2 / (pi r^4) dA,  r^2 = (x-x0)^2 + (y-y0)^2 + h^2,
computed here by adaptive quadrature.  Copies of the oracle that drop the term, swap u
and v, or read the corners' mean instead of interpolating are built on the CPU and
must miss it."""
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

H, W = 1.0, 6.0
X0, Y0 = 1.5, -0.7
C0, AX, BY = 2.0, 0.3, -0.15
ALBEDO = np.float32([0.8, 0.5, 0.25])


def _scene():
    e = 0.001
    v = np.float32([[X0 - e, Y0 - e, 0], [X0 + e, Y0 - e, 0], [X0 + e, Y0 + e, 0], [X0 - e, Y0 + e, 0],
                    [-W, -W, H], [W, -W, H], [W, W, H], [-W, W, H]])
    i = np.uint32([[0, 1, 2], [0, 2, 3], [4, 6, 5], [4, 7, 6]])     # ceiling faces down
    nrm = np.float32([[0, 0, 1]] * 4 + [[0, 0, -1]] * 4)
    field = lambda p: C0 + AX * p[0] + BY * p[1]
    ind = np.zeros((4, 12), np.float32)
    for t in (2, 3):
        for k in range(3):
            ind[t, 4 * k:4 * k + 3] = field(v[i[t, k]])
            ind[t, 4 * k + 3] = ALBEDO[k]
    side = np.uint8([5, 5, 5 | 8, 5 | 8])          # front renders, casts shadows; ceiling fronts take the term
    import oracle_py as O
    arr = dict(verts=v, idx=i, normals=nrm, side_bits=side, exitance=np.zeros((4, 8), np.float32),
               lights=np.zeros(0, O.LIGHT_DTYPE), ind=ind)
    return arr


def _moments():
    from scipy import integrate

    def f(weight):
        g = lambda y, x: weight(x, y) * H * H / (math.pi * ((x - X0) ** 2 + (y - Y0) ** 2 + H * H) ** 2)
        return integrate.dblquad(g, -W, W, -W, W, epsabs=1e-12, epsrel=1e-11)[0]
    return f(lambda x, y: 1.0), f(lambda x, y: x), f(lambda x, y: y)


def expected():
    m0, mx, my = _moments()
    return ALBEDO.astype(np.float64) * (C0 * m0 + AX * mx + BY * my)


def gather_with(gather_fn):
    arr = _scene()
    p = dict(quality=4096, hemisphere=1, ray_length_min=1e-6, locality=1e5, rug_distance=1e-3, light_multiplier=1.0, seed=5)
    tex, sub = _single_texel(0, uv=(0.4, 0.4, 0.6, 0.4, 0.4, 0.6))
    out, _ = gather_fn(arr, p, tex.view(np.uint32).reshape(-1, 4), sub)
    return out[0, :3]


def test_linear_indirect_field_closed_form(oracle):
    got = gather_with(oracle.gather)
    np.testing.assert_allclose(got, expected(), rtol=3e-3)


def test_indirect_term_respects_sides(oracle):
    """Without the side bit the ceiling's hits read only the flat table (zero here)."""
    arr = _scene()
    arr["side_bits"] = np.uint8([5, 5, 5, 5])
    p = dict(quality=256, hemisphere=1, ray_length_min=1e-6, locality=1e5, rug_distance=1e-3, light_multiplier=1.0, seed=5)
    tex, sub = _single_texel(0, uv=(0.4, 0.4, 0.6, 0.4, 0.4, 0.6))
    out, _ = oracle.gather(arr, p, tex.view(np.uint32).reshape(-1, 4), sub)
    assert np.all(out[0, :3] == 0)


SITE = "hitc[k] = e[k] + (c[k] * w + c[4 + k] * u + c[8 + k] * v) * c[4 * k + 3];"
MUTATIONS = [
    "hitc[k] = e[k];",                                                        # term dropped
    "hitc[k] = e[k] + (c[k] * w + c[4 + k] * v + c[8 + k] * u) * c[4 * k + 3];",  # u and v swapped
    "hitc[k] = e[k] + (c[k] + c[4 + k] + c[8 + k]) / 3 * c[4 * k + 3];",      # flat corner mean
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to build the planted copies")
@pytest.mark.parametrize("k", range(len(MUTATIONS)))
def test_planted_indirect_mistakes_fail(oracle, tmp_path, k):
    srcs = list(oracle.SOURCES)
    for f in srcs:
        txt = open(os.path.join(ORACLE, f)).read()
        if f == "lsg_oracle_gather.c":
            assert txt.count(SITE) == 1, "the planted site moved"
            txt = txt.replace(SITE, MUTATIONS[k])
        (tmp_path / f).write_text(txt)
    shutil.copy(os.path.join(ORACLE, "lsg_oracle.h"), tmp_path / "lsg_oracle.h")
    so = tmp_path / "mutant.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=c11", "-fopenmp", "-o", str(so)]
                   + [str(tmp_path / f) for f in srcs] + ["-lm"], check=True)
    saved = oracle._lib
    try:
        oracle._lib = None
        orig = oracle.ORACLE_LIB
        oracle.ORACLE_LIB = str(so)
        got = gather_with(oracle.gather)
    finally:
        oracle.ORACLE_LIB = orig
        oracle._lib = saved
    err = np.max(np.abs(got - expected()) / expected())
    assert err > 0.03, f"planted mistake {k} survived (rel. err {err:.3g})"
