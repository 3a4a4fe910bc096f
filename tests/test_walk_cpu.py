"""Pins of the oracle's subtexel walk (lsgo_subtexel_walk, the selector processTexel uses
to pick each shooter's subtexel, RRSolver/gather.cpp:645-647, 669, 680-688).

The reference loop is re-run here in exact rational arithmetic on the same float32 inputs
(the subtexel areas, their float32 sum and the float32 step areaMax/(n+0.91) of each
series): float rounding can change a decision only where the exact accumulator comes
within the accumulated rounding bound of zero, so every case whose decisions all keep a
wider margin must give the same sequence.  Copies of the oracle with one planted mistake
(the 0.91 margin, the carry from series to series, the accumulator's start) are built
here on the CPU and must fail the same check."""
import ctypes
import os
import shutil
import subprocess
from fractions import Fraction

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


def exact_walk(areas32, series):
    """Returns (sequence, smallest decision margin / rounding bound) of the reference
    loop in exact arithmetic."""
    m = len(areas32)
    amax = np.float32(0)
    for a in areas32:
        amax = np.float32(amax + a)
    A = [Fraction(float(a)) for a in areas32]
    it, accu = 0, -A[0]
    out, worst = [], float("inf")
    ops = 0
    scale = max(float(a) for a in areas32)
    for n in series:
        step32 = np.float32(amax / np.float32(np.float32(n) + np.float32(0.91)))
        step = Fraction(float(step32))
        scale = max(scale, float(step32))
        for _ in range(n):
            accu += step
            ops += 1
            while True:
                # a decision: is the accumulator positive?
                bound = ops * 2.0 ** -24 * scale * 2
                worst = min(worst, abs(float(accu)) / bound if bound else float("inf"))
                if not accu > 0:
                    break
                it = it + 1 if it + 1 < m else 0
                accu -= A[it]
                ops += 1
            out.append(it)
    return np.array(out, np.uint32), worst


def _cases(seed, count):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        m = int(rng.integers(1, 9))
        areas = rng.uniform(0.02, 1.0, size=m) * float(rng.choice([1e-5, 3e-4, 0.01, 1.0]))
        if m > 2 and rng.random() < 0.3:
            areas[rng.integers(0, m)] *= 1e-3          # a sliver the walk may step over
        n1 = int(rng.choice([16, 64, 103, 512, 1024]))
        series = [n1] + [int(rng.choice([2, 11, 103]))] * int(rng.integers(0, 3))
        yield areas.astype(np.float32), series


def check_walk(walk, seed=5, count=300):
    """Number of cases checked, and the first mismatch (or None)."""
    checked = 0
    for areas, series in _cases(seed, count):
        want, margin = exact_walk(areas, series)
        if margin < 1.0:
            continue
        got = walk(areas, series)
        checked += 1
        if not np.array_equal(got, want):
            return checked, (areas, series)
    return checked, None


def test_walk_equals_exact_reference_loop(oracle):
    checked, bad = check_walk(oracle.subtexel_walk)
    assert bad is None, bad
    assert checked >= 250


def test_walk_closed_form_shares():
    """Series 1 covers the texel once, in order: subtexel j gets the shooters whose
    position k*step falls in its part of the area (so its count is a_j/step within 1),
    and the walk never wraps inside series 1 (n*step < areaMax)."""
    import oracle_py as O
    rng = np.random.default_rng(9)
    for _ in range(200):
        m = int(rng.integers(2, 7))
        areas = rng.uniform(0.1, 1.0, size=m).astype(np.float32) * np.float32(1e-4)
        n = int(rng.choice([64, 256, 1024]))
        seq = O.subtexel_walk(areas, [n])
        assert np.all(np.diff(seq.astype(np.int64)) >= 0)
        step = float(areas.astype(np.float64).sum()) / (n + 0.91)
        counts = np.bincount(seq, minlength=m)
        assert np.all(np.abs(counts - areas.astype(np.float64) / step) <= 1.0 + 1e-6), (counts, areas / step)


def test_walk_carries_across_series():
    """A second series starts where the first one stopped, not at subtexel 0."""
    import oracle_py as O
    areas = np.float32([1.0, 1.0, 1.0])
    a = O.subtexel_walk(areas, [10, 10])
    # shooter k (1-based over both series) sits at position k*3/10.91 along the three unit
    # areas, taken cyclically: (0, 1] -> 0, (1, 2] -> 1, (2, 3] -> 2, (3, 4] -> 0, ...
    assert list(a[:10]) == [0, 0, 0, 1, 1, 1, 1, 2, 2, 2]
    assert list(a[10:]) == [0, 0, 0, 0, 1, 1, 1, 1, 2, 2]


MUTATIONS = [
    ("static float walk_step(float area_max, uint32_t n) { return area_max / ((float)n + 0.91f); }",
     "static float walk_step(float area_max, uint32_t n) { return area_max / ((float)n + 0.5f); }"),
    ("        const float step = walk_step(area_max, series_n[k]);\n",
     "        const float step = walk_step(area_max, series_n[k]);\n        walk_init(&w, areas);\n"),
    ("static void walk_init(walk_t* w, const float* areas) { w->it = 0; w->accu = -areas[0]; }",
     "static void walk_init(walk_t* w, const float* areas) { w->it = 0; w->accu = 0; }"),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to build the planted copies")
@pytest.mark.parametrize("k", range(len(MUTATIONS)))
def test_planted_walk_mistakes_fail(oracle, tmp_path, k):
    old, new = MUTATIONS[k]
    srcs = list(oracle.SOURCES)
    for f in srcs:
        txt = open(os.path.join(ORACLE, f)).read()
        if f == "lsg_oracle_gather.c":
            assert txt.count(old) == 1, "the planted site moved: " + old
            txt = txt.replace(old, new)
        (tmp_path / f).write_text(txt)
    shutil.copy(os.path.join(ORACLE, "lsg_oracle.h"), tmp_path / "lsg_oracle.h")
    so = tmp_path / "mutant.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=c11", "-o", str(so)]
                   + [str(tmp_path / f) for f in srcs] + ["-lm"], check=True)
    L = ctypes.CDLL(str(so))
    L.lsgo_subtexel_walk.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]

    def walk(areas, series):
        a = np.ascontiguousarray(areas, np.float32)
        n = np.ascontiguousarray(series, np.uint32)
        out = np.zeros(int(n.sum()), np.uint32)
        L.lsgo_subtexel_walk(a.ctypes.data_as(ctypes.c_void_p), len(a), n.ctypes.data_as(ctypes.c_void_p), len(n),
                             out.ctypes.data_as(ctypes.c_void_p))
        return out
    checked, bad = check_walk(walk)
    assert bad is not None, f"planted mistake survived {checked} cases: {new!r}"
