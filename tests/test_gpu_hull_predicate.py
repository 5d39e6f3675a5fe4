"""The reference's rounded hull predicate where it has the wrong sign (GPU).

Polygonizer::convexHull pops while the double expression
(p2.x-p1.x)*(p3.y-p1.y) - (p2.y-p1.y)*(p3.x-p1.x) <= 0
(lidar_processing_lib/src/polygonizer.cpp:45-48).  For the triples below,
found by a search with exact rational arithmetic, that value rounds to 0
while the exact value is positive (a strict counter-clockwise turn): the
reference drops the middle point, the exact hull keeps it.  The device
detects the uncertain test and recomputes the list as the reference does;
its output must be the oracle's, which must differ from the exact hull.
"""
from fractions import Fraction

import numpy as np
import pytest

from lidar_amd import native as N

pytestmark = pytest.mark.gpu

# (a, b, c) with exact cross(a, b, c) > 0 and the double value == 0.
TRIPLES = [
    ((0.8357651039198697, 0.43276706790505337), (1.6036240951825178, 0.6678592504001499),
     (2.4791852276067825, 0.9359261479317195)),
    ((0.4221165755827173, 0.029040787574867943), (0.8362934456346427, 0.6045497644376704),
     (1.2546540749922699, 1.185872178050726)),
    ((0.2729144368186801, 0.7115899271852729), (0.8453382372539149, 1.1876213544359047),
     (1.454966881887176, 1.694592576766262)),
]


def rounded_cross(a, b, c):
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])


def exact_hull(x, y):
    """Andrew's chain with exact rational orientation (coordinates)."""
    pts = sorted(set(zip(x, y)))
    if len(pts) < 3:
        return pts

    def cross(o, a, b):
        o, a, b = ([Fraction(v) for v in p] for p in (o, a, b))
        return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])

    lower, upper = [], []
    for p in pts:
        while len(lower) >= 2 and cross(lower[-2], lower[-1], p) <= 0:
            lower.pop()
        lower.append(p)
    for p in reversed(pts):
        while len(upper) >= 2 and cross(upper[-2], upper[-1], p) <= 0:
            upper.pop()
        upper.append(p)
    return lower[:-1] + upper[:-1]


def lists():
    out = []
    for a, b, c in TRIPLES:
        assert rounded_cross(a, b, c) == 0.0
        assert (Fraction(b[0]) - Fraction(a[0])) * (Fraction(c[1]) - Fraction(a[1])) - \
            (Fraction(b[1]) - Fraction(a[1])) * (Fraction(c[0]) - Fraction(a[0])) > 0
        # The triple alone, in two input orders, and with points above the
        # line so the near-collinear vertex sits on the lower chain of a
        # larger polygon.
        out.append([a, b, c])
        out.append([c, a, b])
        above = [(a[0] + 0.3, a[1] + 2.0), (c[0] - 0.2, c[1] + 1.5), ((a[0] + c[0]) / 2, (a[1] + c[1]) / 2 + 3.0)]
        out.append([above[0], a, above[1], b, c, above[2]])
    return out


def test_rounded_predicate_cases_match_the_reference(oracle):
    L = lists()
    with N.Context(max_frames=1, max_points=4096) as ctx:
        st, hulls = ctx.convex_hulls([(np.array([p[0] for p in l]), np.array([p[1] for p in l])) for l in L])
    assert st == N.OK
    differs = 0
    for l, h in zip(L, hulls):
        x = np.array([p[0] for p in l])
        y = np.array([p[1] for p in l])
        ref = oracle.convex_hull(x, y)
        assert np.array_equal(np.stack([x[h], y[h]], 1), np.stack([x[ref], y[ref]], 1)), (l, h, ref)
        ex = exact_hull(list(x), list(y))
        differs += int(len(ex) != len(ref))
    # Every case is one where the reference's result is not the exact hull.
    assert differs == len(L)


def test_certified_lists_keep_the_exact_path(oracle):
    """Float-valued, well-separated lists never take the recomputation; the
    results still equal the oracle's."""
    rng = np.random.default_rng(12)
    L = [(rng.uniform(-30, 30, n).astype(np.float32).astype(np.float64),
          rng.uniform(-30, 30, n).astype(np.float32).astype(np.float64)) for n in (3, 10, 300, 2000)]
    with N.Context(max_frames=1, max_points=8192) as ctx:
        st, hulls = ctx.convex_hulls(L)
    assert st == N.OK
    for (x, y), h in zip(L, hulls):
        ref = oracle.convex_hull(x, y)
        assert np.array_equal(np.stack([x[h], y[h]], 1), np.stack([x[ref], y[ref]], 1))


def test_long_nearly_convex_lists_match_the_reference(oracle):
    """Float-rounded straight edges and arcs keep most of their points as hull
    vertices, so pruning barely shrinks them and the chunk hulls take the
    wavefront's block-and-merge chain (and, past three rounds, the
    big-cluster workgroup's).  Both must give the oracle's vertex lists."""
    rng = np.random.default_rng(31)
    lists = []
    for n in (200, 257, 700, 3000):
        t = np.sort(rng.uniform(0.0, 1.0, n))
        # a wall seen from afar: a straight segment in float32
        x = (10.0 + 40.0 * t).astype(np.float32).astype(np.float64)
        y = (5.0 + 17.3 * t).astype(np.float32).astype(np.float64)
        lists.append((x, y))
        # a pole's front: an arc, float-rounded
        a = np.pi * (0.2 + 0.6 * t)
        lists.append(((30.0 + 0.15 * np.cos(a)).astype(np.float32).astype(np.float64),
                      (-12.0 + 0.15 * np.sin(a)).astype(np.float32).astype(np.float64)))
        # both, shuffled
        k = rng.permutation(2 * n)
        lists.append((np.concatenate([x, lists[-1][0]])[k], np.concatenate([y, lists[-1][1]])[k]))
    with N.Context(max_frames=1, max_points=32768) as ctx:
        st, hulls = ctx.convex_hulls(lists)
    assert st == N.OK
    for (x, y), h in zip(lists, hulls):
        ref = oracle.convex_hull(x, y)
        assert np.array_equal(np.stack([x[h], y[h]], 1), np.stack([x[ref], y[ref]], 1)), (len(x), len(h), len(ref))
