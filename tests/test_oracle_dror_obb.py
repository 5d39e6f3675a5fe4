"""Pins of the oracle's noise filter (DROR) and rotating-calipers box.

NoiseRemover::filter: lidar_processing_lib/src/noise_remover.cpp:38-68,
KDTree: lidar_processing_lib/include/lidar_processing_lib/kdtree.hpp:131-143,
170-213, 339-397.  Oriented box: lidar_processing_lib/src/polygonizer.cpp:93-278.

What pins what:
  - the grid-based rule is checked against an O(n^2) brute-force count;
  - the KD-tree restatement is checked against the same brute force in the
    regime where its search never stops early (min_neighbours above every
    neighbourhood), so its tree build and pruning are exact;
  - the stale-stack behaviour only ever adds counts, so NOISE under the
    reference object implies NOISE under the rule;
  - antipodal pairs are checked against hand-traced walks of the reference
    loop and against the geometric definition (parallel supporting lines
    through both vertices);
  - boxes are checked against the least-area rectangle over the candidate
    edges (computed here by projection), containment and rectangle shape,
    and against known rectangles.
"""
import math

import numpy as np
import pytest

f32 = np.float32


def brute_counts(oracle, x, y, z, cfg=None):
    cfg = cfg or oracle.noise_config()
    x, y, z = (np.asarray(a, f32) for a in (x, y, z))
    out = np.zeros(len(x), np.uint32)
    for i in range(len(x)):
        r2 = f32(oracle.noise_radius_sqr(x[i], y[i], cfg))
        d0 = x[i] - x
        d1 = y[i] - y
        d2 = z[i] - z
        d = d0 * d0 + (d1 * d1 + (d2 * d2 + f32(0)))
        out[i] = int(np.sum(d <= r2))
    return out


def cloud(seed, n_clusters=40, n_iso=60):
    rng = np.random.default_rng(seed)
    pts = []
    for _ in range(n_clusters):
        c = rng.uniform(-30, 30, 3) * [1, 1, 0.1]
        m = int(rng.integers(1, 9))
        pts.append(c + rng.normal(0, 0.05, (m, 3)))
    pts.append(rng.uniform(-40, 40, (n_iso, 3)) * [1, 1, 0.1])
    p = np.concatenate(pts).astype(f32)
    rng.shuffle(p)
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()


def test_radius_closed_form(oracle):
    # r = max(0.02 d, 0.1): 0.1 m inside 5 m, 0.02 d beyond.
    for d in (0.0, 1.0, 4.9):
        assert oracle.noise_radius_sqr(d, 0.0) == f32(0.1) * f32(0.1)
    for d in (5.5, 10.0, 37.0, 100.0):
        r2 = oracle.noise_radius_sqr(d, 0.0)
        assert abs(r2 - (0.02 * d) ** 2) <= 1e-6 * (0.02 * d) ** 2
    # The product is formed in double from the float sum and rounded once.
    x, y = f32(3.0), f32(4.0)
    assert oracle.noise_radius_sqr(x, y) == max(f32(np.float64(f32(0.02)) ** 2 * np.float64(x * x + y * y)),
                                                f32(0.1) * f32(0.1))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_rule_matches_brute_force(oracle, seed):
    x, y, z = cloud(seed)
    labels, counts = oracle.noise_definition(x, y, z)
    ref = brute_counts(oracle, x, y, z)
    assert np.array_equal(counts, ref)
    assert np.array_equal(labels, (ref < 4).astype(np.uint8))
    assert 0 < labels.sum() < len(labels)


def test_rule_boundary_is_inclusive(oracle):
    # Neighbours exactly at the search radius count (dist_sqr <= r^2).
    # At the origin r^2 = 0.1f*0.1f, and a neighbour 0.1f away along an axis
    # has dist_sqr = 0.1f*0.1f exactly.
    r = f32(0.1)
    x = np.array([0.0, r, -r, 0.0], f32)
    y = np.array([0.0, 0.0, 0.0, r], f32)
    z = np.zeros(4, f32)
    _, counts = oracle.noise_definition(x, y, z)
    assert np.array_equal(counts, brute_counts(oracle, x, y, z))
    assert list(counts) == [4, 2, 2, 2]


def test_kdtree_search_exact_when_it_never_stops_early(oracle):
    x, y, z = cloud(7)
    big = oracle.noise_config(min_neighbours=1000)
    ref = brute_counts(oracle, x, y, z, big) < 1000
    nr = oracle.NoiseRemover()
    for _ in range(2):  # the stack is empty after every such query
        assert np.array_equal(nr.filter(x, y, z, big), ref.astype(np.uint8))
        assert nr.stack_depth() == 0


def test_stale_stack_only_adds_neighbours(oracle, synth):
    f = synth.workload_c2(0)
    rule, _ = oracle.noise_definition(f.x, f.y, f.z)
    nr = oracle.NoiseRemover()
    ref = nr.filter(f.x, f.y, f.z)
    # NOISE under the reference object implies NOISE under the rule ...
    assert np.all(rule[ref == 1] == 1)
    # ... and the difference is the points rescued by counting a point twice.
    assert (rule != ref).sum() < 0.2 * rule.sum()


def test_stale_stack_rescues_a_sparse_point(oracle):
    # Three mutually close points (3 < 4 neighbours) queried after a dense
    # cluster: the dense queries stop at their 4th neighbour and leave
    # subtrees on the search stack, which the sparse queries then revisit.
    rng = np.random.default_rng(3)
    dense = rng.normal([10, 0, 0], 0.02, (200, 3))
    sparse = np.array([[20, 5, 0], [20.03, 5, 0], [20, 5.03, 0]])
    p = np.concatenate([dense, sparse]).astype(f32)
    x, y, z = p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()
    rule, counts = oracle.noise_definition(x, y, z)
    assert np.all(rule[-3:] == 1) and np.all(counts[-3:] == 3)
    ref = oracle.NoiseRemover().filter(x, y, z)
    assert np.all(ref[:200] == 0)
    assert ref[-3:].sum() < 3


# ------------------------------------------------------------------ boxes
def convex_polygon(rng, n):
    a = np.sort(rng.uniform(0, 2 * math.pi, n))
    r = rng.uniform(1.0, 3.0)
    x = np.cos(a) * r * rng.uniform(0.5, 2.0) + rng.uniform(-50, 50)
    y = np.sin(a) * r + rng.uniform(-50, 50)
    return strictly_convex_ccw(x, y)


def strictly_convex_ccw(x, y):
    pts = sorted(set(zip(x.tolist(), y.tolist())))

    def cross(o, a, b):
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
    h = np.array(lower[:-1] + upper[:-1])
    return h[:, 0].copy(), h[:, 1].copy()


def antipodal_by_definition(x, y, i, j):
    n = len(x)

    def cone(k):
        # Outward normals of vertex k's two edges; a CCW edge (dx, dy) has
        # outward normal (dy, -dx), at angle atan2(-dx, dy).
        e0 = (x[k] - x[k - 1], y[k] - y[k - 1])
        e1 = (x[(k + 1) % n] - x[k], y[(k + 1) % n] - y[k])
        return math.atan2(-e0[0], e0[1]), math.atan2(-e1[0], e1[1])

    def inside(a, lo, hi, eps=1e-12):
        span = (hi - lo) % (2 * math.pi)
        t = (a - lo) % (2 * math.pi)
        return t <= span + eps or t >= 2 * math.pi - eps

    lo_i, hi_i = cone(i)
    lo_j, hi_j = cone(j)
    lo_j, hi_j = lo_j + math.pi, hi_j + math.pi
    return (inside(lo_i, lo_j, hi_j) or inside(hi_i, lo_j, hi_j) or inside(lo_j, lo_i, hi_i)
            or inside(hi_j, lo_i, hi_i))


def test_antipodal_pairs_known_answers(oracle):
    # Traced by hand through polygonizer.cpp:93-161.  The walk starts its
    # pairs at vertex 1 and stops when i reaches j0, and exact area ties add
    # the (i, next(j)) pair, which for a triangle is a vertex with itself.
    sq = oracle.antipodal_pairs(np.array([0.0, 1.0, 1.0, 0.0]), np.array([0.0, 0.0, 1.0, 1.0]))
    assert sq.tolist() == [[1, 2], [1, 3], [1, 0], [2, 3], [2, 0], [2, 1]]
    tri = oracle.antipodal_pairs(np.array([0.0, 1.0, 0.0]), np.array([0.0, 0.0, 1.0]))
    assert tri.tolist() == [[1, 2], [1, 0], [1, 1], [2, 0], [2, 1], [2, 2]]
    assert len(oracle.antipodal_pairs(np.array([0.0]), np.array([0.0]))) == 0


@pytest.mark.parametrize("seed", range(6))
def test_antipodal_pairs_are_antipodal(oracle, seed):
    rng = np.random.default_rng(seed)
    x, y = convex_polygon(rng, 40)
    pairs = oracle.antipodal_pairs(x, y)
    assert len(pairs) >= len(x) // 2
    for i, j in pairs:
        assert antipodal_by_definition(x, y, i, j), (i, j)
    # The walk advances monotonically: i never decreases, j moves forward.
    assert np.all(np.diff(pairs[:, 0]) >= 0)


def candidate_edges(oracle, x, y):
    n = len(x)
    out = []
    for i, j in oracle.antipodal_pairs(x, y):
        for k in (i, j):
            for off in (-1, 1):
                out.append((k, (k + off) % n))
    return out


def min_area_over_edges(x, y, edges=None):
    best = math.inf
    n = len(x)
    for a, b in edges if edges is not None else [(k, (k + 1) % n) for k in range(n)]:
        ex, ey = x[b] - x[a], y[b] - y[a]
        L = math.hypot(ex, ey)
        u = np.array([ex, ey]) / L
        v = np.array([-u[1], u[0]])
        P = np.stack([x, y], 1)
        a = (np.ptp(P @ u)) * (np.ptp(P @ v))
        best = min(best, a)
    return best


@pytest.mark.parametrize("seed", range(8))
def test_box_is_the_minimum_area_rectangle(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    x, y = convex_polygon(rng, int(rng.integers(5, 60)))
    b = oracle.obb_rotating_calipers(x, y)
    assert b["is_valid"]
    # The candidates are the edges next to the antipodal vertices; the box
    # is the smallest of them, never smaller than the best over all edges.
    ref = min_area_over_edges(x, y, candidate_edges(oracle, x, y))
    assert abs(float(b["area"]) - ref) <= 1e-6 * ref
    assert ref >= min_area_over_edges(x, y) * (1 - 1e-12)
    c = b["corners"]
    e1, e2 = c[1] - c[0], c[3] - c[0]
    assert abs(np.dot(e1, e2)) <= 1e-9 * np.linalg.norm(e1) * np.linalg.norm(e2)
    assert abs(np.linalg.norm(e1) * np.linalg.norm(e2) - ref) <= 1e-6 * ref
    # Every vertex inside the box.
    u, v = e1 / np.linalg.norm(e1), e2 / np.linalg.norm(e2)
    P = np.stack([x, y], 1) - c[0]
    su, sv = P @ u, P @ v
    tol = 1e-9 * (1 + np.abs(c).max())
    assert su.min() >= -tol and sv.min() >= -tol
    assert su.max() <= np.linalg.norm(e1) + tol and sv.max() <= np.linalg.norm(e2) + tol
    # The angle is the direction of the chosen edge (atan2Approx error < 1e-4).
    ang = float(b["angle_rad"])
    assert abs(math.atan2(u[1], u[0]) - ang) < 1e-4 or abs(abs(math.atan2(u[1], u[0]) - ang) - 2 * math.pi) < 1e-4


@pytest.mark.parametrize("theta", [0.0, 0.3, 1.0, 2.5])
def test_box_of_a_rectangle_is_the_rectangle(oracle, theta):
    w, h = 4.5, 1.8
    cx, cy = 12.0, -7.0
    loc = np.array([[-w / 2, -h / 2], [w / 2, -h / 2], [w / 2, h / 2], [-w / 2, h / 2]])
    R = np.array([[math.cos(theta), -math.sin(theta)], [math.sin(theta), math.cos(theta)]])
    P = loc @ R.T + [cx, cy]
    x, y = strictly_convex_ccw(P[:, 0], P[:, 1])
    b = oracle.obb_rotating_calipers(x, y)
    assert b["is_valid"]
    assert abs(float(b["area"]) - w * h) < 1e-5
    got = sorted(map(tuple, np.round(b["corners"], 9)))
    want = sorted(map(tuple, np.round(P, 9)))
    assert np.allclose(got, want, atol=1e-9)


def test_box_needs_three_vertices(oracle):
    assert not oracle.obb_rotating_calipers(np.array([0.0, 1.0]), np.array([0.0, 1.0]))["is_valid"]
