"""Pins for the clusterer and hull restatements (CPU only).

Clusterer: the oracle's BFS (clusterer.cpp:122-193) must equal an
independent brute-force connected-components computation (scipy) over the
same voxel adjacency, numbered by each component's minimum point index with
small components removed (clusterer.cpp:195-239); plus known answers for the
azimuth wrap rule and the voxel capacity error.

Hull: Andrew's chain (polygonizer.cpp:33-91) must return Qhull's vertex set
(scipy.spatial.ConvexHull), counter-clockwise, starting at the
lexicographically smallest point; degenerate inputs follow the chain's own
rules (collinear -> [min, max], identical -> two entries, n < 3 -> iota).
"""
import math

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components
from scipy.spatial import ConvexHull

RES_R, RES_AZ, RES_EL = np.float32(0.4), np.float32(1.0) * np.float32(math.pi / 180), np.float32(3.0) * np.float32(
    math.pi / 180)


def brute_clusters(oracle, x, y, z, min_size=3):
    """Voxel keys with host libm (the oracle library's atan2f/atanf), then
    components by brute-force pair tests."""
    L = oracle.lib()
    n = len(x)
    az = np.array([L.lpo_host_atan2f(float(b), float(a)) for a, b in zip(x, y)], np.float32)
    az = np.where(az < 0, (az + np.float32(2 * math.pi)).astype(np.float32), az).astype(np.float32)
    dsq = (x * x + y * y).astype(np.float32)
    d = np.sqrt(dsq).astype(np.float32)
    rng = np.sqrt((dsq + z * z).astype(np.float32)).astype(np.float32)
    el = (np.array([L.lpo_host_atanf(float(v)) for v in (z / d).astype(np.float32)], np.float32)
          + np.float32(math.pi / 2)).astype(np.float32)
    nr = int(np.float32(math.ceil(np.float32(rng.max() / RES_R))) + 1)
    na = int(np.float32(math.ceil(np.float32(az.max() / RES_AZ))) + 1)
    ri = (rng / RES_R).astype(np.float32).astype(np.int64)
    ai = (az / RES_AZ).astype(np.float32).astype(np.int64)
    ei = (el / RES_EL).astype(np.float32).astype(np.int64)
    keys = np.stack([ri, ai, ei], 1)
    uniq, inv = np.unique(keys, axis=0, return_inverse=True)
    inv = inv.ravel()
    U = len(uniq)
    rows, cols = [], []
    for a in range(U):
        for b in range(a + 1, U):
            dr = abs(uniq[a, 0] - uniq[b, 0])
            de = abs(uniq[a, 2] - uniq[b, 2])
            da = abs(uniq[a, 1] - uniq[b, 1])
            # azimuth neighbours wrap by one step of num_azimuth
            da_ok = da <= 1 or da == na - 1
            if dr <= 1 and de <= 1 and da_ok:
                rows.append(a)
                cols.append(b)
    g = coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(U, U))
    _, comp = connected_components(g, directed=False)
    pc = comp[inv]
    first = {}
    for i, c in enumerate(pc):
        first.setdefault(c, i)
    counts = np.bincount(pc)
    order = sorted((i, c) for c, i in first.items() if counts[c] >= min_size)
    rank = {c: k for k, (_, c) in enumerate(order)}
    return np.array([rank.get(c, -1) for c in pc], np.int32), nr


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_clusters_equal_brute_force_components(oracle, seed):
    rng = np.random.default_rng(seed)
    n = 400
    blobs = rng.uniform(-30, 30, (8, 2))
    idx = rng.integers(0, 8, n)
    x = (blobs[idx, 0] + rng.normal(0, 0.6, n)).astype(np.float32)
    y = (blobs[idx, 1] + rng.normal(0, 0.6, n)).astype(np.float32)
    z = rng.uniform(-1.5, 1.0, n).astype(np.float32)
    st, lab = oracle.cluster(x, y, z)
    assert st == 0
    expect, _ = brute_clusters(oracle, x, y, z)
    assert np.array_equal(lab, expect)


def test_azimuth_wrap_joins_first_and_last_column(oracle):
    """The azimuth wrap connects index 0 with index num_azimuth-1 (clusterer.cpp:
    161-169).  With the largest azimuth exactly pi (180 bins of 1 degree),
    num_azimuth = 181, so a point at azimuth pi (index 180) and one at
    azimuth 0.5 deg (index 0) fall in neighbouring voxels even though they
    are on opposite sides of the sensor."""
    x = np.array([-10.0, -10.0, -10.0, 10.0, 10.0, 10.0], np.float32)
    y = np.array([0.0, 0.0, 0.0, 0.0873, 0.0873, 0.0873], np.float32)
    z = np.array([0.0, 0.05, 0.1, 0.0, 0.05, 0.1], np.float32)
    st, lab = oracle.cluster(x, y, z)
    assert st == 0
    assert list(lab) == [0] * 6
    # Moving the far-side points off the exact half turn breaks the wrap.
    y2 = y.copy()
    y2[:3] = -0.001
    st, lab = oracle.cluster(x, y2, z)
    assert sorted(set(lab.tolist())) == [0, 1]


def test_small_clusters_removed_and_renumbered(oracle):
    x = np.array([10, 10.01, 10.02, 30, 30.01, 50, 50.01, 50.02, 50.03], np.float32)
    y = np.zeros(9, np.float32) + 1
    z = np.zeros(9, np.float32)
    st, lab = oracle.cluster(x, y, z)
    assert list(lab) == [0, 0, 0, -1, -1, 1, 1, 1, 1]


def test_voxel_capacity_overflow(oracle):
    x = np.linspace(5, 60, 40).astype(np.float32)
    y = np.zeros(40, np.float32) + 1
    z = np.zeros(40, np.float32)
    st, _ = oracle.cluster(x, y, z, max_voxels=10)
    assert st == oracle.ERR_VOXEL_OVERFLOW
    st, _ = oracle.cluster(x, y, z, max_voxels=200)
    assert st == 0


def test_cluster_empty(oracle):
    e = np.zeros(0, np.float32)
    st, lab = oracle.cluster(e, e, e)
    assert st == 0 and len(lab) == 0


def ccw_from_min(x, y, idx):
    pts = np.stack([x[idx], y[idx]], 1)
    start = min(range(len(idx)), key=lambda k: (pts[k, 0], pts[k, 1]))
    return start, pts


@pytest.mark.parametrize("seed", range(6))
def test_hull_matches_qhull(oracle, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 400))
    x = rng.normal(0, 5, n).astype(np.float32).astype(np.float64)
    y = rng.normal(0, 2, n).astype(np.float32).astype(np.float64)
    h = oracle.convex_hull(x, y)
    q = ConvexHull(np.stack([x, y], 1))
    assert sorted(h.tolist()) == sorted(q.vertices.tolist())
    pts = np.stack([x[h], y[h]], 1)
    area2 = np.sum(pts[:, 0] * np.roll(pts[:, 1], -1) - np.roll(pts[:, 0], -1) * pts[:, 1])
    assert area2 > 0  # counter-clockwise
    lex = min(range(len(h)), key=lambda k: (pts[k, 0], pts[k, 1]))
    assert lex == 0  # starts at the lexicographically smallest point


def test_hull_degenerate_inputs(oracle):
    x = np.array([0.0, 1.0, 2.0, 3.0])
    y = np.array([0.0, 1.0, 2.0, 3.0])
    assert oracle.convex_hull(x, y).tolist() == [0, 3]
    x = np.array([5.0, 5.0, 5.0])
    y = np.array([1.0, 1.0, 1.0])
    h = oracle.convex_hull(x, y)
    assert len(h) == 2
    assert oracle.convex_hull(np.array([1.0, 2.0]), np.array([3.0, 4.0])).tolist() == [0, 1]
    # Collinear points on an edge are dropped.
    x = np.array([0.0, 1.0, 2.0, 2.0, 0.0])
    y = np.array([0.0, 0.0, 0.0, 1.0, 1.0])
    assert oracle.convex_hull(x, y).tolist() == [0, 2, 3, 4]
