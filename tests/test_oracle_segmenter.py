"""Pins for the segmenter restatement (CPU only).

Each test checks the oracle against something other than its own formula:
hand-derived known answers on constructed clouds, a textbook operation
(scipy's binary dilation), an independent RNG restatement, or invariants
the reference's arithmetic guarantees.  The JCP cases are built so that the
two easy mistakes -- reading pending neighbours as resolved (Jacobi instead
of the reference's in-order updates, segmenter.cpp:535-637) and clearing the
24 weight/label slots per pixel instead of once per frame
(segmenter.cpp:523-525) -- give a different label.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

W, H = 2048, 64
UNKNOWN, GROUND, OBSTACLE = 0, 1, 2


def xy_at(col, dist):
    """A point whose image column is `col` (the column is (W-1)*az/2pi)."""
    az = (col + 0.5) * 2.0 * math.pi / (W - 1)
    return dist * math.cos(az), dist * math.sin(az)


def build(points):
    """points: list of (ring, col, dist, z) -> x, y, z, ring arrays."""
    x, y, z, r = [], [], [], []
    for ring, col, dist, zz in points:
        px, py = xy_at(col, dist)
        x.append(px)
        y.append(py)
        z.append(zz)
        r.append(ring)
    return (np.array(x, np.float32), np.array(y, np.float32), np.array(z, np.float32), np.array(r, np.uint16))


def segment(oracle, pts, **cfg):
    x, y, z, r = build(pts) if not isinstance(pts, tuple) else pts
    st, lab, dbg = oracle.segment(x, y, z, r, cfg=oracle.seg_config(**cfg), debug=True)
    return st, lab, dbg


def test_polar_grid_filters_and_pixels(oracle):
    pts = [
        (5, 0, 10.0, -1.73),     # kept: pixel (5, 0)
        (7, 511, 10.0, -1.73),   # kept: pixel (7, 511)
        (9, 100, 10.0, -5.0),    # z below -1.73 - 3.0: dropped
        (9, 100, 1.5, -1.73),    # closer than 2 m: dropped
        (9, 100, 101.0, -1.73),  # beyond 100 m: dropped
        (64, 100, 10.0, -1.73),  # ring >= image height: dropped
    ]
    st, lab, dbg = segment(oracle, pts)
    assert st == 0
    assert list(lab[2:]) == [UNKNOWN] * 4
    assert lab[0] != UNKNOWN and lab[1] != UNKNOWN
    m = dbg["mapping"].reshape(H, W)
    assert m[5, 0] == 0 and m[7, 511] == 1
    assert (m >= 0).sum() == 2


def _equal_depth_pair():
    """Two float32 points with identical x*x+y*y in image column 255, one in
    azimuth slice 44 and one in slice 45 (column 255 spans 44.85..45.02 deg)."""
    r = 10.0
    a1, a2 = math.radians(44.9), math.radians(45.01)
    x1, y1 = np.float32(r * math.cos(a1)), np.float32(r * math.sin(a1))
    d1 = np.float32(x1 * x1 + y1 * y1)
    x2, y2 = np.float32(r * math.cos(a2)), np.float32(r * math.sin(a2))
    k = np.arange(-64, 65, dtype=np.float32)
    xs = (x2 + k * np.spacing(x2)).astype(np.float32)[:, None]
    ys = (y2 + k * np.spacing(y2)).astype(np.float32)[None, :]
    hit = np.argwhere((xs * xs + ys * ys).astype(np.float32) == d1)
    assert len(hit), "no equal-depth pair found"
    i, j = hit[0]
    return (x1, y1), (xs[i, 0], ys[0, j])


def test_pixel_winner_ties_follow_cell_order(oracle):
    (x1, y1), (x2, y2) = _equal_depth_pair()
    # Input order: the slice-1 point first.  The reference visits cells in
    # index order, so the slice-0 point (input index 1) wins the tie.
    x = np.array([x2, x1], np.float32)
    y = np.array([y2, y1], np.float32)
    z = np.array([-1.73, -1.73], np.float32)
    ring = np.array([3, 3], np.uint16)
    st, lab, dbg = oracle.segment(x, y, z, ring, debug=True)
    m = dbg["mapping"].reshape(H, W)
    assert (m >= 0).sum() == 1
    assert m[m >= 0][0] == 1
    assert lab[0] == UNKNOWN and lab[1] != UNKNOWN
    # Exact duplicates in one cell: the earlier input index wins.
    x = np.array([x1, x1], np.float32)
    y = np.array([y1, y1], np.float32)
    st, lab, dbg = oracle.segment(x, y, z, ring, debug=True)
    assert dbg["mapping"][dbg["mapping"] >= 0][0] == 0
    assert lab[0] != UNKNOWN and lab[1] == UNKNOWN


def test_only_pixel_winners_are_labelled(oracle):
    pts = [(10, 700, 20.0, -1.73), (10, 700, 20.5, -1.73)]
    st, lab, _ = segment(oracle, pts)
    assert lab[0] == GROUND and lab[1] == UNKNOWN


def _cell_of(col, dist):
    x, y = xy_at(col, dist)
    az = math.atan2(y, x) % (2 * math.pi)
    return int(az / (math.pi / 180)) * 50 + int(dist / 2)


def test_recm_robust_minimum_known_answers(oracle):
    # One cell (radial 5) with a low outlier: sorted z = [-3.0, -1.75, -1.72,
    # -1.70]; the gap test (i from n/2 down to 1) fires at i=1 and the cell
    # minimum becomes -1.75 instead of -3.0.
    col = 1500
    pts = [(20 + k, col, 10.5, zz) for k, zz in enumerate([-3.0, -1.75, -1.72, -1.70])]
    st, lab, dbg = segment(oracle, pts)
    E = dbg["elevation_map"]
    c = _cell_of(col, 10.5)
    assert E[c] == np.float32(-1.75)
    # A gap in the upper half moves the minimum up to s[i] = 0.5, so the
    # elevation is capped by the slope limit instead: E[r] = E[r-1] + dz.
    pts = [(20 + k, col, 10.5, zz) for k, zz in enumerate([-1.75, -1.74, 0.5, 0.6])]
    st, lab, dbg = segment(oracle, pts)
    E = dbg["elevation_map"]
    dz = np.float32(0.199999884)
    assert E[c] == np.float32(E[c - 1] + dz)
    assert E[c] != np.float32(-1.75)


def test_recm_invariants_on_a_scan(oracle, synth):
    f = synth.workload_c1(3)
    st, lab, dbg = oracle.segment(f.x, f.y, f.z, f.ring, debug=True)
    E = dbg["elevation_map"].reshape(360, 50)
    dz = np.float32(0.199999884)
    assert np.all(E[:, 0] == np.float32(np.float32(-1.73) + np.float32(0.2)))
    assert np.all(E[:, 1:] <= (E[:, :-1] + dz).astype(np.float32))


def _ransac_draws(n):
    from tests.test_oracle_math import lemire, mt19937
    raw = mt19937(42, 2000)
    draws = lemire(raw, n, 1000)
    out, k = [], 0
    for _ in range(60):
        a = draws[k]
        b = draws[k + 1]
        k += 2
        while b == a:
            b = draws[k]
            k += 1
        out += [a, b]
    return out


def test_ransac_draws_and_plane_invariants(oracle, synth):
    f = synth.workload_c1(1)
    st, lab, dbg = oracle.segment(f.x, f.y, f.z, f.ring, debug=True)
    n = dbg["ransac_candidates"]
    assert n > 1000 and dbg["ransac_applied"] == 1
    assert list(dbg["ransac_draws"]) == _ransac_draws(n)
    a, b, c, d = dbg["ransac_plane"]
    cos_max = np.float32(math.cos(np.float32(math.tan(np.float32(0.2)))))
    assert c >= cos_max
    assert d == np.float32(c * np.float32(-1.73))
    assert abs(float(a) ** 2 + float(b) ** 2 + float(c) ** 2 - 1.0) < 1e-5
    assert 0 < dbg["ransac_best_inliers"] <= n


def test_ransac_single_candidate_and_none(oracle):
    # One point in the near cells (radial < 4) within 0.4 m of the elevation
    # map: the reference's redraw loop cannot finish; the oracle reports it.
    near = [(2, 300, 5.0, -1.73)]
    far = [(30, 900, 20.0, -1.73), (31, 900, 20.0, -1.70)]
    st, lab, dbg = segment(oracle, near + far)
    assert st == oracle.ERR_RANSAC_SINGLE
    st, lab, dbg = segment(oracle, far)
    assert st == 0 and dbg["ransac_candidates"] == 0 and dbg["ransac_applied"] == 0


def test_jcp_queue_is_textbook_dilation(oracle, synth):
    f = synth.workload_c2(2)
    st, lab, dbg = oracle.segment(f.x, f.y, f.z, f.ring, debug=True)
    s = dbg["state_recm"].reshape(H, W)
    grown = ndimage.binary_dilation(s == OBSTACLE, structure=np.ones((5, 5), bool))
    expect = np.flatnonzero(((s == GROUND) & grown).ravel())
    assert np.array_equal(dbg["queue"], expect)


def test_jcp_in_order_updates(oracle):
    """G2 (row 9) is queued before G0 (row 10).  G2 sees only the obstacle O1
    (its other near neighbour G0 is still pending) and turns OBSTACLE.  G0
    then weighs G4 (ground, u~0.54) against G2 (now obstacle, u~0.54) plus
    O1 (u~0.12) and turns OBSTACLE; had it read G2 as pending it would stay
    GROUND (0.12 < 0.54)."""
    pts = [
        (10, 600, 20.0, -1.73),  # G0
        (9, 602, 20.0, -1.73),   # G2
        (11, 602, 20.0, -1.33),  # O1
        (10, 598, 20.0, -1.73),  # G4: outside O1's 5x5 window, never queued
    ]
    st, lab, dbg = segment(oracle, pts)
    assert st == 0
    assert list(lab) == [OBSTACLE, OBSTACLE, OBSTACLE, GROUND]
    assert len(dbg["queue"]) == 2


def test_jcp_stale_border_slots(oracle):
    """B sits in image column 0, so its ten slots with column offsets -1/-2
    are outside the image and keep the values last written by the previous
    queue pixel A (row 20), whose slot (0,-2) holds the obstacle OA
    (u~0.34).  B's own weights: only GB (ground, u~0.18), so sum = 0.18 and
    the stale obstacle weight 0.34/0.18 beats 1.0 -> OBSTACLE.  With the
    slots cleared per pixel B would be GROUND."""
    pts = [
        (20, 5, 8.2, -1.73),    # A
        (20, 3, 8.2, -1.52),    # OA (obstacle, 2 columns left of A)
        (30, 0, 50.0, -1.73),   # B (column 0)
        (28, 2, 50.0, -1.58),   # GB (ground, never queued)
        (32, 2, 50.0, 1.5),     # OB: queues B but is > 1 m from it
    ]
    st, lab, dbg = segment(oracle, pts)
    assert st == 0
    assert list(lab) == [OBSTACLE, OBSTACLE, OBSTACLE, GROUND, OBSTACLE]


def test_jcp_unresolved_pixel_stays_unknown(oracle):
    """P is queued (obstacle OP next to it) but every neighbour is farther
    than 1 m, so its weight sum is zero: it keeps the pending colour and its
    point stays UNKNOWN although it won its pixel."""
    pts = [(40, 1000, 20.0, -1.73), (41, 1000, 20.0, 1.27), (42, 1000, 20.0, 1.30)]
    st, lab, dbg = segment(oracle, pts)
    assert st == 0
    assert list(lab) == [UNKNOWN, OBSTACLE, OBSTACLE]
    assert dbg["state_final"].reshape(H, W)[40, 1000] == 3


def test_empty_cloud(oracle):
    z = np.zeros(0, np.float32)
    st, lab, _ = oracle.segment(z, z, z, np.zeros(0, np.uint16))
    assert st == 0 and len(lab) == 0


def test_add_ring_info_closed_form(oracle):
    """Rings count down from 63 on every fourth-to-first quadrant step
    (dataloader.cpp:86-122): ring_i = 63 - min(63, transitions up to i)."""
    rng = np.random.default_rng(5)
    beams = 70
    az = np.concatenate([np.sort(rng.uniform(0.01, 2 * math.pi - 0.01, 300)) for _ in range(beams)])
    x = (20 * np.cos(az)).astype(np.float32)
    y = (20 * np.sin(az)).astype(np.float32)
    ring = oracle.add_ring_info(x, y)
    quad = np.floor(az / (math.pi / 2)).astype(int)
    trans = np.concatenate([[0], ((quad[1:] == 0) & (quad[:-1] == 3)).astype(int)])
    expect = 63 - np.minimum(63, np.cumsum(trans))
    assert np.array_equal(ring, expect.astype(np.uint16))
