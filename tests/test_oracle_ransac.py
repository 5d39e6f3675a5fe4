"""Independent numpy restatement of the RANSAC ground correction (CPU only).

`correctCloseRangeFalsePositivesRANSAC` (lidar_processing_lib/src/segmenter.cpp:321-479)
is recomputed here from scratch and compared with the oracle's debug
outputs:

* the candidate list: cells `slice*50 + r`, r < 4, in cell order and input
  order inside a cell, with `fabs(E[cell] - z) < 2.0f * thr` (:326-351);
* the 60 hypotheses from the index pairs of `mt19937{42}` +
  `uniform_int_distribution` (pinned separately by a pure-Python MT19937
  and Lemire's mapping), the normal in the reference's operand order, the
  `den^2 < 1e-5f` and `fabs(nz) < cosf(tanf(slope))` rejections (:388-418);
* the inlier count with a strict `<` against thr (:421-431), the strict `>`
  argmax that keeps the earliest best hypothesis (:434-442);
* the sign flip when c < 0 (:445-453) and the reclassification of every
  point of the r < 4 cells with `(a*x + b*y + c*z) - d < thr` to GROUND
  (:455-477), observed through the pixel states after RECM + RANSAC.

The inputs it takes from the oracle are the elevation map (pinned by the
RECM known answers and invariants) and the pixel winners (pinned by the
brute-force argmin checks).  Every arithmetic step is float32 with the
reference's association (numpy float32 scalar and array operations round
each operation to single precision; no contraction).  glibc's `tanf` and
`cosf` come straight from libm, and `atan2Approx` from the oracle's copy
that `tests/test_oracle_math.py` pins against its error bound and special
values.
"""
import ctypes
import math

import numpy as np
import pytest

from tests.test_oracle_segmenter import _ransac_draws

F = np.float32
UNKNOWN, GROUND, OBSTACLE = 0, 1, 2

_LIBM = ctypes.CDLL("libm.so.6")
for _fn in ("tanf", "cosf", "atanf"):
    getattr(_LIBM, _fn).argtypes = [ctypes.c_float]
    getattr(_LIBM, _fn).restype = ctypes.c_float

DEG_TO_RAD = F(math.pi / 180.0)     # static_cast<float>(M_PI / 180.0), segmenter.hpp:135
TWO_PI = F(2.0 * math.pi)           # static_cast<float>(2.0 * M_PI), segmenter.hpp:136


def polar_cells(oracle, x, y, z, ring, cfg):
    """Cell index of every point the polar grid keeps (-1 otherwise),
    segmenter.cpp:103-204 with the organized (ring) row."""
    h, zlo, zhi = F(cfg.sensor_height_m), F(cfg.z_min_m), F(cfg.z_max_m)
    z_lo = F(-h + zlo)
    z_hi = F(-h + zhi)
    spacing = F(cfg.grid_radial_spacing_m)
    rings = int(F(cfg.max_distance_m) / spacing)
    res = F(F(cfg.grid_slice_resolution_deg) * DEG_TO_RAD)
    slices = int(TWO_PI / res)
    d = np.sqrt((x * x + y * y).astype(F)).astype(F)
    keep = ~((z < z_lo) | (z > z_hi))
    keep &= ~((d < F(cfg.min_distance_m)) | (d > F(cfg.max_distance_m)))
    radial = np.where(keep, (d / spacing).astype(F), F(0)).astype(np.int64)
    keep &= radial < rings
    keep &= ring < cfg.image_height
    cell = np.full(len(x), -1, np.int64)
    at = oracle.lib().lpo_atan2_approx
    for i in np.flatnonzero(keep):
        az = F(at(float(y[i]), float(x[i])))
        if az < 0:
            az = F(az + TWO_PI)
        s = min(int(F(az / res)), slices - 1)
        cell[i] = s * rings + radial[i]
    return cell, rings


def ransac_numpy(x, y, z, cell, rings, E, n_slices, thr, h, slope):
    """Returns (candidate indices, best inliers, plane a,b,c,d or None, per
    hypothesis inlier counts)."""
    near = (cell >= 0) & (cell % rings < 4)
    idx = np.flatnonzero(near)
    order = idx[np.argsort(cell[idx], kind="stable")]       # cell order, input order inside a cell
    z_min = E[cell[order]]
    cand = order[np.abs((z_min - z[order]).astype(F)) < F(F(2.0) * thr)]
    n = len(cand)
    if n == 0:
        return cand, 0, None, []
    draws = _ransac_draws(n)
    X, Y, Z = x[cand], y[cand], z[cand]
    p1x, p1y, p1z = F(0.0), F(0.0), F(-h)
    cos_max = F(_LIBM.cosf(_LIBM.tanf(float(slope))))
    best, plane, counts = 0, None, []
    for it in range(60):
        i2, i3 = draws[2 * it], draws[2 * it + 1]
        p2 = (X[i2], Y[i2], Z[i2])
        p3 = (X[i3], Y[i3], Z[i3])
        nx = F(F(F(p2[1] - p1y) * F(p3[2] - p1z)) - F(F(p2[2] - p1z) * F(p3[1] - p1y)))
        ny = F(F(F(p2[2] - p1z) * F(p3[0] - p1x)) - F(F(p2[0] - p1x) * F(p3[2] - p1z)))
        nz = F(F(F(p2[0] - p1x) * F(p3[1] - p1y)) - F(F(p2[1] - p1y) * F(p3[0] - p1x)))
        den = F(F(F(nx * nx) + F(ny * ny)) + F(nz * nz))
        if den < F(1.0e-5):
            counts.append(None)
            continue
        norm = F(F(1.0) / np.sqrt(den))
        nz = F(nz * norm)
        if abs(nz) < cos_max:
            counts.append(None)
            continue
        nx = F(nx * norm)
        ny = F(ny * norm)
        d = F(F(F(nx * p1x) + F(ny * p1y)) + F(nz * p1z))
        dist = np.abs((((nx * X).astype(F) + (ny * Y).astype(F)).astype(F) + (nz * Z).astype(F)).astype(F) - d)
        cnt = int(np.count_nonzero(dist.astype(F) < thr))
        counts.append(cnt)
        if cnt > best:
            best, plane = cnt, [nx, ny, nz, d]
    if best > 0 and plane[2] < 0:
        plane = [F(-v) for v in plane]
    return cand, best, plane, counts


def expected_states(x, y, z, cell, rings, E, mapping, thr, plane):
    """Pixel states after RECM's obstacle test (segmenter.cpp:269-283) and the
    RANSAC override of the near cells."""
    st = np.zeros(len(mapping), np.uint8)
    px = np.flatnonzero(mapping >= 0)
    i = mapping[px]
    obstacle = z[i] >= (E[cell[i]] + thr).astype(F)
    lab = np.where(obstacle, OBSTACLE, GROUND).astype(np.uint8)
    if plane is not None:
        a, b, c, d = plane
        sd = (((a * x[i]).astype(F) + (b * y[i]).astype(F)).astype(F) + (c * z[i]).astype(F)).astype(F) - d
        flip = (cell[i] % rings < 4) & (sd.astype(F) < thr)
        lab[flip] = GROUND
    st[px] = lab
    return st


def _frames(synth):
    return [("c1", synth.workload_c1(1)), ("c1", synth.workload_c1(3)), ("c2", synth.workload_c2(0)),
            ("c4", synth.workload_c4(0))]


@pytest.mark.parametrize("k", range(4))
def test_ransac_matches_numpy_restatement(oracle, synth, k):
    name, f = _frames(synth)[k]
    cfg = oracle.seg_config(image_height=f.height, image_width=f.width, elevation_up_deg=f.fov_up_deg,
                            elevation_down_deg=f.fov_down_deg)
    st, lab, dbg = oracle.segment(f.x, f.y, f.z, f.ring, cfg=cfg, debug=True)
    assert st == 0
    thr, h, slope = F(cfg.ground_height_threshold_m), F(cfg.sensor_height_m), F(cfg.road_maximum_slope_m_per_m)
    cell, rings = polar_cells(oracle, f.x, f.y, f.z, f.ring, cfg)
    E = dbg["elevation_map"]
    cand, best, plane, counts = ransac_numpy(f.x, f.y, f.z, cell, rings, E, 360, thr, h, slope)
    assert len(cand) == dbg["ransac_candidates"]
    assert best == dbg["ransac_best_inliers"]
    assert dbg["ransac_applied"] == int(best > 0)
    assert np.array_equal(np.array(plane, F).view(np.uint32), dbg["ransac_plane"].view(np.uint32))
    # The candidate window is 2*thr: points with thr <= |E - z| < 2*thr are in
    # the list, so an oracle with the window at thr would count fewer.
    near = (cell >= 0) & (cell % rings < 4)
    gap = np.abs((E[cell[near]] - f.z[near]).astype(F))
    assert np.count_nonzero((gap >= thr) & (gap < F(2.0) * thr)) > 0
    # Several hypotheses were scored, some rejected by the slope limit.
    scored = [c for c in counts if c is not None]
    assert len(scored) >= 10 and best == max(scored)
    want = expected_states(f.x, f.y, f.z, cell, rings, E, dbg["mapping"], thr, plane)
    assert np.array_equal(want, dbg["state_recm"])
    # The override changed some near-cell pixels from RECM obstacle to ground.
    if name in ("c1", "c2"):
        recm_only = expected_states(f.x, f.y, f.z, cell, rings, E, dbg["mapping"], thr, None)
        assert np.count_nonzero(recm_only != want) > 0


def _tie_cloud():
    """Eight ground points in near cells (r = 1..3) on the plane z = -1.73,
    each a few centimetres off it, so most hypotheses through the pivot and
    two of them keep all eight points within thr and tie at 8 inliers with
    different planes; plus two far points that give the frame an image."""
    rng = np.random.default_rng(3)
    az = rng.uniform(0.1, 2 * math.pi - 0.1, 8)
    dist = rng.uniform(2.5, 7.5, 8)
    x = (dist * np.cos(az)).astype(F)
    y = (dist * np.sin(az)).astype(F)
    z = (-1.73 + rng.uniform(-0.03, 0.03, 8)).astype(F)
    ring = np.full(8, 10, np.uint16)
    x = np.concatenate([x, np.array([20.0, 20.0], F)])
    y = np.concatenate([y, np.array([1.0, 1.1], F)])
    z = np.concatenate([z, np.array([-1.73, -1.72], F)])
    ring = np.concatenate([ring, np.array([30, 31], np.uint16)])
    return x, y, z, ring


def test_ransac_ties_keep_the_first_best(oracle):
    x, y, z, ring = _tie_cloud()
    cfg = oracle.seg_config()
    st, lab, dbg = oracle.segment(x, y, z, ring, cfg=cfg, debug=True)
    assert st == 0
    thr, h, slope = F(0.2), F(1.73), F(0.2)
    cell, rings = polar_cells(oracle, x, y, z, ring, cfg)
    cand, best, plane, counts = ransac_numpy(x, y, z, cell, rings, dbg["elevation_map"], 360, thr, h, slope)
    assert len(cand) == 8 and best == 8
    tied = [i for i, c in enumerate(counts) if c == best]
    assert len(tied) >= 2, "the constructed cloud must produce tied hypotheses"
    assert np.array_equal(np.array(plane, F).view(np.uint32), dbg["ransac_plane"].view(np.uint32))
    # The last of the tied hypotheses (what a >= update would keep) is a
    # different plane, so the comparison above separates > from >=.
    last = _hypothesis_plane(x[cand], y[cand], z[cand], tied[-1], h)
    if last[2] < 0:
        last = [F(-v) for v in last]
    assert not np.array_equal(np.array(last, F).view(np.uint32), dbg["ransac_plane"].view(np.uint32))


def _hypothesis_plane(X, Y, Z, it, h):
    draws = _ransac_draws(len(X))
    i2, i3 = draws[2 * it], draws[2 * it + 1]
    p1z = F(-h)
    nx = F(F(Y[i2] * F(Z[i3] - p1z)) - F(F(Z[i2] - p1z) * Y[i3]))
    ny = F(F(F(Z[i2] - p1z) * X[i3]) - F(X[i2] * F(Z[i3] - p1z)))
    nz = F(F(X[i2] * Y[i3]) - F(Y[i2] * X[i3]))
    norm = F(F(1.0) / np.sqrt(F(F(F(nx * nx) + F(ny * ny)) + F(nz * nz))))
    nx, ny, nz = F(nx * norm), F(ny * norm), F(nz * norm)
    return [nx, ny, nz, F(F(F(nx * F(0)) + F(ny * F(0))) + F(nz * p1z))]
