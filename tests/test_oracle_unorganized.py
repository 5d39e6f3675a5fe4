"""Pins for the ring-less image row of the polar grid (CPU only).

Point types without a ring field (`PointXYZ`, `PointXYZI`) and organized
types with `assume_unorganized_cloud` take the image row from the elevation
angle instead of the ring (lidar_processing_lib/src/segmenter.cpp:155-193;
the `HasRing` trait, segmenter.hpp:53-65):

    row = static_cast<int32_t>((atanf(z / d) - down) / ((up - down) / H)),
    dropped unless 0 <= row < H.

The cast truncates toward zero, so an elevation less than one row below
`down` still lands in row 0.  The checks below are (1) hand-placed points
whose rows follow from their construction, including that quirk and both
ends of the [0, H) filter, and (2) a numpy restatement of the projection
with glibc's `atanf` called directly and a brute-force pixel winner,
compared with the oracle's pixel mapping on whole synthetic scans.
"""
import ctypes
import math

import numpy as np
import pytest

F = np.float32
_LIBM = ctypes.CDLL("libm.so.6")
_LIBM.atanf.argtypes = [ctypes.c_float]
_LIBM.atanf.restype = ctypes.c_float
DEG_TO_RAD = F(math.pi / 180.0)
TWO_PI = F(2.0 * math.pi)


def _geometry(H=64, up=2.0, down=-24.8):
    up_r = F(F(up) * DEG_TO_RAD)
    down_r = F(F(down) * DEG_TO_RAD)
    per = F(F(up_r - down_r) / F(H))
    return up_r, down_r, per


def _point(elev_rad, dist=5.0, az_deg=30.0):
    a = math.radians(az_deg)
    return (F(dist * math.cos(a)), F(dist * math.sin(a)), F(dist * math.tan(elev_rad)))


def _mapping_rows(dbg, H, W):
    m = dbg["mapping"].reshape(H, W)
    rows = {}
    for r, c in zip(*np.nonzero(m >= 0)):
        rows[int(m[r, c])] = int(r)
    return rows


@pytest.mark.parametrize("organized_flag", [False, True])
def test_rows_from_elevation_hand_placed(oracle, organized_flag):
    H = 64
    up, down, per = _geometry(H)
    cases = [
        (float(down) + 10.5 * float(per), 10),        # middle of row 10
        (float(down) + 0.5 * float(per), 0),          # middle of row 0
        (float(down) + 63.5 * float(per), 63),        # middle of the top row
        (float(down) - 0.5 * float(per), 0),          # half a row below: truncation toward zero keeps it
        (float(down) - 1.5 * float(per), None),       # one and a half rows below: row -1, dropped
        (float(up) + 0.5 * float(per), None),         # above the field of view: row 64, dropped
    ]
    x, y, z = [], [], []
    for k, (e, _) in enumerate(cases):
        px, py, pz = _point(e, az_deg=10.0 + 40.0 * k)
        x.append(px)
        y.append(py)
        z.append(pz)
    x, y, z = (np.array(v, F) for v in (x, y, z))
    if organized_flag:
        # Rings present but ignored: 200 would drop every point in organized mode.
        ring = np.full(len(x), 200, np.uint16)
        cfg = oracle.seg_config(assume_unorganized_cloud=1)
    else:
        ring = None
        cfg = oracle.seg_config()
    st, lab, dbg = oracle.segment(x, y, z, ring, cfg=cfg, debug=True)
    assert st == 0
    rows = _mapping_rows(dbg, H, 2048)
    for i, (_, want) in enumerate(cases):
        if want is None:
            assert i not in rows and lab[i] == 0
        else:
            assert rows.get(i) == want, (i, rows.get(i), want)
    # The same points with the rings honoured (organized, rings 200 >= H)
    # are all dropped.
    if organized_flag:
        st, lab, dbg = oracle.segment(x, y, z, ring, cfg=oracle.seg_config(), debug=True)
        assert (dbg["mapping"] >= 0).sum() == 0


def projection_numpy(oracle, x, y, z, H, W, up_deg, down_deg):
    """(pixel, cell, depth^2) per point, -1 pixel for dropped points, with the
    elevation row (segmenter.cpp:120-196)."""
    _, down, per = _geometry(H, up_deg, down_deg)
    z_lo, z_hi = F(F(-1.73) + F(-3.0)), F(F(-1.73) + F(4.0))
    d = np.sqrt((x * x + y * y).astype(F)).astype(F)
    keep = ~((z < z_lo) | (z > z_hi)) & ~((d < F(2.0)) | (d > F(100.0)))
    radial = np.where(keep, (d / F(2.0)).astype(F), F(0)).astype(np.int64)
    keep &= radial < 50
    pix = np.full(len(x), -1, np.int64)
    cell = np.full(len(x), -1, np.int64)
    at = oracle.lib().lpo_atan2_approx
    res = F(F(1.0) * DEG_TO_RAD)
    for i in np.flatnonzero(keep):
        elev = F(_LIBM.atanf(float(F(z[i] / d[i]))))
        row = int(F(F(elev - down) / per))      # int() truncates toward zero, like the C++ cast
        if row < 0 or row >= H:
            continue
        az = F(at(float(y[i]), float(x[i])))
        if az < 0:
            az = F(az + TWO_PI)
        col = int(F(F(F(W - 1) * az) / TWO_PI))
        pix[i] = row * W + col
        cell[i] = min(int(F(az / res)), 359) * 50 + radial[i]
    depth = (x * x + y * y).astype(F)
    return pix, cell, depth


def brute_force_winners(pix, cell, depth, P):
    """Per pixel: the smallest depth^2, ties to the earliest (cell, input)
    position (segmenter.cpp:289-318 visits cells in index order)."""
    idx = np.flatnonzero(pix >= 0)
    order = np.lexsort((idx, cell[idx], depth[idx], pix[idx]))
    s = idx[order]
    first = np.ones(len(s), bool)
    first[1:] = pix[s][1:] != pix[s][:-1]
    m = np.full(P, -1, np.int32)
    m[pix[s][first]] = s[first]
    return m


@pytest.mark.parametrize("which", ["c1", "c2", "c4"])
def test_ringless_projection_matches_numpy(oracle, synth, which):
    f = {"c1": lambda: synth.workload_c1(2), "c2": lambda: synth.workload_c2(1),
         "c4": lambda: synth.workload_c4(1)}[which]()
    cfg = oracle.seg_config(image_height=f.height, image_width=f.width, elevation_up_deg=f.fov_up_deg,
                            elevation_down_deg=f.fov_down_deg)
    st, lab, dbg = oracle.segment(f.x, f.y, f.z, None, cfg=cfg, debug=True)
    assert st == 0
    pix, cell, depth = projection_numpy(oracle, f.x, f.y, f.z, f.height, f.width, f.fov_up_deg, f.fov_down_deg)
    want = brute_force_winners(pix, cell, depth, f.height * f.width)
    assert np.array_equal(want, dbg["mapping"])
    # The elevation row differs from the beam index for some points, so the
    # ring-less image is not the organized one.
    st2, _, dbg2 = oracle.segment(f.x, f.y, f.z, f.ring, cfg=cfg, debug=True)
    assert not np.array_equal(dbg2["mapping"], dbg["mapping"])
    # assume_unorganized_cloud with the rings present is the ring-less branch.
    cfg.assume_unorganized_cloud = 1
    st3, lab3, dbg3 = oracle.segment(f.x, f.y, f.z, f.ring, cfg=cfg, debug=True)
    assert st3 == 0 and np.array_equal(dbg3["mapping"], dbg["mapping"]) and np.array_equal(lab3, lab)
