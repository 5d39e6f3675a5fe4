"""Real data: the reference's sample KITTI scan (tests/golden, written by
make_kitti_fixture.py from data/0000000000.pcd) through the oracle, and the
PCD reader (CPU)."""
import os

import numpy as np

from lidar_amd.pcd import read_pcd

HERE = os.path.dirname(os.path.abspath(__file__))


FIXTURE = os.path.join(HERE, "golden", "kitti_0000000000_xyz.npz")


def kitti():
    """The KITTI scan (see tests/golden/README.md); tests skip without it."""
    import pytest
    if not os.path.exists(FIXTURE):
        pytest.skip("KITTI fixture not present")
    d = np.load(FIXTURE)
    return d["x"], d["y"], d["z"]


def test_pcd_reader_binary_and_ascii(tmp_path):
    pts = np.zeros(3, np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("intensity", "<f4")]))
    pts["x"], pts["y"], pts["z"], pts["intensity"] = [1.5, -2.0, 3.25], [0.5, 7.0, -1.0], [0.0, 1.0, 2.0], [0.1, 0.2, 0.3]
    head = ("# .PCD v0.7 - Point Cloud Data file format\nVERSION 0.7\nFIELDS x y z intensity\nSIZE 4 4 4 4\n"
            "TYPE F F F F\nCOUNT 1 1 1 1\nWIDTH 3\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\nPOINTS 3\n")
    b = tmp_path / "b.pcd"
    b.write_bytes((head + "DATA binary\n").encode() + pts.tobytes())
    got = read_pcd(b)
    assert got.dtype.names == ("x", "y", "z", "intensity") and np.array_equal(got, pts)
    a = tmp_path / "a.pcd"
    a.write_text(head + "DATA ascii\n" + "\n".join(" ".join(repr(float(v)) for v in p) for p in pts) + "\n")
    assert np.array_equal(read_pcd(a), pts)


def test_kitti_ring_partitioning(oracle):
    x, y, z = kitti()
    assert len(x) == 123398
    ring = oracle.add_ring_info(x, y)
    # Scan order: rings never increase, start at 63 and use (nearly) all 64.
    assert ring[0] == 63 and np.all(np.diff(ring.astype(int)) <= 0)
    assert len(np.unique(ring)) >= 60


def test_kitti_full_frame_through_the_oracle(oracle):
    x, y, z = kitti()
    ring = oracle.add_ring_info(x, y)
    r = oracle.process_frame(x, y, z, ring)
    assert r["status"] == 0
    seg = r["seg_labels"]
    assert 0.3 < np.mean(seg == 1) < 0.9 and 0.05 < np.mean(seg == 2) < 0.6
    assert 50 < r["num_clusters"] < 1000
