"""Record layouts staged through the device (GPU).

AoS clouds are copied record by record span into a device staging area and
de-interleaved there.  The copy must read nothing past the last field of
the last record (the caller's array may end there), and clouds larger than
the staging area must go through in chunks.  Both are checked against the
SoA path and the oracle.
"""
import numpy as np
import pytest

from lidar_amd import native as N
from lidar_amd import synth

pytestmark = pytest.mark.gpu


def _records(f, dtype):
    rec = np.zeros(f.n, dtype=dtype)
    rec["x"], rec["y"], rec["z"] = f.x, f.y, f.z
    if "ring" in dtype.names:
        rec["ring"] = f.ring
    return rec


def test_leading_field_before_x(oracle):
    """{t, x, y, z, ring}: x sits 8 bytes into each 24-byte record, and the
    cloud ends at the last record's ring field (no trailing padding)."""
    f = synth.workload_c1(4)
    dt = np.dtype({"names": ["t", "x", "y", "z", "ring"], "formats": ["<f8", "<f4", "<f4", "<f4", "<u2"],
                   "offsets": [0, 8, 12, 16, 20], "itemsize": 24})
    rec = _records(f, dt)
    # The view handed over starts at the first record and the buffer is cut
    # right after the last ring field.
    raw = rec.view(np.uint8)[: (f.n - 1) * 24 + 22].copy()
    base = raw.ctypes.data
    cloud = N.Cloud()
    cloud.x, cloud.y, cloud.z, cloud.ring = base + 8, base + 12, base + 16, base + 20
    cloud.xyz_stride = cloud.ring_stride = 24
    cloud.num_points = f.n
    cloud.on_device = 0
    with N.Context(max_frames=1, max_points=f.n) as ctx:
        st, lab, _ = ctx.segment(None, None, None, cloud=cloud)
        st2, lab2, _ = ctx.segment(f.x, f.y, f.z, f.ring)
    _, ref, _ = oracle.segment(f.x, f.y, f.z, f.ring)
    assert st == 0 and st2 == 0
    assert np.array_equal(lab, lab2) and np.array_equal(lab, ref)


def test_records_larger_than_the_staging_area(oracle):
    """A 16x256 image leaves a staging area of 16*256*96 bytes; 100k records
    of 32 bytes (3.2 MB) go through in chunks."""
    f = synth.workload_c1(5)
    n = min(f.n, 100000)
    x, y, z = f.x[:n], f.y[:n], f.z[:n]
    ring = (f.ring[:n] % 16).astype(np.uint16)
    dt = np.dtype({"names": ["x", "y", "z", "intensity", "ring"], "formats": ["<f4", "<f4", "<f4", "<f4", "<u2"],
                   "offsets": [0, 4, 8, 16, 20], "itemsize": 32})
    rec = np.zeros(n, dtype=dt)
    rec["x"], rec["y"], rec["z"], rec["ring"] = x, y, z, ring
    cfg = dict(image_height=16, image_width=256)
    with N.Context(max_frames=1, max_points=n, max_height=16, max_width=256) as ctx:
        st, lab, _ = ctx.segment(None, None, None, cfg=N.segmenter_config(**cfg), cloud=N.aos_cloud(rec))
    _, ref, _ = oracle.segment(x, y, z, ring, cfg=oracle.seg_config(**cfg))
    assert st == 0
    assert np.array_equal(lab, ref)
