"""PointCloud2 field dispatch (Processor::run, processor.cpp:494-549, and
convert<PointT>, processor.cpp:42-179) as zero-copy lp_cloud views.

The CPU tests check the dispatch rules through the C ABI (no GPU work is
involved); the GPU test segments through the view and compares with the
same points passed as arrays."""
import ctypes

import numpy as np
import pytest

from lidar_amd import native as N
from lidar_amd import synth

LAYOUTS = {
    "xyz": ([("x", 0), ("y", 4), ("z", 8)], 16),
    "xyzi": ([("x", 0), ("y", 4), ("z", 8), ("intensity", 16)], 32),
    "xyzr": ([("x", 0), ("y", 4), ("z", 8), ("ring", 12)], 16),
    "xyzir": ([("x", 0), ("y", 4), ("z", 8), ("intensity", 16), ("ring", 20)], 32),
}


def message(x, y, z, ring, fields, step):
    n = len(x)
    buf = np.zeros(n * step, np.uint8)
    rec = buf.reshape(n, step)
    for name, off in fields:
        src = {"x": x, "y": y, "z": z, "intensity": np.full(n, 0.5, np.float32), "ring": ring}[name]
        rec[:, off:off + src.itemsize] = np.ascontiguousarray(src).view(np.uint8).reshape(n, src.itemsize)
    return buf


def read_back(c, dtype, n):
    """Values the view addresses, read on the host."""
    def col(p, stride, dt):
        if not p:
            return None
        return np.array([ctypes.cast(p + i * stride, ctypes.POINTER(dt)).contents.value for i in range(n)])
    return (col(c.x, c.xyz_stride, ctypes.c_float), col(c.y, c.xyz_stride, ctypes.c_float),
            col(c.z, c.xyz_stride, ctypes.c_float), col(c.ring, c.ring_stride, ctypes.c_uint16))


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_dispatch_by_field_names(layout):
    fields, step = LAYOUTS[layout]
    rng = np.random.default_rng(3)
    n = 37
    x, y, z = (rng.normal(0, 10, n).astype(np.float32) for _ in range(3))
    ring = rng.integers(0, 64, n).astype(np.uint16)
    buf = message(x, y, z, ring, fields, step)
    st, c = N.pointcloud2_cloud(buf, fields, n, 1, step, n * step)
    assert st == 0
    assert c.num_points == n and c.xyz_stride == step and c.on_device == 0
    gx, gy, gz, gr = read_back(c, None, n)
    assert np.array_equal(gx, x) and np.array_equal(gy, y) and np.array_equal(gz, z)
    if "r" in layout[3:]:
        assert np.array_equal(gr, ring)
    else:
        assert gr is None


def test_organized_rows_and_errors():
    fields, step = LAYOUTS["xyzir"]
    buf = np.zeros(64 * 8 * step, np.uint8)
    st, c = N.pointcloud2_cloud(buf, fields, 8, 64, step, 8 * step)
    assert st == 0 and c.num_points == 512
    # padded rows are not packed
    st, _ = N.pointcloud2_cloud(buf, fields, 8, 32, step, 8 * step + 4)
    assert st == N.ERR_INVALID_ARGUMENT
    # the reference's unsupported formats: other names, other order, extra fields
    for bad in ([("x", 0), ("y", 4)], [("y", 0), ("x", 4), ("z", 8)], [("x", 0), ("y", 4), ("z", 8), ("rgb", 12)],
                [("x", 0), ("y", 4), ("z", 8), ("ring", 12), ("intensity", 16)]):
        st, _ = N.pointcloud2_cloud(buf, bad, 8, 1, step, 8 * step)
        assert st == N.ERR_INVALID_ARGUMENT
        assert "Unsupported point format" in N.lib().lp_last_error().decode()
    # a field outside the point
    st, _ = N.pointcloud2_cloud(buf, [("x", 0), ("y", 4), ("z", 30)], 8, 1, step, 8 * step)
    assert st == N.ERR_INVALID_ARGUMENT
    # empty message
    st, c = N.pointcloud2_cloud(np.zeros(0, np.uint8), fields, 0, 0, step, 0)
    assert st == 0 and c.num_points == 0


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["xyzir", "xyz"])
def test_segment_through_pointcloud2_view(layout):
    f = synth.workload_c1(2)
    fields, step = LAYOUTS[layout]
    buf = message(f.x, f.y, f.z, f.ring, fields, step)
    st, cloud = N.pointcloud2_cloud(buf, fields, f.n, 1, step, f.n * step)
    assert st == 0
    cfg = N.segmenter_config()
    with N.Context(max_points=f.n) as ctx:
        st1, lab_view, _ = ctx.segment(None, None, None, cfg=cfg, cloud=cloud)
        ring = f.ring if "r" in layout[3:] else None
        st2, lab_arr, _ = ctx.segment(f.x, f.y, f.z, ring, cfg=cfg)
    assert st1 == 0 and st2 == 0
    assert np.array_equal(lab_view, lab_arr)
