"""The C ABI library loads, exports every function include/lp_amd.h
declares, and its configuration defaults equal the reference's
(segmenter.hpp:91-108, clusterer.hpp:61-68, processor.param.yaml:31-35).
No GPU needed: only host-side entry points are called."""
import os
import re

import numpy as np

from lidar_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lp_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w]+\s*\**\s*(lp_\w+)\s*\(", text, flags=re.M)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 20
    lib = N.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(N.EXPORTED) == names


def test_segmenter_defaults_match_reference():
    c = N.segmenter_config()
    assert c.elevation_up_deg == np.float32(2.0)
    assert c.elevation_down_deg == np.float32(-24.8)
    assert (c.image_width, c.image_height, c.assume_unorganized_cloud) == (2048, 64, 0)
    expect = dict(grid_radial_spacing_m=2.0, grid_slice_resolution_deg=1.0, ground_height_threshold_m=0.2,
                  road_maximum_slope_m_per_m=0.2, min_distance_m=2.0, max_distance_m=100.0, sensor_height_m=1.73,
                  kernel_threshold_distance_m=1.0, amplification_factor=5.0, z_min_m=-3.0, z_max_m=4.0)
    for k, v in expect.items():
        assert getattr(c, k) == np.float32(v), k


def test_clusterer_defaults_match_reference():
    h = N.clusterer_config(shipped=False)
    assert (h.voxel_grid_range_resolution_m, h.voxel_grid_azimuth_resolution_deg) == (np.float32(0.4), 1.0)
    assert h.voxel_grid_elevation_resolution_deg == np.float32(1.5) and h.min_cluster_size == 3
    s = N.clusterer_config()
    assert s.voxel_grid_elevation_resolution_deg == np.float32(3.0)


def test_capacity_defaults_match_reference_reserves():
    cap = N.Capacity()
    N.lib().lp_capacity_default(cap)
    assert cap.max_points_per_frame == 200000 and cap.max_voxels == 200000


def test_host_math_entry_needs_no_gpu():
    out = N.math_eval(N.MATH_ATAN2_APPROX, np.array([1.0], np.float32), np.array([0.0], np.float32),
                      on_device=False)
    assert out[0] == np.float32(1.57079637)
    assert N.lib().lp_math_eval(7, None, None, None, 0, 0) == N.ERR_INVALID_ARGUMENT


def test_invalid_arguments_are_reported_not_crashing():
    assert N.lib().lp_context_create(None, None) == N.ERR_INVALID_ARGUMENT
    assert b"null" in N.lib().lp_last_error()
    assert N.lib().lp_batch_run(None, 0) == N.ERR_INVALID_ARGUMENT
