"""lidar_amd -- MI355X-native lidar_processing_lib hot path.

The compute lives in liblp_amd.so (HIP kernels for gfx950 behind the C ABI of
include/lp_amd.h); this package is the Python-side binding used by the tests,
the benchmark and the serving loop.
"""
from . import native  # noqa: F401
from .native import (Context, LidarError, clusterer_config, segmenter_config,  # noqa: F401
                     LABEL_GROUND, LABEL_OBSTACLE, LABEL_UNKNOWN)
