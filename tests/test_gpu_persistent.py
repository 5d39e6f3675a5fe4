"""Persistent (frame, tile) work-queue launches (GPU): the same outputs as
the grid launches, on batches of uneven frames, with graphs replayed."""
import numpy as np
import pytest

from lidar_amd import native as N
from lidar_amd import synth

pytestmark = pytest.mark.gpu

KEYS = ("seg_labels", "obstacle_index", "cluster_labels", "hull_offsets", "hull_indices")


def run(frames, persistent, B, streams=1):
    cfgs = [N.segmenter_config(image_height=f.height, image_width=f.width, elevation_up_deg=f.fov_up_deg,
                               elevation_down_deg=f.fov_down_deg) for f in frames]
    with N.Context(max_frames=B, max_points=max(f.n for f in frames), max_height=max(f.height for f in frames),
                   max_width=2048, streams=streams) as ctx:
        ctx.set_persistent(persistent)
        ctx.stage([(f.x, f.y, f.z, f.ring) for f in frames], cfgs)
        for _ in range(3):
            ctx.run(use_graph=True)
        ctx.synchronize()
        return ctx.fetch()


@pytest.mark.parametrize("count,streams", [(12, 1), (3, 1), (16, 2)])
def test_persistent_equals_grid(count, streams):
    frames = [synth.workload_c3(i) for i in range(count)]
    print(f"persistent vs grid: {count} frames, {streams} streams", flush=True)
    a = run(frames, False, count, streams)
    b = run(frames, True, count, streams)
    for fa, fb in zip(a, b):
        assert fa["status"] == 0 and fb["status"] == 0
        for k in KEYS:
            assert np.array_equal(fa[k], fb[k]), k
