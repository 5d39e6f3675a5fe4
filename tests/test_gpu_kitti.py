"""Real data on the GPU (GPU): the reference's sample KITTI scan through
ring partitioning (Dataloader::addRingInfo), the full segment -> cluster ->
hull path, and the DROR + boxes variant, against the oracle."""
import numpy as np
import pytest

from lidar_amd import native as N

from tests.test_gpu_noise_boxes import box_equal, full_frame_reference
from tests.test_gpu_parity import hull_coords
from tests.test_kitti_fixture import kitti

pytestmark = pytest.mark.gpu


class Frame:
    def __init__(self, x, y, z, ring):
        self.x, self.y, self.z, self.ring = x, y, z, ring
        self.n = len(x)


def test_kitti_bit_exact(oracle):
    x, y, z = kitti()
    with N.Context(max_points=len(x)) as ctx:
        st, ring = ctx.add_ring_info(x, y)
        assert st == 0
        assert np.array_equal(ring, oracle.add_ring_info(x, y))
        sco, scg = oracle.seg_config(), N.segmenter_config()
        out = ctx.process([(x, y, z, ring)], scg)[0]
        ref = oracle.process_frame(x, y, z, ring, scfg=sco)
        assert out["status"] == 0
        assert np.array_equal(out["seg_labels"], ref["seg_labels"])
        assert np.array_equal(out["cluster_labels"], ref["cluster_labels"])
        obs = out["obstacle_index"]
        ox, oy = x[obs].astype(np.float64), y[obs].astype(np.float64)
        gh = hull_coords(ox, oy, out["cluster_labels"], out["num_clusters"], out["hull_offsets"], out["hull_indices"])
        rh = hull_coords(ox, oy, ref["cluster_labels"], ref["num_clusters"], ref["hull_offsets"], ref["hull_indices"])
        assert len(gh) == len(rh) and all(np.array_equal(a, b) for a, b in zip(gh, rh))

        ctx.set_noise_filter(N.noise_config())
        ctx.set_oriented_boxes(True)
        out = ctx.process([(x, y, z, ring)], scg)[0]
        noise, (kx, ky), ref, hulls, boxes = full_frame_reference(oracle, Frame(x, y, z, ring), sco)
        assert out["status"] == 0
        assert np.array_equal(out["noise_labels"], noise)
        assert np.array_equal(out["seg_labels"], ref["seg_labels"])
        assert np.array_equal(out["cluster_labels"], ref["cluster_labels"])
        obs = out["obstacle_index"]
        gh = hull_coords(kx[obs].astype(np.float64), ky[obs].astype(np.float64), out["cluster_labels"],
                         out["num_clusters"], out["hull_offsets"], out["hull_indices"])
        assert all(np.array_equal(a, b) for a, b in zip(gh, hulls))
        assert all(box_equal(a, b) for a, b in zip(out["boxes"], boxes))
