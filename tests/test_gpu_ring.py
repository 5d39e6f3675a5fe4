"""Ring partitioning of scan-ordered clouds on the GPU vs the oracle (GPU)."""
import math

import numpy as np
import pytest

from lidar_amd import native as N

pytestmark = pytest.mark.gpu


def test_ring_info_matches_oracle(oracle, synth):
    f = synth.workload_c1(2)
    # Re-order the scan like a KITTI file: top beam first, azimuth sweeping.
    order = np.lexsort((np.arctan2(f.y, f.x), -f.ring.astype(int)))
    x, y = f.x[order], f.y[order]
    with N.Context(max_points=f.n) as ctx:
        st, ring = ctx.add_ring_info(x, y)
    assert st == 0
    assert np.array_equal(ring, oracle.add_ring_info(x, y))


def test_ring_info_quadrant_edges(oracle):
    # Points exactly on the quadrant boundaries and a long sweep (> 64 turns).
    a = np.concatenate([np.linspace(0, 2 * math.pi, 97)[:-1] for _ in range(70)])
    x = np.cos(a).astype(np.float32) * 10
    y = np.sin(a).astype(np.float32) * 10
    x[:4] = [10.0, 0.0, -10.0, 0.0]
    y[:4] = [0.0, 10.0, 0.0, -10.0]
    with N.Context(max_points=len(x)) as ctx:
        st, ring = ctx.add_ring_info(x, y)
    assert st == 0
    assert np.array_equal(ring, oracle.add_ring_info(x, y))
    assert ring.min() == 0
