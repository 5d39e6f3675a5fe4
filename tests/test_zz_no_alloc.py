"""Allocation-free ring partitioning (GPU): lp_add_ring_info_noalloc against
lp_add_ring_info, through the C ABI, on device arrays.

Two calls in a row, the second on a smaller cloud, so the shared device
scratch array holds the first call's tile prefix when the second starts.
Device memory for the inputs and outputs comes from torch (plumbing only).
"""
import ctypes
import math

import numpy as np
import pytest

from lidar_amd import native as N

pytestmark = pytest.mark.gpu


def _entry():
    f = N.lib().lp_add_ring_info_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int
    return f


def _scan(beams, per_beam, seed):
    """A scan-ordered cloud: `beams` sweeps of ascending azimuth."""
    rng = np.random.default_rng(seed)
    az = np.concatenate([np.sort(rng.uniform(0.001, 2 * math.pi - 0.001, per_beam)) for _ in range(beams)])
    r = rng.uniform(3.0, 60.0, len(az))
    return (r * np.cos(az)).astype(np.float32), (r * np.sin(az)).astype(np.float32)


def test_noalloc_ring_info_matches_context_entry_twice():
    import torch
    f = _entry()
    stream = torch.cuda.current_stream().cuda_stream
    with N.Context(max_frames=1, max_points=300000) as ctx:
        for beams, per_beam, seed in ((80, 3000, 1), (23, 1500, 2)):
            x, y = _scan(beams, per_beam, seed)
            n = len(x)
            print(f"noalloc ring info: {n} points ({beams} beams)", flush=True)
            st, want = ctx.add_ring_info(x, y)
            assert st == N.OK
            xd = torch.from_numpy(x).cuda()
            yd = torch.from_numpy(y).cuda()
            rd = torch.full((n,), -1, dtype=torch.int16, device="cuda")
            assert f(xd.data_ptr(), yd.data_ptr(), n, rd.data_ptr(), stream) == N.OK
            torch.cuda.synchronize()
            got = rd.cpu().numpy().view(np.uint16)
            assert np.array_equal(got, want)
            # More than 63 fourth-to-first steps: the rings saturate at 0.
            assert (beams < 64) or got[-1] == 0


def test_noalloc_ring_info_rejects_without_launching():
    import torch
    f = _entry()
    xd = torch.zeros(16, dtype=torch.float32, device="cuda")
    rd = torch.zeros(16, dtype=torch.int16, device="cuda")
    # Beyond the scratch array's 16384 tiles of 1024 points.
    assert f(xd.data_ptr(), xd.data_ptr(), 16777217, rd.data_ptr(), None) == N.ERR_CAPACITY
    assert b"16777216" in N.lib().lp_last_error()
    # Host arrays are refused, not dereferenced on the device.
    h = np.zeros(16, np.float32)
    hr = np.zeros(16, np.uint16)
    assert f(h.ctypes.data, h.ctypes.data, 16, hr.ctypes.data, None) == N.ERR_INVALID_ARGUMENT
    assert f(None, None, 0, None, None) == N.OK
    assert f(None, None, -1, None, None) == N.ERR_INVALID_ARGUMENT
