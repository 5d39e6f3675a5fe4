"""Allocation-free ring partitioning (GPU): lp_add_ring_info_noalloc against
lp_add_ring_info, through the C ABI, on device arrays.

Two calls in a row, the second on a smaller cloud, so the shared device
scratch array holds the first call's tile prefix when the second starts.
Device memory for the inputs and outputs comes from the HIP runtime the
library itself runs on (hipMalloc through ctypes, plumbing only): a second
HIP runtime in the process, such as the one torch bundles, would not share
the library's streams and device state.
"""
import ctypes
import math

import numpy as np
import pytest

from lidar_amd import native as N

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / streams of the library's own HIP runtime (the
    loaded libamdhip64.so.7 its dependency resolved to)."""

    H2D, D2H = 1, 2

    def __init__(self):
        N.lib()
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def check(self, e):
        assert e == 0, f"HIP error {e}"

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check(self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(nbytes, 1))))
        self.bufs.append(p)
        return p.value

    def upload(self, a):
        d = self.malloc(a.nbytes)
        self.check(self.h.hipMemcpy(ctypes.c_void_p(d), a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes),
                                    self.H2D))
        return d

    def fill(self, d, byte, nbytes):
        self.check(self.h.hipMemset(ctypes.c_void_p(d), byte, ctypes.c_size_t(nbytes)))

    def download(self, d, a):
        self.check(self.h.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(d), ctypes.c_size_t(a.nbytes),
                                    self.D2H))
        return a

    def stream(self):
        s = ctypes.c_void_p()
        self.check(self.h.hipStreamCreate(ctypes.byref(s)))
        return s

    def sync(self, s):
        self.check(self.h.hipStreamSynchronize(s))

    def close(self, streams=()):
        for s in streams:
            self.h.hipStreamDestroy(s)
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


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
    hip = _Hip()
    f = _entry()
    stream = hip.stream()
    try:
        with N.Context(max_frames=1, max_points=300000) as ctx:
            for beams, per_beam, seed in ((80, 3000, 1), (23, 1500, 2)):
                x, y = _scan(beams, per_beam, seed)
                n = len(x)
                print(f"noalloc ring info: {n} points ({beams} beams)", flush=True)
                st, want = ctx.add_ring_info(x, y)
                assert st == N.OK
                xd = hip.upload(x)
                yd = hip.upload(y)
                rd = hip.malloc(2 * n)
                hip.fill(rd, 0xFF, 2 * n)
                assert f(xd, yd, n, rd, stream) == N.OK
                hip.sync(stream)
                got = hip.download(rd, np.empty(n, np.uint16))
                assert np.array_equal(got, want)
                # More than 63 fourth-to-first steps: the rings saturate at 0.
                assert (beams < 64) or got[-1] == 0
    finally:
        hip.close([stream])


def test_noalloc_ring_info_rejects_without_launching():
    hip = _Hip()
    f = _entry()
    xd = hip.malloc(64)
    rd = hip.malloc(32)
    try:
        # Beyond the scratch array's 16384 tiles of 1024 points.
        assert f(xd, xd, 16777217, rd, None) == N.ERR_CAPACITY
        assert b"16777216" in N.lib().lp_last_error()
        # Host arrays are refused, not dereferenced on the device.
        h = np.zeros(16, np.float32)
        hr = np.zeros(16, np.uint16)
        assert f(h.ctypes.data, h.ctypes.data, 16, hr.ctypes.data, None) == N.ERR_INVALID_ARGUMENT
        assert f(None, None, 0, None, None) == N.OK
        assert f(None, None, -1, None, None) == N.ERR_INVALID_ARGUMENT
    finally:
        hip.close()
