"""Seeded synthetic LiDAR frames (ctypes binding of liblp_synth.so).

Workload presets restate the benchmark configurations: an HDL-64E-like
64x2048 scan (15 cars, 20 poles, 2 walls, 10% dropout, 200 below-ground
returns), the full-pipeline variant with 40 cars, 60 poles and 2000 rain
points, the uneven batched 64-beam set, the 128x2048 dense-urban set, and the
32/64/128-beam serving sensors.  Seeds follow seed(c, f) =
0x9E3779B97F4A7C15*(c+1) + f.
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class SynthSpec(ctypes.Structure):
    _fields_ = [
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("fov_up_deg", ctypes.c_float),
        ("fov_down_deg", ctypes.c_float),
        ("beam_layout", ctypes.c_int32),
        ("dropout", ctypes.c_float),
        ("n_cars", ctypes.c_int32),
        ("n_poles", ctypes.c_int32),
        ("n_walls", ctypes.c_int32),
        ("facades", ctypes.c_int32),
        ("obj_range_min", ctypes.c_float),
        ("obj_range_max", ctypes.c_float),
        ("n_below", ctypes.c_int32),
        ("n_rain", ctypes.c_int32),
        ("max_range", ctypes.c_float),
        ("enclosure_radius", ctypes.c_float),
        ("seed", ctypes.c_uint64),
    ]


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblp_synth.so")
        if not os.path.exists(path):
            raise RuntimeError("liblp_synth.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.lp_synth_frame.argtypes = [ctypes.POINTER(SynthSpec), ctypes.c_int32, P, P, P, P, P]
        L.lp_synth_frame.restype = ctypes.c_int32
        _LIB = L
    return _LIB


@dataclass
class Frame:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    intensity: np.ndarray
    ring: np.ndarray
    height: int
    width: int
    fov_up_deg: float
    fov_down_deg: float

    @property
    def n(self):
        return len(self.x)


def seed_for(config_index, frame_index):
    return (0x9E3779B97F4A7C15 * (config_index + 1) + frame_index) & 0xFFFFFFFFFFFFFFFF


def make_frame(height=64, width=2048, fov_up=2.0, fov_down=-24.8, layout=0, dropout=0.10, cars=15,
               poles=20, walls=2, facades=0, rmin=5.0, rmax=60.0, below=200, rain=0, max_range=120.0,
               enclosure=92.0, seed=1):
    spec = SynthSpec(height, width, fov_up, fov_down, layout, dropout, cars, poles, walls, facades, rmin, rmax,
                     below, rain, max_range, enclosure, seed)
    cap = height * width + rain + 16
    x = np.empty(cap, np.float32)
    y = np.empty(cap, np.float32)
    z = np.empty(cap, np.float32)
    it = np.empty(cap, np.float32)
    rg = np.empty(cap, np.uint16)
    n = _lib().lp_synth_frame(ctypes.byref(spec), cap, *(a.ctypes.data_as(ctypes.c_void_p) for a in (x, y, z, it, rg)))
    if n < 0:
        raise RuntimeError("synth capacity exceeded")
    return Frame(x[:n].copy(), y[:n].copy(), z[:n].copy(), it[:n].copy(), rg[:n].copy(), height, width, fov_up,
                 fov_down)


def workload_c1(frame_index=0):
    """64x2048 HDL-64E-like frame, 15 cars, 20 poles, 2 walls, 10% dropout."""
    return make_frame(seed=seed_for(0, frame_index))


def workload_c2(frame_index=0):
    """Full pipeline frame: 40 cars, 60 poles, 2000 rain/snow points."""
    return make_frame(cars=40, poles=60, dropout=0.04, rain=2000, seed=seed_for(1, frame_index))


def workload_c3(frame_index):
    """Batched set: W in {1024,1800,2048}, dropout 5-30%."""
    rng = np.random.default_rng(seed_for(2, frame_index) & 0xFFFFFFFF)
    width = int(rng.choice([1024, 1800, 2048]))
    dropout = float(rng.uniform(0.05, 0.30))
    return make_frame(width=width, dropout=dropout, seed=seed_for(2, frame_index))


def workload_c4(frame_index):
    """128x2048 dense urban: 80 cars, 150 poles, facades; +15..-25 deg uniform beams."""
    return make_frame(height=128, width=2048, fov_up=15.0, fov_down=-25.0, layout=1, dropout=0.03, cars=80,
                      poles=150, walls=0, facades=1, rmin=5.0, rmax=60.0, seed=seed_for(3, frame_index))


def workload_sensor(kind, frame_index, sensor_index=0):
    """Serving sensors: kind in {32, 64, 128} beams."""
    seed = seed_for(4, sensor_index * 100003 + frame_index)
    if kind == 32:
        return make_frame(height=32, width=1024, fov_up=10.67, fov_down=-30.67, layout=1, dropout=0.05, seed=seed)
    if kind == 64:
        return make_frame(dropout=0.10, seed=seed)
    return make_frame(height=128, width=2048, fov_up=15.0, fov_down=-25.0, layout=1, dropout=0.05, cars=40,
                      poles=80, facades=1, seed=seed)
