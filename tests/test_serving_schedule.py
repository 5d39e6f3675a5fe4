"""The simulated sensor arrivals (CPU)."""
import numpy as np

from lidar_amd import serving


def test_schedule_rate_jitter_and_order():
    sch = serving.arrival_schedule(3.0, seed=1)
    assert len(sch) == 8 * 30
    t = np.array([a[0] for a in sch])
    assert np.all(np.diff(t) >= 0)
    for s in range(8):
        ts = np.array([a[0] for a in sch if a[1] == s])
        ks = [a[2] for a in sch if a[1] == s]
        assert ks == list(range(30))
        gaps = np.diff(ts)
        assert np.all(gaps > 0.09 - 1e-9) and np.all(gaps < 0.11 + 1e-9)
    assert sch == serving.arrival_schedule(3.0, seed=1)


def test_sensor_mix():
    assert sorted(serving.SENSORS) == [32, 32, 32, 32, 64, 64, 64, 128]
