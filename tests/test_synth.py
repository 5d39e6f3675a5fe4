"""The seeded scan generator: deterministic, organized, and shaped like the
benchmark configurations."""
import numpy as np


def test_deterministic_and_seed_sensitive(synth):
    a = synth.workload_c1(0)
    b = synth.workload_c1(0)
    c = synth.workload_c1(1)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.ring, b.ring)
    assert a.n != c.n or not np.array_equal(a.x, c.x)


def test_c1_shape(synth):
    f = synth.workload_c1(0)
    assert (f.height, f.width) == (64, 2048)
    # 64 x 2048 rays with 10% dropout, ~118k returns.
    assert 110_000 < f.n < 125_000
    assert f.ring.max() == 63 and f.ring.min() == 0
    assert np.all(np.diff(f.ring.astype(int)) >= 0)  # ring-major order


def test_c2_has_rain_points(synth):
    f = synth.workload_c2(0)
    assert f.n > 120_000
    rain = f.n - 2000
    d = np.hypot(f.x[rain:], f.y[rain:])
    assert d.max() < 30.5


def test_c3_widths_vary(synth):
    widths = {synth.workload_c3(i).width for i in range(12)}
    assert widths <= {1024, 1800, 2048} and len(widths) >= 2


def test_c4_shape(synth):
    f = synth.workload_c4(0)
    assert (f.height, f.width) == (128, 2048)
    assert 230_000 < f.n <= 262_144
