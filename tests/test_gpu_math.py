"""Device restatements of the C library functions against the host C
library, bit for bit (GPU)."""
import numpy as np
import pytest

from lidar_amd import native as N

pytestmark = pytest.mark.gpu

CHUNK = 1 << 22


def _bits(v):
    return int(np.array([v], np.float32).view(np.uint32)[0])


@pytest.mark.parametrize("start", [-5.0, -4.0, -2.0, -1.0, -0.5, -0.1, -1e-3, -1e-6])
def test_device_expf_on_jcp_domain(oracle, start):
    """JCP evaluates expf(-5 * dist) for dist in [0, 1]: contiguous runs of
    4M consecutive floats starting at several points of [-5, 0]."""
    lo = _bits(start)
    ref = oracle.host_math_range(0, lo, CHUNK)
    assert N.math_check_range(N.MATH_EXPF, lo, ref) == 0


@pytest.mark.parametrize("start", [1e-30, 1e-8, 0.3, 0.4375, 0.6875, 1.1875, 2.4375, 100.0, 3e7,
                                   -1e-8, -0.5, -1.2, -3.0, -1e5])
def test_device_atanf_reduction_branches(oracle, start):
    lo = _bits(start)
    ref = oracle.host_math_range(1, lo, CHUNK)
    assert N.math_check_range(N.MATH_ATANF, lo, ref) == 0


def test_device_atan2f_and_atan2_approx(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(7)
    y = (rng.standard_normal(200000) * 50).astype(np.float32)
    x = (rng.standard_normal(200000) * 50).astype(np.float32)
    y[:50] = 0.0
    x[50:100] = 0.0
    x[100:150] = 1.0
    dev = N.math_eval(N.MATH_ATAN2F, y, x)
    ref = np.array([L.lpo_host_atan2f(float(a), float(b)) for a, b in zip(y[:20000], x[:20000])], np.float32)
    assert np.array_equal(dev[:20000].view(np.uint32), ref.view(np.uint32))
    host_copy = N.math_eval(N.MATH_ATAN2F, y, x, on_device=False)
    assert np.array_equal(dev.view(np.uint32), host_copy.view(np.uint32))
    ap = N.math_eval(N.MATH_ATAN2_APPROX, y, x)
    apr = np.array([L.lpo_atan2_approx(float(a), float(b)) for a, b in zip(y[:20000], x[:20000])], np.float32)
    assert np.array_equal(ap[:20000].view(np.uint32), apr.view(np.uint32))
