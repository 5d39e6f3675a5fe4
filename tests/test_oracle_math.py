"""Pins for the arithmetic the hot path depends on (CPU only).

* MT19937: a pure-Python restatement of the published generator (Matsumoto &
  Nishimura 1998) is checked against the C++ standard's known answer
  ([rand.predef]: the 10000th output of a default-seeded mt19937 is
  4123659995), then against libstdc++'s std::mt19937{42} through the oracle.
* uniform_int_distribution<uint32_t>: a Python restatement of Lemire's
  nearly-divisionless mapping (libstdc++ >= 11 for a 32-bit engine) must
  reproduce std::uniform_int_distribution draws for many ranges, including
  ranges that force rejections.  The device RANSAC sampler uses this mapping.
* atan2Approx (common.hpp:33-62): its documented maximum relative error
  (3.6e-5, common.hpp:31) and exact special values.
* The libm restatements compiled for the host (same source as the device
  code) against the host C library on strided sweeps of their domains.
"""
import math

import numpy as np
import pytest


def mt19937(seed, count):
    mt = [0] * 624
    mt[0] = seed & 0xFFFFFFFF
    for i in range(1, 624):
        mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
    idx = 624
    out = []
    for _ in range(count):
        if idx >= 624:
            for i in range(624):
                y = (mt[i] & 0x80000000) | (mt[(i + 1) % 624] & 0x7FFFFFFF)
                mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            idx = 0
        y = mt[idx]
        idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        out.append(y & 0xFFFFFFFF)
    return out


def lemire(raw, n, count):
    it = iter(raw)
    out = []
    for _ in range(count):
        prod = next(it) * n
        low = prod & 0xFFFFFFFF
        if low < n:
            thr = (2**32 - n) % n
            while low < thr:
                prod = next(it) * n
                low = prod & 0xFFFFFFFF
        out.append(prod >> 32)
    return out


def test_mt19937_known_answer():
    assert mt19937(5489, 10000)[-1] == 4123659995


def test_mt19937_seed42_matches_libstdcxx(oracle):
    assert list(oracle.mt19937_raw(2000)) == mt19937(42, 2000)


@pytest.mark.parametrize("n", [2, 3, 5, 7, 60, 1000, 37911, 65536, 2**31, 2**31 + 1, 3_000_000_000, 2**32 - 1])
def test_lemire_matches_std_uniform_int_distribution(oracle, n):
    raw = mt19937(42, 4000)
    assert list(oracle.std_uniform_draws(n, 500)) == lemire(raw, n, 500)


def test_atan2_approx_error_bound_and_special_values(oracle):
    L = oracle.lib()
    assert L.lpo_atan2_approx(0.0, 1.0) == 0.0
    assert np.float32(L.lpo_atan2_approx(1.0, 0.0)) == np.float32(1.57079637)
    assert np.float32(L.lpo_atan2_approx(0.0, -1.0)) == np.float32(3.14159274)
    assert np.float32(L.lpo_atan2_approx(-1.0, 0.0)) == np.float32(-1.57079637)
    rng = np.random.default_rng(3)
    ang = rng.uniform(-math.pi, math.pi, 4000)
    rad = rng.uniform(0.5, 150.0, 4000)
    worst = 0.0
    for a, r in zip(ang, rad):
        y, x = np.float32(r * math.sin(a)), np.float32(r * math.cos(a))
        exact = math.atan2(float(y), float(x))
        if abs(exact) < 1e-3:
            continue
        got = L.lpo_atan2_approx(float(y), float(x))
        worst = max(worst, abs(got - exact) / abs(exact))
    # 3.6e-5 polynomial error plus float32 rounding of the evaluation.
    assert worst < 4.0e-5


def _strided(lo_bits, hi_bits, step):
    return np.arange(lo_bits, hi_bits, step, dtype=np.uint64).astype(np.uint32).view(np.float32)


def test_expf_restatement_matches_host_libm(oracle):
    from lidar_amd import native as N
    # Negative floats in [-20, -0] (JCP evaluates expf(-5 * dist), dist in [0, 1]).
    xs = _strided(0x80000000, 0xC1A00000, 4099)
    mine = N.math_eval(N.MATH_EXPF, xs, on_device=False)
    L = oracle.lib()
    ref = np.array([L.lpo_host_expf(float(v)) for v in xs[::7]], np.float32)
    assert np.array_equal(mine[::7].view(np.uint32), ref.view(np.uint32))


def test_atanf_atan2f_restatements_match_host_libm(oracle):
    from lidar_amd import native as N
    L = oracle.lib()
    xs = _strided(0, 0xFFFFFFFF, 65537)
    xs = xs[np.isfinite(xs)]
    mine = N.math_eval(N.MATH_ATANF, xs, on_device=False)
    ref = np.array([L.lpo_host_atanf(float(v)) for v in xs], np.float32)
    assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))
    rng = np.random.default_rng(4)
    y = (rng.standard_normal(30000) * 60).astype(np.float32)
    x = (rng.standard_normal(30000) * 60).astype(np.float32)
    y[:100] = 0.0
    x[100:200] = 0.0
    x[200:300] = 1.0
    mine2 = N.math_eval(N.MATH_ATAN2F, y, x, on_device=False)
    ref2 = np.array([L.lpo_host_atan2f(float(a), float(b)) for a, b in zip(y, x)], np.float32)
    assert np.array_equal(mine2.view(np.uint32), ref2.view(np.uint32))
