"""Accuracy of the engine's FP64 math layer (geliosphere_amd/csrc/gs_math.hpp)
against mpmath at 50 digits: host evaluation of the polynomial parts (CPU)
and device evaluation of everything, Newton-refined seeds included (GPU)."""
import math
import random

import mpmath
import pytest

from geliosphere_amd import engine

mpmath.mp.dps = 50
E = engine


def ulps(y, ref):
    ref = float(ref)
    if ref == 0.0:
        return abs(y) / 5e-324
    return abs(y - ref) / math.ulp(ref)


def samples(seed=1):
    rng = random.Random(seed)
    log_x = [rng.uniform(0.5, 2.0) for _ in range(300)] + \
            [10.0 ** rng.uniform(-300, 300) for _ in range(700)] + [1.0, 2.0, math.sqrt(2), 1e-300]
    exp_x = [rng.uniform(-700, 700) for _ in range(600)] + [rng.uniform(-1, 1) for _ in range(400)] + [0.0]
    # random words; the top of the range, where ln u1 is tiny and must stay
    # accurate relative to itself; both sides of the sin/cos table's
    # interval edges (j 2^24 +- 2^23) and of powers of two
    u32 = [float(rng.getrandbits(32)) for _ in range(1000)] + \
          [0.0, 1.0, 2.0 ** 29, 2.0 ** 29 - 1, 2.0 ** 30, 2.0 ** 31, 3 * 2.0 ** 30, 2.0 ** 32 - 1] + \
          [float(2 ** 32 - k) for k in range(2, 200)] + \
          [float((j << 24) + d) for j in range(0, 256, 5) for d in (-(1 << 23) - 1, -(1 << 23), (1 << 23) - 1)
           if 0 <= (j << 24) + d < 2 ** 32] + \
          [float((1 << k) + d) for k in range(1, 32) for d in (-2, -1, 0)]
    # sin/cos: the 2D steps' ranges (theta in [0, pi], sheet angles of tens of
    # radians), larger arguments, and both sides of the table's interval
    # edges (k + 1/2) pi/128
    trig = [rng.uniform(-10, 10) for _ in range(400)] + [rng.uniform(0, math.pi) for _ in range(400)] + \
           [rng.uniform(-1e5, 1e5) for _ in range(100)] + \
           [(k + 0.5) * math.pi / 128 * (1 + e) for k in range(0, 300, 7) for e in (-1e-15, 1e-15)] + \
           [0.0, math.pi / 4, math.pi / 2, math.pi, 2 * math.pi, 1e-10, -3.0, 1000.5]
    at = [s * 10.0 ** rng.uniform(-8, 8) for s in (1, -1) for _ in range(400)] + \
         [0.0, 1.0, -1.0, 0.19891236737965801, 0.66817863791929892, 0.4142135623730950, 1e300, -1e300]
    return log_x, exp_x, u32, trig, at


def frac_samples(seed):
    """(n, d) with 0 <= n <= d: random ratios, the reduction's interval edges
    (tan(pi/16), tan(3 pi/16)) from both sides, 0 and 1, over wide scales."""
    rng = random.Random(seed)
    pairs = []
    for _ in range(600):
        d = 10.0 ** rng.uniform(-10, 10)
        pairs.append((d * rng.random(), d))
    for q in (0.19891236737965801, 0.66817863791929892):
        for e in (-3e-16, 0.0, 3e-16, 1e-9, -1e-9):
            pairs.append((q * (1 + e), 1.0))
    pairs += [(0.0, 1.0), (1.0, 1.0), (3.0, 3.0), (1e-300, 1.0), (0.5, 1.0)]
    return pairs


def check_atan_frac(fn_eval, pairs):
    y = fn_eval(E.DBG_ATAN_FRAC, [p[0] for p in pairs] + [p[1] for p in pairs])
    for a, (num, den) in zip(y, pairs):
        ref = mpmath.atan(mpmath.mpf(num) / mpmath.mpf(den))
        assert ulps(a, ref) <= 2.0, (num, den, a, float(ref))


def sheet_samples(seed):
    """(sin of the sheet angle, k = 1 - 2 theta / pi) pairs: random, both
    signs, tiny and edge values."""
    rng = random.Random(seed)
    pairs = [(rng.uniform(-1, 1), rng.uniform(-1, 1)) for _ in range(400)]
    pairs += [(s * 10.0 ** rng.uniform(-12, 0), rng.uniform(-1, 1)) for s in (1, -1) for _ in range(50)]
    pairs += [(1.0, 1.0), (-1.0, -1.0), (1.0, -1.0), (0.5, 1e-9), (-0.5, -1e-9), (1e-9, 1.0), (0.3, 0.0)]
    return pairs


def check_sheet(fn_eval, pairs):
    """The drift factors in fraction form against the reference's own chain
    (SolarPropLikeCpuModel.cpp:183-209) evaluated at 50 digits."""
    xs = [p[0] for p in pairs] + [p[1] for p in pairs]
    f = fn_eval(E.DBG_SHEET_F, xs)
    g = fn_eval(E.DBG_SHEET_G, xs)
    for a, b, (sy, k) in zip(f, g, pairs):
        sy, k = mpmath.mpf(sy), mpmath.mpf(k)
        x = sy / (mpmath.pi - sy)
        ah = mpmath.acos(x)
        tanh = mpmath.tan(ah)
        arg = k * tanh
        rf = mpmath.atan(arg) / ah
        rg = tanh / (ah * (1 + arg * arg))
        assert abs(a - float(rf)) <= 8 * math.ulp(float(rf)) + 1e-300, (sy, k, a, float(rf))
        assert abs(b - float(rg)) <= 8 * math.ulp(float(rg)) + 1e-300, (sy, k, b, float(rg))


def check_trig(fn_eval, trig, at):
    for fn, f in ((E.DBG_SIN, mpmath.sin), (E.DBG_COS, mpmath.cos)):
        y = fn_eval(fn, trig)
        assert max(abs(a - float(f(x))) for a, x in zip(y, trig)) <= 2.5e-16
    y = fn_eval(E.DBG_ATAN, at)
    assert max(ulps(a, mpmath.atan(x)) for a, x in zip(y, at)) <= 2.0
    assert fn_eval(E.DBG_ATAN, [float("inf")])[0] == math.pi / 2


def check(fn_eval, log_x, exp_x, u32, trig, at):
    check_trig(fn_eval, trig, at)
    y = fn_eval(E.DBG_LOG, log_x)
    assert max(ulps(a, mpmath.log(x)) for a, x in zip(y, log_x)) <= 2.0
    y = fn_eval(E.DBG_EXP, exp_x)
    assert max(ulps(a, mpmath.exp(x)) for a, x in zip(y, exp_x)) <= 2.0
    for fn, f in ((E.DBG_SIN_TURN, mpmath.sin), (E.DBG_COS_TURN, mpmath.cos)):
        y = fn_eval(fn, u32)
        err = max(abs(a - float(f(2 * mpmath.pi * mpmath.mpf(int(k)) / 2 ** 32))) for a, k in zip(y, u32))
        assert err <= 2.5e-16
    y = fn_eval(E.DBG_LOG_U32, u32)
    for a, k in zip(y, u32):
        ref = mpmath.log((mpmath.mpf(int(k)) + 1) / 2 ** 32)
        assert abs(a - float(ref)) <= 2 * math.ulp(float(ref)) + 1e-300
    assert y[u32.index(2.0 ** 32 - 1)] == 0.0    # u1 = 1 gives exactly ln 1 = 0


def test_host_math_accuracy():
    check(E.host_math, *samples(1))
    check_atan_frac(E.host_math, frac_samples(1))
    check_sheet(E.host_math, sheet_samples(1))


@pytest.mark.gpu
def test_device_math_accuracy(engine):
    check(engine.debug_math, *samples(2))
    check_atan_frac(engine.debug_math, frac_samples(2))
    check_sheet(engine.debug_math, sheet_samples(2))
    rng = random.Random(3)
    xs = [10.0 ** rng.uniform(-200, 200) for _ in range(1000)] + [0.3, 1.0, 99.9999]
    for fn, f in ((E.DBG_RCP, lambda v: 1 / mpmath.mpf(v)), (E.DBG_RSQRT, lambda v: 1 / mpmath.sqrt(v)),
                  (E.DBG_SQRT, mpmath.sqrt)):
        y = engine.debug_math(fn, xs)
        assert max(ulps(a, f(x)) for a, x in zip(y, xs)) <= 1.0
    assert engine.debug_math(E.DBG_SQRT, [0.0]) == [0.0]
    # x^(-1/4) (the 1D steps' E^(-1/2) from E^2): two seeds and one correction
    q = engine.debug_math(E.DBG_RSQRT4, xs)
    assert max(ulps(a, mpmath.mpf(x) ** mpmath.mpf(-0.25)) for a, x in zip(q, xs)) <= 2.0


@pytest.mark.gpu
def test_device_normals_match_oracle(engine, oracle):
    for seed, pid in ((1, 0), (0xDEADBEEF12345678, 987654321), (7, 2 ** 40 + 3)):
        z = engine.debug_normals(seed, pid, 64)
        for n, v in enumerate(z):
            assert v == pytest.approx(oracle.normal(seed, pid, n), abs=2e-15)
