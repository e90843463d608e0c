"""Pins the oracle's libm-backed binops, `power` and `atan2`
(lib/PDL/Ops.pd:321-322: `bifunc('power',['pow','op**'],...)` and
`bifunc('atan2','atan2',...)`, both through <tgmath.h>, Ops.pd:66), which
the reference's own tests fix at three points only (t/ops.t:33, 126-127).

Each result is held to the mathematics, not to the oracle's own formula:
  * random operands in every sign/quadrant, scored in units in the last
    place of a 120-bit mpmath value (glibc documents < 1 ulp for
    powf/pow/atan2f/atan2 on x86-64); a swapped-argument or sign mistake
    misses by many ulps on these inputs;
  * the special cases C99 Annex F fixes exactly (F.9.4.4 pow, F.9.1.4 atan2):
    zeros of both signs, infinities, NaN, negative bases with integer and
    non-integer exponents.
"""
import math

import numpy as np
import pytest

mpmath = pytest.importorskip("mpmath")

from oracle import pdlcpu as O

MP = mpmath.mp
MP.prec = 120

RNG = np.random.default_rng(20261203)
FMT = {O.F: (np.float32, 24, -126, -149), O.D: (np.float64, 53, -1022, -1074)}


def run_op(op, a, b, t):
    pa, pb, pc = O.OPdl.from_numpy(a, t), O.OPdl.from_numpy(b, t), O.OPdl.null()
    O.run(op, pa, pb, pc)
    assert pc.dtype == t
    return pc.numpy()


def ulp_error(got, truth, t):
    """|got - truth| in units of the type's ulp at truth (exact mpmath arithmetic)"""
    _, p, emin, qmin = FMT[t]
    if truth == 0:
        return 0.0 if got == 0 else math.inf
    e = int(mpmath.floor(mpmath.log(abs(truth), 2)))
    q = max(e - (p - 1), qmin) if e >= emin else qmin
    return float(abs(mpmath.mpf(float(got)) - truth) / mpmath.ldexp(1, q))


def finite_max(t):
    return float(np.finfo(FMT[t][0]).max)


def pow_truth(x, y):
    x, y = mpmath.mpf(float(x)), mpmath.mpf(float(y))
    if x < 0:
        if y != mpmath.floor(y):
            return None                      # NaN in C: negative base, non-integer exponent
        r = (-x) ** y
        return -r if int(y) % 2 else r
    return x ** y


@pytest.mark.parametrize("t", [O.F, O.D], ids=["float", "double"])
def test_power_within_one_ulp_of_the_exact_value(t):
    dt = FMT[t][0]
    n = 1500
    # positive bases over many binades, real exponents; negative bases with
    # integer exponents (odd and even, positive and negative)
    xp = np.exp(RNG.uniform(-20, 20, n)).astype(dt)
    yp = RNG.uniform(-6, 6, n).astype(dt)
    xn = -np.exp(RNG.uniform(-6, 6, n)).astype(dt)
    yn = RNG.integers(-9, 10, n).astype(dt)
    xs = np.concatenate([xp, xn, xp[:200]])
    ys = np.concatenate([yp, yn, np.full(200, 2.0, dt)])     # squares too
    got = run_op("power", xs, ys, t)
    worst = 0.0
    for x, y, g in zip(xs, ys, got):
        tr = pow_truth(x, y)
        if tr is None:
            assert np.isnan(g), (x, y, g)
            continue
        if abs(tr) > finite_max(t):
            assert np.isinf(g) and (g > 0) == (tr > 0), (x, y, g)
            continue
        err = ulp_error(g, tr, t)
        worst = max(worst, err)
        assert err <= 1.0, (float(x), float(y), float(g), str(tr), err)
    assert worst > 0.0           # not all exact: the comparison has teeth


@pytest.mark.parametrize("t", [O.F, O.D], ids=["float", "double"])
def test_atan2_within_one_ulp_in_every_quadrant(t):
    dt = FMT[t][0]
    n = 1000
    mag_y = np.exp(RNG.uniform(-10, 10, n))
    mag_x = np.exp(RNG.uniform(-10, 10, n))
    sy = RNG.choice([-1.0, 1.0], n)
    sx = RNG.choice([-1.0, 1.0], n)
    ys = (sy * mag_y).astype(dt)
    xs = (sx * mag_x).astype(dt)
    got = run_op("atan2", ys, xs, t)              # atan2(a, b) = atan2(y, x)
    for y, x, g in zip(ys, xs, got):
        tr = mpmath.atan2(mpmath.mpf(float(y)), mpmath.mpf(float(x)))
        assert ulp_error(g, tr, t) <= 1.0, (float(y), float(x), float(g), str(tr))
        # the quadrant: sign follows y, |angle| > pi/2 exactly when x < 0
        assert (g > 0) == (y > 0)
        if abs(abs(float(g)) - math.pi / 2) > 1e-3:
            assert (abs(g) > math.pi / 2) == (x < 0)


def _bits_equal(a, b):
    if np.isnan(a) and np.isnan(b):
        return True
    return a == b and np.signbit(a) == np.signbit(b)


INF, NAN = math.inf, math.nan
# C99 Annex F.9.4.4 (pow)
POW_SPECIAL = [
    (+0.0, -3.0, INF), (-0.0, -3.0, -INF), (+0.0, -2.0, INF), (-0.0, -2.0, INF),
    (+0.0, -0.5, INF), (-0.0, -INF, INF), (+0.0, 3.0, 0.0), (-0.0, 3.0, -0.0),
    (-0.0, 2.0, 0.0), (+0.0, 0.5, 0.0), (-1.0, INF, 1.0), (-1.0, -INF, 1.0),
    (1.0, NAN, 1.0), (1.0, -INF, 1.0), (NAN, 0.0, 1.0), (NAN, -0.0, 1.0), (5.0, 0.0, 1.0),
    (-2.0, 0.5, NAN), (0.5, -INF, INF), (2.0, -INF, 0.0), (0.5, INF, 0.0), (2.0, INF, INF),
    (-INF, -3.0, -0.0), (-INF, -2.0, 0.0), (-INF, 3.0, -INF), (-INF, 2.0, INF),
    (INF, -1.0, 0.0), (INF, 0.5, INF), (NAN, 2.0, NAN), (2.0, NAN, NAN),
    (-8.0, 1.0 / 3.0, NAN), (-2.0, 3.0, -8.0), (-2.0, -2.0, 0.25),
]
# C99 Annex F.9.1.4 (atan2(y, x))
ATAN2_SPECIAL = [
    (+0.0, -0.0, math.pi), (-0.0, -0.0, -math.pi), (+0.0, +0.0, +0.0), (-0.0, +0.0, -0.0),
    (+0.0, -1.0, math.pi), (-0.0, -1.0, -math.pi), (+0.0, 1.0, +0.0), (-0.0, 1.0, -0.0),
    (-1.0, 0.0, -math.pi / 2), (1.0, -0.0, math.pi / 2), (1.0, -INF, math.pi), (-1.0, -INF, -math.pi),
    (1.0, INF, 0.0), (-1.0, INF, -0.0), (INF, 1.0, math.pi / 2), (-INF, 1.0, -math.pi / 2),
    (INF, -INF, 3 * math.pi / 4), (-INF, -INF, -3 * math.pi / 4), (INF, INF, math.pi / 4),
    (-INF, INF, -math.pi / 4), (NAN, 1.0, NAN), (1.0, NAN, NAN),
]


@pytest.mark.parametrize("t", [O.F, O.D], ids=["float", "double"])
def test_power_annex_f_special_cases(t):
    dt = FMT[t][0]
    xs = np.array([c[0] for c in POW_SPECIAL], dt)
    ys = np.array([c[1] for c in POW_SPECIAL], dt)
    got = run_op("power", xs, ys, t)
    for (x, y, want), g in zip(POW_SPECIAL, got):
        assert _bits_equal(g, dt(want)), ("pow", x, y, g, want)


@pytest.mark.parametrize("t", [O.F, O.D], ids=["float", "double"])
def test_atan2_annex_f_special_cases(t):
    dt = FMT[t][0]
    ys = np.array([c[0] for c in ATAN2_SPECIAL], dt)
    xs = np.array([c[1] for c in ATAN2_SPECIAL], dt)
    got = run_op("atan2", ys, xs, t)
    for (y, x, want), g in zip(ATAN2_SPECIAL, got):
        assert _bits_equal(g, dt(want)), ("atan2", y, x, g, want)


def test_power_swap_reverses_operands():
    """`$b ** $a` with swap (Ops.pd bifunc's swap OtherPar) = pow(b, a)"""
    a = np.array([2.0, 3.0, 0.5], np.float64)
    b = np.array([5.0, 2.0, 4.0], np.float64)
    pa, pb, pc = O.OPdl.from_numpy(a, O.D), O.OPdl.from_numpy(b, O.D), O.OPdl.null()
    O.run("power", pa, pb, pc, swap=1)
    np.testing.assert_array_equal(pc.numpy(), [25.0, 8.0, 2.0])


def test_square_via_pow_differs_from_the_product_rarely():
    """glibc's powf(x, 2) is not always the correctly rounded x*x: measure
    how often (this is the tolerance the device's `** 2 -> x*x` shortcut is
    held to, and the reason that shortcut is within 1 ulp, not bit-exact)"""
    x = RNG.standard_normal(200000).astype(np.float32)
    got = run_op("power", x, np.full_like(x, 2.0), O.F)
    sq = x * x
    diff = got != sq
    rate = float(diff.mean())
    assert rate < 2e-3
    # wherever they differ, the two are adjacent floats
    gi, si = got[diff].view(np.int32), sq[diff].view(np.int32)
    assert np.all(np.abs(gi.astype(np.int64) - si) == 1)
