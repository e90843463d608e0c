"""Pins the oracle's ndtri (the inverse normal distribution function that
PDL's grandom applies, Primitive.pd:2256-2261 -> Math.pd:346-356 ->
lib/PDL/Math/ndtri.c) and the library's jump-ahead of PDL's xoshiro256+
stream, all on the CPU.

  * ndtri: bit-exact against the reference's own ndtri.c (with its polevl.c,
    mtherr.c, const.c) compiled into oracle/_ref where the reference sources
    are present, and, independently of it, within 1e-15 relative of the
    exact inverse (one Newton step on mpmath's 120-bit normal CDF from the
    returned value) across the middle and both tails -- a wrong coefficient,
    branch or sign misses by orders of magnitude -- plus its closed forms
    (ndtri(1/2) = 0, odd symmetry, the domain ends at +-MAXNUM);
  * the stream: the library's state after n draws (a GF(2) matrix power of
    xoshiro256's linear state transition) equals the state the oracle's
    sequential generator reaches, across chunk boundaries and far ahead.
"""
import ctypes
import math
import os

import numpy as np
import pytest

from oracle import pdlcpu as O

REF = os.path.join(os.path.dirname(O.HERE), "oracle", "_ref", "libndtri_ref.so")
MAXNUM = 1.79769313486231570815E308


def sample_y(rng, n):
    """the middle, both tails down to 1e-300, the branch points"""
    mid = rng.uniform(0.0, 1.0, n)
    lo = np.exp(-rng.uniform(0, 690, n))
    hi = 1.0 - np.exp(-rng.uniform(0, 36, n))
    e2 = math.exp(-2.0)
    edges = np.array([0.5, e2, np.nextafter(e2, 1), np.nextafter(e2, 0), 1 - e2, np.nextafter(1 - e2, 1),
                      math.exp(-32.0), np.nextafter(math.exp(-32.0), 0), 2.0 ** -53, 1 - 2.0 ** -53, 5e-324])
    y = np.concatenate([mid, lo, hi, edges])
    return y[(y > 0) & (y < 1)]


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (needs the reference sources)")
def test_ndtri_matches_the_reference_source_bit_for_bit():
    ref = ctypes.CDLL(REF)
    ref.ndtri.argtypes = [ctypes.c_double]
    ref.ndtri.restype = ctypes.c_double
    y = sample_y(np.random.default_rng(1), 30000)
    got = O.ndtri(y)
    want = np.array([ref.ndtri(float(v)) for v in y])
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))


def test_ndtri_against_the_exact_inverse():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.prec = 120
    y = sample_y(np.random.default_rng(2), 1500)
    x = O.ndtri(y)
    worst = 0.0
    for yi, xi in zip(y, x):
        X, Y = mpmath.mpf(float(xi)), mpmath.mpf(float(yi))
        # one Newton step from the returned value onto the exact inverse
        dx = (mpmath.ncdf(X) - Y) / mpmath.npdf(X)
        truth = X - dx
        if truth == 0:
            assert xi == 0.0
            continue
        rel = float(abs(dx / truth))
        worst = max(worst, rel)
        assert rel <= 1e-15, (float(yi), float(xi), rel)
    assert worst > 0.0


def test_ndtri_closed_forms():
    assert O.ndtri(np.array([0.5]))[0] == 0.0
    # odd symmetry where 1 - y is exact
    y = np.random.default_rng(3).uniform(0.5, 1.0, 2000)
    assert np.all((1.0 - (1.0 - y)) == y)
    np.testing.assert_array_equal(O.ndtri(1.0 - y), -O.ndtri(y))
    # the domain ends (Cephes returns -+MAXNUM, PDL passes it on)
    assert O.ndtri(np.array([0.0]))[0] == -MAXNUM
    assert O.ndtri(np.array([1.0]))[0] == MAXNUM
    # textbook quantiles
    np.testing.assert_allclose(O.ndtri(np.array([0.975, 0.8413447460685429, 0.001])),
                               [1.959963984540054, 1.0, -3.090232306167813], rtol=1e-14)


def test_grandom_float_rounds_the_draw_first():
    """grandom(float) = float(ndtri(double(float(draw)))): a draw within
    2^-25 of 1 rounds to 1.0f, and ndtri(1) = MAXNUM becomes +Inf, as in PDL"""
    u = O.random_from(200000, 7)
    g = O.grandom(np.float32, 200000, 7)
    want = O.ndtri(u.astype(np.float32).astype(np.float64)).astype(np.float32)
    np.testing.assert_array_equal(g.view(np.uint32), want.view(np.uint32))
    assert np.all(np.isfinite(g[u.astype(np.float32) < 1.0]))


# ------------------------------------------------------------ the stream's jump-ahead
def _lib_or_skip():
    from pdl_amd import _lib as L
    try:
        return L.lib()
    except Exception as e:                      # noqa: BLE001
        pytest.skip("libpdlhip.so not built: %s" % e)


@pytest.mark.parametrize("seed", [0, 5, 2**63 + 11])
def test_stream_jump_ahead_matches_sequential_draws(seed):
    lib = _lib_or_skip()
    from pdl_amd import _lib as L
    total = 0
    for skip in (0, 1, 4095, 1, 4096, 4097, 123457, 2**20 + 3):
        L.check(lib.pdlhip_srandom(seed))
        L.check(lib.pdlhip_rand_skip(skip))
        st = (ctypes.c_uint64 * 4)()
        L.check(lib.pdlhip_rand_state(st))
        assert [int(v) for v in st] == [int(v) for v in O.rand_state_after(seed, skip)], skip
        total += skip
    # skips add up: one call of a+b draws ends where two calls end
    L.check(lib.pdlhip_srandom(seed))
    L.check(lib.pdlhip_rand_skip(3000))
    L.check(lib.pdlhip_rand_skip(9000))
    st = (ctypes.c_uint64 * 4)()
    L.check(lib.pdlhip_rand_state(st))
    assert [int(v) for v in st] == [int(v) for v in O.rand_state_after(seed, 12000)]


def test_stream_far_ahead_is_consistent():
    """2^40 draws ahead: the matrix power equals 2^20 draws taken 2^20 times"""
    lib = _lib_or_skip()
    from pdl_amd import _lib as L
    L.check(lib.pdlhip_srandom(9))
    L.check(lib.pdlhip_rand_skip(2 ** 40))
    a = (ctypes.c_uint64 * 4)()
    L.check(lib.pdlhip_rand_state(a))
    L.check(lib.pdlhip_srandom(9))
    for _ in range(2 ** 10):
        L.check(lib.pdlhip_rand_skip(2 ** 30))
    b = (ctypes.c_uint64 * 4)()
    L.check(lib.pdlhip_rand_state(b))
    assert list(a) == list(b)
