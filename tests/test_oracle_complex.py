"""Pins the oracle's float/double complex ops (C99 complex arithmetic as
GCC compiles PDL's PP code) on the CPU: the reference's known answer for
complex matmult and its mismatch text (t/primitive-matmult.t:29-38), the
product's definition (ac - bd) + i(ad + bc) evaluated with separate numpy
real ops (each step rounded once, no FMA: a contracted product fails it) and
numpy's complex sum for finite values, infinities the C99 Annex G recovery
gives (worked by hand from G.5.1), and the conversions between complex and
real types (the real part survives, through intmax_t for an unsigned
target).  Division: libgcc's Smith's method (GCC 11 calls __divsc3 /
__divdc3 for `/`) restated with separate numpy real ops on values of wide
range, and the Annex G recovery cases worked by hand."""
import numpy as np
import pytest

from oracle import pdlcpu as O


def run(op, *arrs, types=None):
    types = types or [None] * len(arrs)
    ps = [O.OPdl.from_numpy(a, t) for a, t in zip(arrs, types)]
    out = O.OPdl.null()
    O.run(op, *ps, out)
    return out


def test_complex_matmult_known_answer():
    """t/primitive-matmult.t:32-34: pdl('1 1+i 1') x pdl('2 3 i')->transpose
    = [[5+4i]] (the PMCode dummies a 1-d left operand to (3,1))"""
    a = O.OPdl.from_numpy(np.array([1, 1 + 1j, 1]), O.CD).dummy(1)      # dims (3,1)
    b = O.OPdl.from_numpy(np.array([[2], [3], [1j]]), O.CD)             # dims (1,3)
    c = O.OPdl.null()
    O.run("matmult", a, b, c)
    assert c.dims == [1, 1]
    assert c.gather().tolist() == [[5 + 4j]]


def test_complex_matmult_dim_mismatch_text():
    """t/primitive-matmult.t:35-37: $cm1->transpose x $cm2 barfs in the
    PMCode (the signature alone would broadcast the size-1 t); the package's
    matmult gives the same text before any device work (plan-only here)"""
    from pdl_amd import core as C
    C.set_plan_only(True)
    try:
        a = C.shell(O.CD, [1, 3])
        b = C.shell(O.CD, [1, 3])
        with pytest.raises(C.PDLError) as e:
            C.matmult(a, b)
        assert str(e.value) == "Dim mismatch in matmult of [1x3] x [1x3]: 1 != 3"
    finally:
        C.set_plan_only(False)


@pytest.mark.parametrize("t,npt", [(O.CF, np.complex64), (O.CD, np.complex128)], ids=["cfloat", "cdouble"])
def test_elementwise_equals_numpy_on_finite_values(t, npt):
    rng = np.random.default_rng(3)
    n = 20000
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(npt)
    y = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(npt)
    y[:50] = x[:50]                                                      # some equal pairs
    # the product from its definition, each real step a separate numpy op
    # (numpy's own complex multiply may fuse a*c - b*d into an FMA)
    a, b, c, d = x.real, x.imag, y.real, y.imag
    prod = np.empty_like(x)
    prod.real = a * c - b * d
    prod.imag = a * d + b * c
    for op, want in (("plus", x + y), ("minus", x - y), ("mult", prod),
                     ("eq", (x == y).astype(npt)), ("ne", (x != y).astype(npt))):
        got = run(op, x, y, types=[t, t]).gather()
        assert got.dtype == npt, op
        np.testing.assert_array_equal(got.view(np.uint8), want.view(np.uint8), err_msg=op)


def test_matmult_is_the_sequential_complex_sum():
    rng = np.random.default_rng(4)
    a = (rng.standard_normal((5, 7)) + 1j * rng.standard_normal((5, 7))).astype(np.complex64)   # a(t=7,h=5)
    b = (rng.standard_normal((7, 3)) + 1j * rng.standard_normal((7, 3))).astype(np.complex64)   # b(w=3,t=7)
    got = run("matmult", a, b, types=[O.CF, O.CF]).gather()
    want = np.zeros((5, 3), np.complex64)
    for h in range(5):
        for w in range(3):
            cc = np.complex64(0)
            for t in range(7):
                x, y = a[h, t], b[t, w]
                re = np.float32(x.real * y.real) - np.float32(x.imag * y.imag)
                im = np.float32(x.real * y.imag) + np.float32(x.imag * y.real)
                cc = np.complex64(complex(np.float32(cc.real + re), np.float32(cc.imag + im)))
            want[h, w] = cc
    np.testing.assert_array_equal(got.view(np.uint8), want.view(np.uint8))


def test_annex_g_infinities():
    """products whose plain formula gives NaN + iNaN: (inf + i inf)(nan + i)
    boxes the infinite factor to (1 + i) and the NaN of the other to 0:
    inf * ((1*0 - 1*1) + i(1*1 + 1*0)) = -inf + i inf; (inf + 0i)(nan + 0i)
    stays NaN + iNaN (inf * 0); finite products that overflow: (1e300 +
    1e300 i)^2 = NaN + i inf (only one part NaN, no recovery)"""
    inf, nan = np.inf, np.nan
    x = np.array([complex(inf, inf), complex(inf, 0), complex(1e300, 1e300)])
    y = np.array([complex(nan, 1), complex(nan, 0), complex(1e300, 1e300)])
    got = run("mult", x, y, types=[O.CD, O.CD]).gather()
    assert got[0].real == -inf and got[0].imag == inf
    assert np.isnan(got[1].real) and np.isnan(got[1].imag)
    assert np.isnan(got[2].real) and got[2].imag == inf


def test_conversions_with_a_complex_end():
    z = np.array([1.75 - 2j, -3.5 + 0.25j, 1e10 + 1j])
    # complex to real: the real part (long truncates; ulong through intmax_t)
    c = run("assgn", z, types=[O.CD])
    np.testing.assert_array_equal(c.gather(), z)
    out = O.OPdl.from_numpy(np.zeros(3, np.int64), O.LL)
    O.run("assgn", O.OPdl.from_numpy(z, O.CD), out)
    np.testing.assert_array_equal(out.gather(), [1, -3, 10000000000])
    out = O.OPdl.from_numpy(np.zeros(3, np.float32), O.F)
    O.run("assgn", O.OPdl.from_numpy(z, O.CD), out)
    np.testing.assert_array_equal(out.gather(), z.real.astype(np.float32))
    # real to complex: (x, 0); a long plus a cfloat computes in cfloat
    s = run("plus", np.array([1, 2, 3], np.int32), np.array([0.5j, 1, 2 + 2j], np.complex64), types=[O.L, O.CF])
    assert s.dtype == O.CF
    np.testing.assert_array_equal(s.gather(), np.array([1 + 0.5j, 3, 5 + 2j], np.complex64))


@pytest.mark.parametrize("t,npt", [(O.CF, np.complex64), (O.CD, np.complex128)], ids=["cfloat", "cdouble"])
def test_complex_sums_and_products(t, npt):
    """Ufunc.pd:88-140 for a complex type (int+ is the type itself): the sum
    and the cumulative sum are numpy's sequential cumsum; the product is the
    sequential product of the definition; prodover stops at a zero running
    product, so a NaN after a zero does not reach the result"""
    rng = np.random.default_rng(5)
    x = (rng.standard_normal((6, 33)) + 1j * rng.standard_normal((6, 33))).astype(npt)
    cs = run("cumusumover", x, types=[t]).gather()
    np.testing.assert_array_equal(cs.view(np.uint8), np.cumsum(x, axis=1).view(np.uint8))
    s = run("sumover", x, types=[t]).gather()
    np.testing.assert_array_equal(s.view(np.uint8), np.cumsum(x, axis=1)[:, -1].copy().view(np.uint8))
    rt = x.real.dtype.type

    def cmul(p, q):
        return complex(rt(rt(p.real * q.real) - rt(p.imag * q.imag)), rt(rt(p.real * q.imag) + rt(p.imag * q.real)))

    y = x * npt(0.9)
    want = np.empty_like(y)
    for r in range(y.shape[0]):
        acc = npt(1)
        for k in range(y.shape[1]):
            acc = npt(cmul(acc, y[r, k]))
            want[r, k] = acc
    cp = run("cumuprodover", y, types=[t]).gather()
    np.testing.assert_array_equal(cp.view(np.uint8), want.view(np.uint8))
    p = run("prodover", y, types=[t]).gather()
    np.testing.assert_array_equal(p.view(np.uint8), want[:, -1].copy().view(np.uint8))
    z = np.array([[2 + 1j, 0, complex(np.nan, 0), 3]], npt)
    assert run("prodover", z, types=[t]).gather().tolist() == [0j]
    assert np.isnan(run("cumuprodover", z, types=[t]).gather()[0, 2].real)


def smith(x, y):
    """Smith's method, each real step a separate numpy op in the type: the
    ratio of y's smaller part to its larger"""
    a, b, c, d = x.real, x.imag, y.real, y.imag
    with np.errstate(all="ignore"):
        first = np.abs(c) < np.abs(d)
        r1 = c / d
        den1 = c * r1 + d
        re1, im1 = (a * r1 + b) / den1, (b * r1 - a) / den1
        r2 = d / c
        den2 = d * r2 + c
        re2, im2 = (b * r2 + a) / den2, (b - a * r2) / den2
    out = np.empty_like(x)
    out.real = np.where(first, re1, re2)
    out.imag = np.where(first, im1, im2)
    return out


@pytest.mark.parametrize("t,npt", [(O.CF, np.complex64), (O.CD, np.complex128)], ids=["cfloat", "cdouble"])
def test_divide_is_smiths_method(t, npt):
    rng = np.random.default_rng(6)
    n = 20000
    rt = np.float32 if t == O.CF else np.float64
    span = 60 if t == O.CF else 400
    def wide():
        return (rng.standard_normal(n) * 2.0 ** rng.integers(-span, span, n)).astype(rt)
    x = (wide() + 1j * wide()).astype(npt)
    y = (wide() + 1j * wide()).astype(npt)
    y[:40] = y[:40].real.astype(npt)                                     # real divisors
    y[40:80] = (1j * y[40:80].imag).astype(npt)                          # imaginary divisors
    got = run("divide", x, y, types=[t, t]).gather()
    want = smith(x, y)
    ok = ~(np.isnan(want.real) & np.isnan(want.imag))
    np.testing.assert_array_equal(got[ok].view(rt), want[ok].view(rt))


def test_divide_annex_g_cases():
    """(1+i)/(±0+0i) = ±inf(1+i); (inf+i inf)/2 boxes x to (1+i): inf(2) + i
    inf(2); (1+i)/(inf+i inf) boxes y: 0(1+1) + i 0(1-1) = 0 + 0i; 0/0 stays
    NaN + iNaN; (inf+i inf)/(1+i) has one NaN part only (no recovery)"""
    inf, nan = np.inf, np.nan
    x = np.array([1 + 1j, 1 + 1j, complex(inf, inf), 1 + 1j, 0j, complex(inf, inf), 3 - 4j])
    y = np.array([0j, complex(-0.0, 0.0), 2 + 0j, complex(inf, inf), 0j, 1 + 1j, 1 + 2j])
    for t in (O.CF, O.CD):
        got = run("divide", x, y, types=[t, t]).gather()
        assert got[0] == complex(inf, inf) and got[1] == complex(-inf, -inf)
        assert got[2] == complex(inf, inf)
        assert got[3] == 0j and not np.signbit(got[3].real) and not np.signbit(got[3].imag)
        assert np.isnan(got[4].real) and np.isnan(got[4].imag)
        assert got[5].real == inf and np.isnan(got[5].imag)
        assert got[6] == -1 - 2j                                         # (3-4i)/(1+2i), exact
