"""PDL's random numbers on the device (Primitive.pd:2091-2262) against the
oracle's sequential xoshiro256+ stream (itself pinned against the
reference's Primitive-xoshiro256plus.c), through the C ABI.

Bars: random and randsym bit-exact (every element is one draw converted to
the type); grandom (ndtri of the randsym draw rounded to the type) within 1
ulp for float and 8 ulps for double of the oracle's, whose ndtri is the
reference's ndtri.c bit for bit, with over 90% of elements bit-exact: in
the tails (y < exp(-2) or 1 - y < exp(-2)) ndtri takes log twice and
sqrt, and where the device's log rounds differently from glibc's in the
last place, the subtraction x - log(x)/x - P/Q carries that ulp into a few
ulps of a double result (8 ulps = 1.8e-15 relative, about Cephes' own
stated accuracy, 7.2e-16 peak)."""
import numpy as np
import pytest

from pdl_amd import core as C
from pdl_amd import _lib as L
from oracle import pdlcpu as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    C.init(0)
    yield


def ulps(x, y):
    it = np.int32 if x.dtype == np.float32 else np.int64
    xi, yi = x.view(it).astype(np.int64), y.view(it).astype(np.int64)
    lo = np.int64(np.iinfo(it).min)
    xi = np.where(xi < 0, lo - xi, xi)
    yi = np.where(yi < 0, lo - yi, yi)
    d = np.abs(xi.astype(object) - yi.astype(object)).astype(np.float64)
    d[np.isnan(x) & np.isnan(y)] = 0
    d[np.isnan(x) ^ np.isnan(y)] = np.inf
    return d


def test_srandom_repeats_the_stream():
    """t/primitive-random.t:9-22: the same seed gives the same values"""
    C.srandom(5)
    r1 = C.random(10).numpy()
    C.srandom(5)
    r2 = C.random(10).numpy()
    np.testing.assert_array_equal(r1, r2)
    np.testing.assert_array_equal(r1, O.random_from(10, 5))
    C.srandom(10)
    g1 = C.grandom(10).numpy()
    C.srandom(10)
    g2 = C.grandom(10).numpy()
    np.testing.assert_array_equal(g1, g2)


def test_types_and_empty():
    """t/primitive-random.t:26-43: double by default; an empty ndarray draws nothing"""
    C.srandom(1)
    assert C.random().dtype == L.D
    assert C.randsym().dtype == L.D
    before = C.rand_state()
    x = C.random(1, 1, 0)
    assert x.dims == [1, 1, 0]
    assert C.rand_state() == before


@pytest.mark.parametrize("n", [1, 15, 16, 17, 4095, 4096, 4097, 3 * 4096 + 17, 1000003])
@pytest.mark.parametrize("t", [L.D, L.F], ids=["double", "float"])
def test_random_bit_exact_across_chunks(n, t):
    C.srandom(42)
    got = C.random(n, dtype=t).numpy()
    want = O.random_from(n, 42).astype(np.float64 if t == L.D else np.float32)
    np.testing.assert_array_equal(got.view(np.uint8), want.view(np.uint8))
    # the stream moved on by n draws
    assert C.rand_state() == [int(v) for v in O.rand_state_after(42, n)]


def test_stream_continues_across_calls_and_types():
    C.srandom(3)
    a = C.random(7, 1000, 3).numpy()             # PDL dims (7,1000,3): numpy shape (3,1000,7)
    b = C.random(5000, dtype=L.F).numpy()
    c = C.randsym(33, dtype=L.D).numpy()
    u = O.random_from(21000 + 5000 + 33, 3)
    np.testing.assert_array_equal(a.reshape(-1), u[:21000])
    np.testing.assert_array_equal(b, u[21000:26000].astype(np.float32))
    np.testing.assert_array_equal(c, u[26000:])


@pytest.mark.parametrize("t", [L.B, L.S, L.L, L.LL, L.UL], ids=["byte", "short", "long", "longlong", "ulong"])
def test_random_integer_types_truncate(t):
    """random in an integer type converts [0,1) draws: all zero, one draw each"""
    C.srandom(8)
    x = C.random(3000, dtype=t)
    assert not np.any(x.numpy())
    assert C.rand_state() == [int(v) for v in O.rand_state_after(8, 3000)]


def test_random_into_views_in_pdl_order():
    """a view fills in its own element order, through its dims and incs;
    the parent's other elements keep their values"""
    base = C.ndarray.from_numpy(np.full((40, 30), -1.0), L.D)   # PDL dims (30,40)
    v = base.xchg(0, 1).slice((2, 31, 3), (1, 25, 2))           # dims (10,13), strided, transposed
    C.srandom(11)
    C.random_into(v)
    u = O.random_from(10 * 13, 11).reshape(13, 10)               # u[j, i] = draw i + 10 j
    want = np.full((40, 30), -1.0)
    for j in range(13):
        for i in range(10):
            # v(i, j) = base(1 + 2 j, 2 + 3 i) in PDL dims -> numpy [2 + 3 i, 1 + 2 j]
            want[2 + 3 * i, 1 + 2 * j] = u[j, i]
    np.testing.assert_array_equal(base.numpy(), want)


def test_large_grid_stride_sampled():
    """2^26 + 5 floats (16385 chunks, more than one per block): sampled
    positions equal the sequential stream"""
    n = 2 ** 26 + 5
    C.srandom(2026)
    x = C.random(n, dtype=L.F).numpy()
    for pos in (0, 4095, 4096, 4096 * 2047, 4096 * 2048 + 1, 4096 * 5000 + 77, n - 4100):
        want = O.random_from(4100, 2026, pos).astype(np.float32)
        np.testing.assert_array_equal(x[pos:pos + 4100], want, err_msg=str(pos))
    assert C.rand_state() == [int(v) for v in O.rand_state_after(2026, n)]


@pytest.mark.parametrize("t,bound", [(L.F, 1), (L.D, 8)], ids=["float", "double"])
def test_grandom_within_ulps_of_the_oracle(t, bound):
    n = 300000
    C.srandom(77)
    got = C.grandom(n, dtype=t).numpy()
    npt = np.float32 if t == L.F else np.float64
    want = O.grandom(npt, n, 77)
    d = ulps(got, want)
    assert d.max() <= bound, (d.max(), np.mean(d == 0))
    assert np.mean(d == 0) > 0.9, np.mean(d == 0)
    assert C.rand_state() == [int(v) for v in O.rand_state_after(77, n)]


def test_grandom_refuses_integer_types():
    C.srandom(1)
    with pytest.raises(L.PDLError) as e:
        C.grandom(10, dtype=L.L)
    assert e.value.kind == L.EUSERERROR
    assert "grandom" in str(e.value)
