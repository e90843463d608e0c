"""Parity of the Ops.pd ufuncs (Ops.pd:222-265, 327-381) and the Bad.pd
elementwise ops (Bad.pd:343-395, 584-840) on the device with the oracle,
through the C ABI.  Bars: bit-exact (values, types and bad flags) for
not/bitnot/abs, sqrt (correctly rounded on both sides), every integer-type
result and every Bad.pd op; sin/cos/exp/log/log10 on float and double within
a stated number of ulps of glibc (the oracle, itself held within glibc's
documented bound of the exact value by tests/test_oracle_libm.py)."""
import numpy as np
import pytest

from pdl_amd import core as C
from oracle import pdlcpu as O

pytestmark = pytest.mark.gpu

INT_T = [O.SB, O.B, O.S, O.US, O.L, O.UL, O.IND, O.ULL, O.LL]
ALL_T = INT_T + [O.F, O.D]
NAMES = O.TYPE_NAMES


@pytest.fixture(scope="module", autouse=True)
def device():
    C.init(0)
    yield


def pair(arr, t, badflag=False, badvalue=None):
    return (O.OPdl.from_numpy(arr, t, badflag=badflag, badvalue=badvalue),
            C.ndarray.from_numpy(arr, t, badflag=badflag, badvalue=badvalue))


def view(p, v):
    for step in v:
        p = getattr(p, step[0])(*step[1:])
    return p


def run_both(op, inputs, nout=1, dpar=None):
    oo = [O.OPdl.null() for _ in range(nout)]
    dd = [C.ndarray() for _ in range(nout)]
    O.run(op, *[i[0] for i in inputs], *oo, dpar=dpar)
    C._run(op, [i[1] for i in inputs] + dd, dpar=dpar)
    return list(zip(oo, dd))


def same(o, d, what=""):
    assert NAMES[o.dtype] == C.TYPE_NAMES[d.dtype], what
    assert list(o.dims) == list(d.dims), what
    a, b = o.gather(), d.numpy()
    if a.dtype.kind == "f":
        u = {4: np.uint32, 8: np.uint64}[a.dtype.itemsize]
        an, bn = np.isnan(a), np.isnan(b)
        np.testing.assert_array_equal(an, bn, err_msg=what)
        np.testing.assert_array_equal(np.where(an, 0, a).view(u), np.where(bn, 0, b).view(u), err_msg=what)
    else:
        np.testing.assert_array_equal(a, b, err_msg=what)
    assert bool(o.state & O.BADVAL) == d.badflag(), what


def ulp_distance(x, y):
    it = np.int32 if x.dtype == np.float32 else np.int64
    xi, yi = x.view(it).astype(np.int64), y.view(it).astype(np.int64)
    lo = np.int64(np.iinfo(it).min)
    xi = np.where(xi < 0, lo - xi, xi)
    yi = np.where(yi < 0, lo - yi, yi)
    d = np.abs(xi.astype(object) - yi.astype(object)).astype(np.float64)
    d[np.isnan(x) & np.isnan(y)] = 0
    d[np.isnan(x) ^ np.isnan(y)] = np.inf
    return d


def rand(t, shape, rng, lo=None, hi=None):
    dt = np.dtype(O.NP_TYPES[t])
    if dt.kind in "iu":
        info = np.iinfo(dt)
        lo = info.min if lo is None else lo
        hi = info.max if hi is None else hi
        return rng.integers(lo, hi, size=shape, endpoint=True, dtype=dt)
    return (rng.standard_normal(shape) * 10).astype(dt)


LAYOUTS = {"flat": ((37, 91), []), "xchg": ((130, 70), [("xchg", 0, 1)]),
           "slice": ((40, 60), [("slice", (1, 38, 3), None)]), "dummy": ((50,), [("dummy", 1, 7)])}


@pytest.mark.parametrize("layout", list(LAYOUTS))
@pytest.mark.parametrize("t", ALL_T, ids=[NAMES[t] for t in ALL_T])
def test_exact_ufuncs_bit_exact(t, layout):
    """not, abs (PDL_ABS: -0.0 stays, NaN flips sign, the most negative
    integer wraps), bitnot (integer types; a float input computes in
    longlong), sqrt of non-negative values"""
    rng = np.random.default_rng(60 + t)
    shape, v = LAYOUTS[layout]
    a = rand(t, shape, rng)
    if t >= O.F:
        a.reshape(-1)[:4] = [0.0, -0.0, np.nan, -np.inf]
    else:
        a.reshape(-1)[:2] = [0, np.iinfo(a.dtype).min]
    A = pair(a, t)
    A = (view(A[0], v), view(A[1], v))
    for op in ("not", "abs"):
        (o, d), = run_both(op, [A])
        same(o, d, "%s %s %s" % (op, NAMES[t], layout))
    # bitnot of a float input computes in longlong: keep the converted values
    # defined (NaN and infinities convert to integers by no rule of C)
    f = np.nan_to_num(a, nan=1.0, posinf=2.0, neginf=-3.0) if t >= O.F else a
    Fp = pair(f, t)
    Fp = (view(Fp[0], v), view(Fp[1], v))
    (o, d), = run_both("bitnot", [Fp])
    same(o, d, "bitnot %s %s" % (NAMES[t], layout))
    s = np.abs(a) if a.dtype.kind != "u" else a
    if a.dtype.kind == "i":
        s = np.where(s < 0, 0, s).astype(a.dtype)       # |INT_MIN| wraps negative: keep sqrt's input >= 0
    S = pair(s, t)
    S = (view(S[0], v), view(S[1], v))
    (o, d), = run_both("sqrt", [S])
    same(o, d, "sqrt %s %s" % (NAMES[t], layout))


@pytest.mark.parametrize("t", INT_T, ids=[NAMES[t] for t in INT_T])
def test_transcendentals_on_integer_types_bit_exact(t):
    """sin/cos/log10 on integer types: ($GENERIC(b))f((double)a), the
    truncated double result; log10 of exact powers of ten stays exact"""
    rng = np.random.default_rng(70 + t)
    dt = np.dtype(O.NP_TYPES[t])
    hi = min(int(np.iinfo(dt).max), 10 ** 12)
    a = rng.integers(1, hi, 600, endpoint=True).astype(dt)
    p10 = [10 ** k for k in range(0, 19) if 10 ** k <= hi]
    a[:len(p10)] = np.array(p10, dtype=object).astype(dt)
    A = pair(a, t)
    for op in ("sin", "cos", "log10"):
        (o, d), = run_both(op, [A])
        same(o, d, "%s %s" % (op, NAMES[t]))


# device libm vs glibc, in ulps, on the ranges below (measured bound + margin)
TRANS_ULP = {"sin": 2, "cos": 2, "exp": 2, "log": 2, "log10": 3}


@pytest.mark.parametrize("t", [O.F, O.D], ids=["float", "double"])
@pytest.mark.parametrize("op", ["sin", "cos", "exp", "log", "log10"])
def test_transcendentals_within_ulps_of_glibc(op, t):
    rng = np.random.default_rng(80)
    dt = np.dtype(O.NP_TYPES[t]).type
    n = 20000
    if op in ("sin", "cos"):
        a = np.concatenate([rng.uniform(-100, 100, n), rng.uniform(-1e-4, 1e-4, 500)]).astype(dt)
    elif op == "exp":
        lim = 87.0 if t == O.F else 705.0
        a = rng.uniform(-lim, lim, n).astype(dt)
    else:
        a = np.exp(rng.uniform(-80, 80, n)).astype(dt)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0], dt)
    a = np.concatenate([a, special])
    (o, d), = run_both(op, [pair(a, t)])
    x, y = o.gather(), d.numpy()
    dist = ulp_distance(x, y)
    print("%s %s: max %.0f ulp, %d of %d differ" % (op, NAMES[t], dist.max(), int(np.count_nonzero(dist)), len(a)))
    assert dist.max() <= TRANS_ULP[op], (op, float(dist.max()), a[np.argmax(dist)])
    sp = slice(len(a) - len(special), len(a))
    assert np.all(dist[sp] == 0), (op, x[sp], y[sp])


@pytest.mark.parametrize("t", ALL_T, ids=[NAMES[t] for t in ALL_T])
def test_ufuncs_with_bad_values(t):
    """bad mode: ISBAD(a) -> SETBAD(b) (Ops.pd:257-259), the flag inherited"""
    rng = np.random.default_rng(90 + t)
    a = np.abs(rand(t, (12, 40), rng)) if np.dtype(O.NP_TYPES[t]).kind != "u" else rand(t, (12, 40), rng)
    if np.dtype(O.NP_TYPES[t]).kind == "i":
        a = np.where(a < 0, 1, a).astype(a.dtype)
    bv = O.default_badval(t)
    a.reshape(-1)[::7] = bv
    A = pair(a, t, badflag=True)
    for op in ("sqrt", "not", "abs") + (("bitnot",) if t in INT_T else ()):
        (o, d), = run_both(op, [A])
        same(o, d, "%s %s bad" % (op, NAMES[t]))


@pytest.mark.parametrize("t", ALL_T, ids=[NAMES[t] for t in ALL_T])
def test_isbad_isgood(t):
    rng = np.random.default_rng(100 + t)
    a = rand(t, (9, 33), rng)
    bv = O.default_badval(t)
    a.reshape(-1)[::5] = bv
    for flag in (True, False):
        A = pair(a, t, badflag=flag)
        for op in ("isbad", "isgood"):
            (o, d), = run_both(op, [(view(A[0], [("xchg", 0, 1)]), view(A[1], [("xchg", 0, 1)]))])
            same(o, d, "%s %s flag=%s" % (op, NAMES[t], flag))


@pytest.mark.parametrize("t", ALL_T, ids=[NAMES[t] for t in ALL_T])
def test_setbadif_and_setvaltobad(t):
    """setbadif with masks of another type (converted to int), bad mask
    values in bad mode, a broadcast mask; setvaltobad of a present and an
    absent value; the output's bad flag set either way"""
    rng = np.random.default_rng(110 + t)
    a = rand(t, (20, 30), rng, 0, 50) if t in INT_T else np.round(rand(t, (20, 30), rng))
    m = rng.integers(0, 3, (20, 30)).astype(np.float64)        # a double mask: converted to int
    (o, d), = run_both("setbadif", [pair(a, t), pair(m, O.D)])
    same(o, d, "setbadif %s" % NAMES[t])
    mb = rng.integers(-2, 2, (30,)).astype(np.int32)          # broadcast over dim 1
    mb[::4] = -2 ** 31                                         # int's bad value
    a2 = a.copy()
    a2.reshape(-1)[::9] = O.default_badval(t)
    for fa, fm in ((True, False), (False, True), (False, False)):
        (o, d), = run_both("setbadif", [pair(a2, t, badflag=fa), pair(mb, O.L, badflag=fm)])
        same(o, d, "setbadif %s bad a=%s mask=%s" % (NAMES[t], fa, fm))
    for val in (float(a.reshape(-1)[3]), 12345.5):
        (o, d), = run_both("setvaltobad", [pair(a, t)], dpar=val)
        same(o, d, "setvaltobad %s %s" % (NAMES[t], val))


@pytest.mark.parametrize("t", [O.F, O.D], ids=["float", "double"])
def test_nan_inf_to_bad_and_back(t):
    """setnantobad / setinftobad / setnonfinitetobad set the output's bad
    flag only when they found such a value; setbadtonan and setbadtoval
    clear it; a NaN bad value"""
    rng = np.random.default_rng(120 + t)
    dt = np.dtype(O.NP_TYPES[t]).type
    a = rand(t, (16, 50), rng)
    clean = a.copy()
    a.reshape(-1)[::11] = np.nan
    a.reshape(-1)[5::13] = np.inf
    a.reshape(-1)[7::17] = -np.inf
    for op in ("setnantobad", "setinftobad", "setnonfinitetobad"):
        for arr in (a, clean):
            (o, d), = run_both(op, [pair(arr, t)])
            same(o, d, "%s %s" % (op, NAMES[t]))
    b = a.copy()
    b.reshape(-1)[3::7] = O.default_badval(t)
    for flag in (True, False):
        (o, d), = run_both("setbadtonan", [pair(b, t, badflag=flag)])
        same(o, d, "setbadtonan %s flag=%s" % (NAMES[t], flag))
        (o, d), = run_both("setbadtoval", [pair(b, t, badflag=flag)], dpar=-7.25)
        same(o, d, "setbadtoval %s flag=%s" % (NAMES[t], flag))
    # a per-ndarray NaN bad value: ISBAD is isnan
    (o, d), = run_both("setbadtoval", [pair(a, t, badflag=True, badvalue=dt(np.nan))], dpar=3.5)
    same(o, d, "setbadtoval NaN badvalue %s" % NAMES[t])


@pytest.mark.parametrize("t", INT_T, ids=[NAMES[t] for t in INT_T])
def test_setbadtoval_integer_types(t):
    rng = np.random.default_rng(130 + t)
    a = rand(t, (10, 21), rng)
    a.reshape(-1)[::4] = O.default_badval(t)
    (o, d), = run_both("setbadtoval", [pair(a, t, badflag=True)], dpar=42.9)
    same(o, d, "setbadtoval %s" % NAMES[t])
