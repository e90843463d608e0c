"""Parity of the HIP path (through the C ABI) with the oracle on the same
seeded inputs.  Bars: bit-exact for integer types, indices, bad-value
propagation, IEEE single/double elementwise ops, min/max, float products and
scans (PDL's sequential order); float sums within the stated compensated-sum
bound against an FP64 truth; libm functions (power, atan2) within 2 ulp of
glibc, whose own results the oracle tests hold within 1 ulp of the exact
value."""
import numpy as np
import pytest

import pdl_amd
from pdl_amd import core as C
from oracle import pdlcpu as O

from harness import check_case, load_cases
from device_backend import DeviceBackend

pytestmark = pytest.mark.gpu

DOC = load_cases()
INT_T = [O.SB, O.B, O.S, O.US, O.L, O.UL, O.IND, O.ULL, O.LL]
ALL_T = INT_T + [O.F, O.D]
NAMES = O.TYPE_NAMES


@pytest.fixture(scope="module", autouse=True)
def device():
    C.init(0)
    C.set_matmult_precision("split16")
    yield
    C.set_matmult_precision("split16")


DEVICE_CASES = [c for c in DOC["cases"] if c["op"] != "append"]   # append pins the oracle's planner only


@pytest.mark.parametrize("case", DEVICE_CASES, ids=[c["id"] for c in DEVICE_CASES])
def test_golden_vector_on_device(case):
    err = check_case(DeviceBackend(), case)
    assert err is None, "%s (%s): %s" % (case["id"], case["src"], err)


# ------------------------------------------------------------------ helpers
def rand(t, shape, rng, lo=None, hi=None):
    dt = np.dtype(O.NP_TYPES[t])
    if dt.kind in "iu":
        info = np.iinfo(dt)
        lo = info.min if lo is None else lo
        hi = info.max if hi is None else hi
        return rng.integers(lo, hi, size=shape, endpoint=True, dtype=dt)
    return (rng.standard_normal(shape) * (10 if lo is None else 1)).astype(dt)


def pair(arr, t, badflag=False):
    return O.OPdl.from_numpy(arr, t, badflag=badflag), C.ndarray.from_numpy(arr, t, badflag=badflag)


def view(p, v):
    for step in v:
        p = getattr(p, step[0])(*step[1:])
    return p


def run_both(op, inputs, nout=1, swap=0, m_size=None):
    """inputs: list of (oracle_pdl, device_pdl); returns [(oracle_out, device_out)]"""
    oo = [O.OPdl.null() for _ in range(nout)]
    dd = [C.ndarray() for _ in range(nout)]
    O.run(op, *[i[0] for i in inputs], *oo, swap=swap, m_size=m_size)
    C._run(op, [i[1] for i in inputs] + dd, swap=swap, m_size=m_size)
    return list(zip(oo, dd))


def same(o, d, what=""):
    assert O.TYPE_NAMES[o.dtype] == C.TYPE_NAMES[d.dtype], what
    assert list(o.dims) == list(d.dims), what
    a, b = o.gather(), d.numpy()
    if a.dtype.kind == "f":
        # bit-exact, with every NaN counted equal (NaN sign/payload carries no meaning in PDL)
        u = {4: np.uint32, 8: np.uint64}[a.dtype.itemsize]
        an, bn = np.isnan(a), np.isnan(b)
        np.testing.assert_array_equal(an, bn, err_msg=what)
        np.testing.assert_array_equal(np.where(an, 0, a).view(u), np.where(bn, 0, b).view(u), err_msg=what)
    else:
        np.testing.assert_array_equal(a, b, err_msg=what)
    assert bool(o.state & O.BADVAL) == d.badflag(), what


# ------------------------------------------------------------------ binops
BIOPS_ALL = ["plus", "mult", "minus", "divide", "gt", "lt", "le", "ge", "eq", "ne", "modulo", "spaceship"]
BIOPS_INT = ["shiftleft", "shiftright", "or2", "and2", "xor"]
LAYOUTS = {
    "flat": ([7, 131], [7, 131], []),
    "scalar": ([1000], [], []),
    "bcast": ([33, 1, 5], [33, 6, 1], []),
    "xchg": ([70, 130, 3], [130, 1, 3], [("xchg", 0, 1)]),
    # 11 tile-rows of 64 (one whole group of 8 and a part group), 3 ragged tile-columns
    "xchg_groups": ([300, 700, 2], [700, 1, 2], [("xchg", 0, 1)]),
    # both inputs transposed against the output
    "xchg_both": ([300, 700, 2], [300, 700, 2], [("xchg", 0, 1)], [("xchg", 0, 1)]),
    "slice": ([40, 30], [20, 30], [("slice", (0, 39, 2), None)]),
}


def binop_inputs(op, t, layout, rng):
    sa, sb, va = LAYOUTS[layout][:3]
    vb = LAYOUTS[layout][3] if len(LAYOUTS[layout]) > 3 else []
    a = rand(t, sa[::-1], rng)
    b = rand(t, sb[::-1] if sb else (), rng)
    dt = np.dtype(O.NP_TYPES[t])
    if op in ("divide", "modulo") and dt.kind in "iu":
        b = np.where(b == 0, 1, b).astype(dt)
        if dt.kind == "i":
            b = np.where(b == -1, 3, b).astype(dt)
    if op in ("shiftleft", "shiftright"):
        b = rand(t, sb[::-1] if sb else (), rng, 0, 7)
        if op == "shiftleft" and dt.kind == "i":
            a = np.abs(a // 4).astype(dt)
    if op == "modulo" and dt.kind == "f":
        a = np.round(a * 10).astype(dt)
        b = np.round(b).astype(dt)
    A, Bp = pair(a, t), pair(b, t)
    A = (view(A[0], va), view(A[1], va))
    Bp = (view(Bp[0], vb), view(Bp[1], vb))
    return [A, Bp]


@pytest.mark.parametrize("layout", list(LAYOUTS))
@pytest.mark.parametrize("t", ALL_T, ids=[NAMES[t] for t in ALL_T])
def test_binops_bit_exact(t, layout):
    rng = np.random.default_rng(t * 100 + len(layout))
    ops = BIOPS_ALL + (BIOPS_INT if t in INT_T else [])
    for op in ops:
        for swap in (0, 1):
            ins = binop_inputs(op, t, layout, rng)
            if op in ("divide", "modulo", "shiftleft", "shiftright") and swap:
                continue
            (o, d), = run_both(op, ins, swap=swap)
            same(o, d, "%s %s %s swap=%d" % (op, NAMES[t], layout, swap))


def ulp_distance(x, y):
    """elementwise distance in units in the last place between two float32 or
    float64 arrays (0 for equal values, NaN vs NaN and +0 vs -0)"""
    it = np.int32 if x.dtype == np.float32 else np.int64
    xi, yi = x.view(it).astype(np.int64), y.view(it).astype(np.int64)
    # map the sign-magnitude bit patterns onto one monotone integer line
    lo = np.int64(np.iinfo(it).min)
    xi = np.where(xi < 0, lo - xi, xi)
    yi = np.where(yi < 0, lo - yi, yi)
    d = np.abs(xi.astype(object) - yi.astype(object)).astype(np.float64)
    both_nan = np.isnan(x) & np.isnan(y)
    d[both_nan] = 0
    d[np.isnan(x) ^ np.isnan(y)] = np.inf
    return d


POW_SPECIAL = [(0.0, -3.0), (-0.0, -3.0), (-0.0, -2.0), (0.0, 0.5), (-0.0, 3.0), (-1.0, np.inf),
               (1.0, np.nan), (np.nan, 0.0), (-2.0, 0.5), (0.5, -np.inf), (2.0, np.inf), (-np.inf, -3.0),
               (-np.inf, 3.0), (np.inf, -1.0), (np.nan, 2.0), (-2.0, 3.0), (-2.0, -2.0), (-0.0, 2.0)]
ATAN2_SPECIAL = [(0.0, -0.0), (-0.0, -0.0), (0.0, 0.0), (-0.0, 0.0), (0.0, -1.0), (-0.0, -1.0), (-1.0, 0.0),
                 (1.0, -np.inf), (-1.0, np.inf), (np.inf, -np.inf), (-np.inf, -np.inf), (np.inf, np.inf),
                 (np.nan, 1.0), (1.0, np.nan)]


@pytest.mark.parametrize("t", [O.F, O.D], ids=["float", "double"])
def test_power_atan2_within_two_ulp_of_glibc(t):
    """libm on the device vs glibc (the oracle, itself pinned within 1 ulp of
    the exact value by tests/test_oracle_libm.py): at most 2 ulp apart, the
    C99 Annex F special cases bit-identical.  `x ** 2` is computed as the
    correctly rounded x*x, which glibc's powf(x, 2) misses by one ulp on a
    few elements in ten thousand (the count is printed)."""
    rng = np.random.default_rng(5)
    dt = np.dtype(O.NP_TYPES[t]).type
    n = 4000
    pos = np.exp(rng.uniform(-20, 20, n)).astype(dt)
    neg = -np.exp(rng.uniform(-6, 6, n)).astype(dt)
    exps_real = rng.uniform(-6, 6, n).astype(dt)
    exps_int = rng.integers(-9, 10, n).astype(dt)
    sq = rng.standard_normal(n).astype(dt)
    base = np.concatenate([pos, neg, sq, np.array([c[0] for c in POW_SPECIAL], dt)])
    expo = np.concatenate([exps_real, exps_int, np.full(n, 2.0, dt), np.array([c[1] for c in POW_SPECIAL], dt)])
    ys = np.concatenate([(rng.choice([-1.0, 1.0], n) * np.exp(rng.uniform(-10, 10, n))).astype(dt),
                         np.array([c[0] for c in ATAN2_SPECIAL], dt)])
    xs = np.concatenate([(rng.choice([-1.0, 1.0], n) * np.exp(rng.uniform(-10, 10, n))).astype(dt),
                         np.array([c[1] for c in ATAN2_SPECIAL], dt)])
    for op, a, b, nspec in (("power", base, expo, len(POW_SPECIAL)), ("atan2", ys, xs, len(ATAN2_SPECIAL))):
        (o, d), = run_both(op, [pair(a, t), pair(b, t)])
        x, y = o.gather(), d.numpy()
        dist = ulp_distance(x, y)
        assert dist.max() <= 2, (op, float(dist.max()), a[np.argmax(dist)], b[np.argmax(dist)])
        spec = slice(len(a) - nspec, len(a))
        assert np.all(dist[spec] == 0), (op, x[spec], y[spec])
        xs_, ys_ = x[spec], y[spec]
        keep = ~np.isnan(xs_)
        assert np.array_equal(np.signbit(xs_[keep]), np.signbit(ys_[keep])), (op, xs_, ys_)
        if op == "power":
            print("power: %d of %d squares differ from glibc by one ulp"
                  % (int(np.count_nonzero(dist[2 * n:3 * n])), n))


@pytest.mark.parametrize("t", ALL_T, ids=[NAMES[t] for t in ALL_T])
def test_binops_bad_values(t):
    rng = np.random.default_rng(11 + t)
    a = rand(t, (9, 64), rng)
    b = rand(t, (9, 64), rng)
    bv = O.default_badval(t)
    a.reshape(-1)[::7] = bv
    b.reshape(-1)[::5] = bv
    for op in ("plus", "mult", "gt", "power" if t >= O.F else "or2", "modulo"):
        if op == "modulo" and t >= O.F:
            continue
        bb = np.where(b == 0, 1, b).astype(b.dtype) if op == "modulo" else b
        for fa, fb in ((True, False), (False, True), (True, True)):
            (o, d), = run_both(op, [pair(a, t, fa), pair(bb, t, fb)])
            what = "%s %s bad %s%s" % (op, NAMES[t], fa, fb)
            if op == "power":
                # libm: bad positions exact, values within the stated power() bound
                x, y = o.gather().astype(np.float64), d.numpy().astype(np.float64)
                bvv = np.float64(np.asarray(bv).astype(a.dtype))
                assert np.array_equal(x == bvv, y == bvv), what
                assert ulp_distance(o.gather(), d.numpy()).max() <= 2, what
                assert bool(o.state & O.BADVAL) == d.badflag(), what
            else:
                same(o, d, what)


def test_mixed_types_convert_on_device():
    rng = np.random.default_rng(3)
    m = rng.integers(0, 2, size=(10, 300)).astype(np.uint8)
    x = rng.standard_normal((10, 300)).astype(np.float32)
    (o, d), = run_both("mult", [pair(x, O.F), pair(m, O.B)])
    same(o, d, "float*byte")
    (o, d), = run_both("plus", [pair(m, O.B), pair(rng.integers(-5, 5, (300,)).astype(np.int16), O.S)])
    same(o, d, "byte+short")
    # output of another type: computed in the output's type, written back
    oo, dd = O.OPdl.from_numpy(np.zeros((10, 300)), O.D), C.ndarray.from_numpy(np.zeros((10, 300)), O.D)
    O.run("plus", *pair(x, O.F)[:1], O.OPdl.from_numpy(m, O.B), oo)
    C._run("plus", [pair(x, O.F)[1], C.ndarray.from_numpy(m, O.B), dd])
    same(oo, dd, "into double")


MIXED_PAIRS = [(O.F, O.B), (O.B, O.S), (O.L, O.D), (O.SB, O.UL), (O.US, O.F), (O.LL, O.F), (O.D, O.IND),
               (O.ULL, O.L)]


@pytest.mark.parametrize("ta,tb", MIXED_PAIRS, ids=["%s-%s" % (NAMES[a], NAMES[b]) for a, b in MIXED_PAIRS])
def test_mixed_types_fused_conversion_bit_exact(ta, tb):
    """inputs of other types than the op's are converted in registers
    (one pass, no copies): bit-exact with the oracle's converted copies for
    arithmetic, comparisons, bit ops and a ufunc, through a transposed view
    and a broadcast operand, with bad-flagged sources (their bad elements
    become the computing type's bad value) and an inplace output that
    still takes a copy"""
    rng = np.random.default_rng(300 + 16 * ta + tb)
    a = rand(ta, (24, 70), rng, 0 if np.dtype(O.NP_TYPES[ta]).kind == "u" else -50, 50)
    b = rand(tb, (70,), rng, 1, 40)
    ops = ["plus", "mult", "minus", "gt", "eq", "divide"]
    if ta < O.F and tb < O.F:
        ops += ["or2", "xor", "shiftleft"]
    for op in ops:
        for lay in ((), (("xchg", 0, 1),)):
            A = pair(a, ta)
            A = (view(A[0], lay), view(A[1], lay))
            B = pair(b, tb) if not lay else pair(rng.integers(1, 40, (24,)).astype(O.NP_TYPES[tb]), tb)
            (o, d), = run_both(op, [A, B])
            same(o, d, "%s %s %s %s" % (op, NAMES[ta], NAMES[tb], lay))
    # bad-flagged sources: a bad element of either becomes a bad output
    a2, b2 = a.copy(), b.copy()
    a2.reshape(-1)[::9] = O.default_badval(ta)
    b2[::7] = O.default_badval(tb)
    for fa, fb in ((True, False), (False, True), (True, True)):
        for op in ("plus", "modulo" if ta < O.F and tb < O.F else "power"):
            (o, d), = run_both(op, [pair(a2, ta, fa), pair(b2, tb, fb)])
            what = "%s %s %s bad %s %s" % (op, NAMES[ta], NAMES[tb], fa, fb)
            if op == "power":
                x, y = o.gather(), d.numpy()
                assert x.dtype == y.dtype and ulp_distance(x, y).max() <= 2, what
                assert bool(o.state & O.BADVAL) == d.badflag(), what
            else:
                same(o, d, what)
    # a ufunc of another type's input (sqrt of an integer input computes in its type; exp in double)
    (o, d), = run_both("exp", [pair(np.abs(a) % 20, ta)])
    x, y = o.gather(), d.numpy()
    assert o.dtype == d.dtype and ulp_distance(x, y).max() <= 2
    # inplace: the input that is also the output still converts through a copy
    oa, da = pair(a.copy(), ta)
    O.run("plus", oa, O.OPdl.from_numpy(b, tb), oa)
    C._run("plus", [da, C.ndarray.from_numpy(b, tb), da])
    same(oa, da, "inplace %s %s" % (NAMES[ta], NAMES[tb]))


def test_inplace_through_slice():
    rng = np.random.default_rng(4)
    x = rand(O.L, (6, 50), rng)
    ox, dx = pair(x, O.L)
    ov, dv = ox.slice(None, (1, 5, 2)), dx.slice(None, (1, 5, 2))
    one_o, one_d = pair(np.int8(3), O.SB)
    O.run("plus", ov, one_o, ov)
    C._run("plus", [dv, one_d, dv])
    same(ox, dx, "inplace slice")


# ------------------------------------------------------------------ reductions
RED_EXACT = ["minimum", "maximum", "minimum_ind", "maximum_ind"]


def red_layouts(rng, t):
    return {
        "rows": (rand(t, (37, 500), rng), []),               # wave kernel
        "short": (rand(t, (300, 9), rng), []),                # thread kernel
        "xchg": (rand(t, (400, 70), rng), [("xchg", 0, 1)]),  # strided rows
        "long": (rand(t, (3, 200000), rng), []),              # split kernel
        "rows_long": (rand(t, (40, 3000), rng), []),          # wave kernel, several load groups per lane
        "rows_unaligned": (rand(t, (20, 2051), rng), [("slice", (1, 2050, 1), None)]),
        "empty": (rand(t, (4, 0), rng), []),
    }


def truth_sum(a_view, op, badval=None):
    """sum / mean over PDL dim 0 in long double (numpy pairwise): the truth a
    float result is scored against"""
    g = a_view.gather()
    # PDL's dsumover converts each element to double before adding (tmp += $a())
    x = (g.astype(np.float64) if g.dtype.kind in "iu" and g.dtype.itemsize == 8 else g).astype(np.longdouble)
    good = np.ones(x.shape, bool) if badval is None else (a_view.gather() != np.asarray(badval).astype(a_view.gather().dtype))
    s = np.where(good, x, 0).sum(axis=-1)
    if op in ("average", "daverage"):
        cnt = good.sum(axis=-1)
        # empty rows: 0 for an integer generic type, NaN for a float one (Ufunc.pd:430-431)
        empty = 0.0 if g.dtype.kind in "iu" else np.nan
        with np.errstate(all="ignore"):
            s = np.where(cnt > 0, s / np.maximum(cnt, 1), empty)
    absum = np.where(good, np.abs(x), 0).sum(axis=-1)
    return s.astype(np.float64), absum.astype(np.float64)


def check_sum(d, truth, absum, what, bad_expected=None):
    """|got - truth| <= 2 eps |truth| + 64 eps^2 sum|a|  (compensated sum,
    one rounding to the output type, one more for average's division)"""
    y = d.numpy().astype(np.float64)
    eps = 2.0 ** -23 if d.dtype == O.F else 2.0 ** -52
    bound = 2 * eps * np.abs(truth) + 64 * eps * eps * absum + 1e-300
    ok = (np.abs(truth - y) <= bound) | (np.isnan(truth) & np.isnan(y))
    if bad_expected is not None:
        ok = np.where(bad_expected, y == bad_expected_value(d), ok)
    assert ok.all(), (what, truth[~ok][:5], y[~ok][:5])


def bad_expected_value(d):
    return np.float64(np.asarray(C.default_badvalue(d.dtype)).astype(C.NP_TYPES[d.dtype]))


@pytest.mark.parametrize("t", ALL_T, ids=[NAMES[t] for t in ALL_T])
def test_reductions(t):
    rng = np.random.default_rng(21 + t)
    for lname, (a, v) in red_layouts(rng, t).items():
        if t >= O.F and a.size:
            a.reshape(-1)[::97] = np.nan if lname != "long" else a.reshape(-1)[::97]
        A = pair(a, t)
        A = (view(A[0], v), view(A[1], v))
        ops = ["sumover", "prodover", "dprodover", "dsumover", "average", "daverage", "cumusumover",
               "dcumusumover", "cumuprodover", "dcumuprodover"] + RED_EXACT
        for op in ops:
            if lname == "long" and op in ("cumusumover", "dcumusumover", "cumuprodover", "dcumuprodover",
                                          "prodover", "dprodover"):
                continue
            (o, d), = run_both(op, [A])
            what = "%s %s %s" % (op, NAMES[t], lname)
            floatsum = op in ("sumover", "average", "dsumover", "daverage") and t >= O.F
            if floatsum:
                assert list(o.dims) == list(d.dims) and o.dtype == d.dtype, what
                truth, absum = truth_sum(A[0], op)
                check_sum(d, truth, absum, what)
            elif op in ("dsumover", "daverage") and t in (O.IND, O.ULL, O.LL):
                # 64-bit integers do not convert to double exactly: compare to truth
                truth, absum = truth_sum(A[0], op)
                check_sum(d, truth, absum, what)
            else:
                same(o, d, what)
        outs = run_both("minmaximum", [A], nout=4)
        for k, (o, d) in enumerate(outs):
            same(o, d, "minmaximum %d %s %s" % (k, NAMES[t], lname))


@pytest.mark.parametrize("shape", [(50, 300), (20, 3000)], ids=["short_rows", "long_rows"])
@pytest.mark.parametrize("t", ALL_T, ids=[NAMES[t] for t in ALL_T])
def test_reductions_bad_values(t, shape):
    rng = np.random.default_rng(31 + t)
    a = rand(t, shape, rng)
    bv = O.default_badval(t)
    a.reshape(-1)[::3] = bv
    a[7, :] = bv                       # an all-bad row
    A = pair(a, t, badflag=True)
    for op in ["sumover", "prodover", "average", "cumusumover", "minimum", "maximum", "minimum_ind",
               "maximum_ind", "dsumover"]:
        (o, d), = run_both(op, [A])
        what = "%s %s bad" % (op, NAMES[t])
        if (t >= O.F and op in ("sumover", "average")) or (op == "dsumover" and t in (O.IND, O.ULL, O.LL, O.F, O.D)):
            truth, absum = truth_sum(A[0], op, badval=bv)
            allbad = (a == np.asarray(bv).astype(a.dtype)).all(axis=-1)
            check_sum(d, truth, absum, what, bad_expected=allbad)
            assert d.badflag() == bool(o.state & O.BADVAL), what
        else:
            same(o, d, what)
    for k, (o, d) in enumerate(run_both("minmaximum", [A], nout=4)):
        same(o, d, "minmaximum bad %d" % k)
    for op in ("minimum_n_ind", "maximum_n_ind"):
        (o, d), = run_both(op, [pair(a[:, :40], t, True)], m_size=5)
        same(o, d, op)


# ------------------------------------------------------------------ matmult
def mm_pair(h, t_, w, t, rng, views=((), ())):
    lo = 0 if np.dtype(O.NP_TYPES[t]).kind == "u" else -3
    a = rand(t, (h, t_), rng, lo, 3) if t in INT_T else rand(t, (h, t_), rng)
    b = rand(t, (t_, w), rng, lo, 3) if t in INT_T else rand(t, (t_, w), rng)
    A, B = pair(a, t), pair(b, t)
    A = (view(A[0], views[0]), view(A[1], views[0]))
    B = (view(B[0], views[1]), view(B[1], views[1]))
    return A, B, a, b


@pytest.mark.parametrize("t", ALL_T, ids=[NAMES[t] for t in ALL_T])
def test_matmult_pdl_order_bit_exact(t):
    rng = np.random.default_rng(41 + t)
    C.set_matmult_precision("pdlorder")
    try:
        for (h, tt, w) in [(5, 5, 5), (70, 130, 90), (1, 64, 3), (64, 1, 64), (0, 4, 4)]:
            A, B, _, _ = mm_pair(h, tt, w, t, rng)
            (o, d), = run_both("matmult", [A, B])
            same(o, d, "matmult %s %s" % (NAMES[t], (h, tt, w)))
        # batched over a broadcast dim, transposed operand view
        lo = 0 if np.dtype(O.NP_TYPES[t]).kind == "u" else -3
        a = rand(t, (6, 20, 30), rng, lo, 3)    # a(t=30, h=20) x 6
        b = rand(t, (6, 40, 30), rng, lo, 3)    # (t=30, w=40) -> xchg -> b(w=40, t=30)
        A, B = pair(a, t), pair(b, t)
        B = (B[0].xchg(0, 1), B[1].xchg(0, 1))
        (o, d), = run_both("matmult", [A, B])
        same(o, d, "batched/transposed")
    finally:
        C.set_matmult_precision("split16")


# relative Frobenius error against the FP64 product, per row of C (rows of A
# are scaled over many decades, so a whole-matrix norm would hide small rows)
MM_TOL = {"fp32": 4e-6, "split16": 4e-6, "fp8": 0.08}


@pytest.mark.parametrize("mode", ["fp32", "split16", "fp8"])
@pytest.mark.parametrize("shape", [(300, 257, 190), (600, 100, 130), (1030, 64, 333), (1100, 100, 1030)],
                         ids=["small", "tall", "tall ragged", "256x256 tiles ragged"])
@pytest.mark.parametrize("layout", ["dense", "a transposed"])
def test_matmult_float_modes_accuracy(mode, shape, layout):
    h, k, w = shape
    rng = np.random.default_rng(7 + h)
    C.set_matmult_precision(mode)
    try:
        a = rng.standard_normal((h, k)).astype(np.float32)
        a *= np.exp(3.0 * rng.standard_normal((h, 1))).astype(np.float32)     # rows over ~25 decades
        b = rng.standard_normal((k, w)).astype(np.float32)
        if layout == "dense":
            da = C.ndarray.from_numpy(a)
        else:   # a(t,h) as the transpose of a dense (t,h) array: its k index is strided
            da = C.ndarray.from_numpy(np.ascontiguousarray(a.T)).xchg(0, 1)
        d = C.matmult(da, C.ndarray.from_numpy(b)).numpy().astype(np.float64)
        truth = a.astype(np.float64) @ b.astype(np.float64)
        err = np.linalg.norm(d - truth, axis=1) / np.linalg.norm(truth, axis=1)
        assert err.max() < MM_TOL[mode], (mode, shape, layout, err.max())
    finally:
        C.set_matmult_precision("split16")


def plant_range_and_nonfinite(a, b, rng):
    """Inf/NaN rows of A and columns of B (an Inf meeting a zero gives NaN,
    +Inf and -Inf in one row), and wide-range rows/columns: the element
    pattern A[h] = [1e6, 1e-7, 0...] x B[:, w] = [1e-9, 1e7, 0...] whose
    product 1.001 a single per-row scale would flush to zero, and a row and
    a column spread over ~70 binades with random signs"""
    h, k = a.shape
    w = b.shape[1]
    a[3, 5] = np.inf
    b[5, 6] = 0.0                      # Inf * 0 -> NaN at C[3, 6]
    a[7, 0] = np.nan
    a[9, 10], a[9, 11] = np.inf, -np.inf
    b[2, 4] = -np.inf
    a[12, :] = 0.0
    a[12, 0], a[12, 1] = 1e6, 1e-7
    b[:, 13] = 0.0
    b[0, 13], b[1, 13] = 1e-9, 1e7
    a[14, :] = (rng.choice([-1.0, 1.0], k) * np.exp(rng.uniform(-24, 24, k))).astype(np.float32)
    b[:, 15] = (rng.choice([-1.0, 1.0], k) * np.exp(rng.uniform(-24, 24, k))).astype(np.float32)
    return {"nonfinite_rows": [3, 7, 9], "nonfinite_cols": [4], "range_rows": [12, 14], "range_cols": [13, 15]}


def check_range_and_nonfinite(d, a, b, marks, mode, what):
    """d (float32, h x w) against PDL's FP32 loop (the oracle) and the FP64
    product: rows/columns holding a non-finite element equal PDL's loop bit
    for bit; every finite output of split16 is within 2^-21 |c| +
    2^-19 sum|a||b| of the exact value (the range-loss recompute bound)"""
    oa, ob, oc = O.OPdl.from_numpy(a, O.F), O.OPdl.from_numpy(b, O.F), O.OPdl.null()
    O.run("matmult", oa, ob, oc)
    pdl = oc.gather()
    for r in marks["nonfinite_rows"]:
        np.testing.assert_array_equal(d[r], pdl[r], err_msg="%s row %d" % (what, r))
    for col in marks["nonfinite_cols"]:
        np.testing.assert_array_equal(d[:, col], pdl[:, col], err_msg="%s col %d" % (what, col))
    fin_a = np.isfinite(a).all(axis=1)
    fin_b = np.isfinite(b).all(axis=0)
    a64 = np.where(np.isfinite(a), a, 0).astype(np.float64)
    b64 = np.where(np.isfinite(b), b, 0).astype(np.float64)
    truth = a64 @ b64
    absp = np.abs(a64) @ np.abs(b64)
    ok = np.outer(fin_a, fin_b)
    if mode == "split16":
        # an element is either within the split bound or recomputed by
        # PDL's loop (then equal to the oracle's)
        err = np.abs(d.astype(np.float64) - truth)
        bound = 2.0 ** -21 * np.abs(truth) + 2.0 ** -19 * absp + 1e-37
        bad = ok & ~(err <= bound) & ~(d == pdl)
        assert not bad.any(), (what, np.argwhere(bad)[:5], d[bad][:5], truth[bad][:5])
        pdl_err = np.abs(pdl.astype(np.float64) - truth)
        for r in marks["range_rows"]:
            for col in marks["range_cols"]:
                assert err[r, col] <= max(2.0 ** -20 * absp[r, col], pdl_err[r, col]), (what, r, col, d[r, col])
    else:
        # FP8: rows and columns without planted values keep the FP8 bound
        rows = [r for r in range(a.shape[0]) if r not in sum(marks.values(), [])]
        cols = [c for c in range(b.shape[1]) if c not in sum(marks.values(), [])]
        sub, tr = d[np.ix_(rows, cols)].astype(np.float64), truth[np.ix_(rows, cols)]
        err = np.linalg.norm(sub - tr, axis=1) / np.linalg.norm(tr, axis=1)
        assert err.max() < MM_TOL["fp8"], (what, err.max())
    assert np.isfinite(d[ok]).all(), what


@pytest.mark.parametrize("mode", ["split16", "fp8"])
@pytest.mark.parametrize("shape", [(300, 257, 190), (1100, 100, 1030)], ids=["128x128 blocks", "256x256 blocks"])
@pytest.mark.parametrize("layout", ["dense", "a transposed"])
def test_matmult_float_nonfinite_and_wide_range(mode, shape, layout):
    h, k, w = shape
    rng = np.random.default_rng(900 + h)
    a = rng.standard_normal((h, k)).astype(np.float32)
    b = rng.standard_normal((k, w)).astype(np.float32)
    marks = plant_range_and_nonfinite(a, b, rng)
    C.set_matmult_precision(mode)
    try:
        if layout == "dense":
            da = C.ndarray.from_numpy(a)
        else:
            da = C.ndarray.from_numpy(np.ascontiguousarray(a.T)).xchg(0, 1)
        with np.errstate(invalid="ignore", over="ignore"):
            d = C.matmult(da, C.ndarray.from_numpy(b)).numpy()
            check_range_and_nonfinite(d, a, b, marks, mode, "%s %s %s" % (mode, shape, layout))
    finally:
        C.set_matmult_precision("split16")


def test_matmult_small_batches_nonfinite_and_wide_range():
    """the one-block-per-product kernel (64x64x64 and 40x40x40 batches):
    the same planted rows/columns in some products of the batch"""
    rng = np.random.default_rng(77)
    for n in (64, 40):
        nb = 6
        a = rng.standard_normal((nb, n, n)).astype(np.float32)
        b = rng.standard_normal((nb, n, n)).astype(np.float32)
        marks = {}
        for i in (1, 4):
            marks[i] = plant_range_and_nonfinite(a[i], b[i], rng)
        with np.errstate(invalid="ignore", over="ignore"):
            d = C.matmult(C.ndarray.from_numpy(a), C.ndarray.from_numpy(b)).numpy()
            for i in range(nb):
                m = marks.get(i, {"nonfinite_rows": [], "nonfinite_cols": [], "range_rows": [], "range_cols": []})
                check_range_and_nonfinite(d[i], a[i], b[i], m, "split16", "small %d product %d" % (n, i))


@pytest.mark.parametrize("shape", [(64, 64, 64), (17, 33, 50), (48, 16, 64), (64, 20, 31), (16, 16, 16), (15, 40, 40)],
                         ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("layout", ["dense", "a transposed", "b broadcast"])
def test_matmult_small_batches_split16(shape, layout):
    """broadcast batches of small products (16 <= h, k, w <= 64 run one block
    per product on the matrix cores; 15 stays on the FP32 path): per-row
    relative error against the float64 product, rows of A over ~25 decades"""
    h, k, w = shape
    nb = 37
    rng = np.random.default_rng(100 + h + 3 * k + 7 * w)
    a = rng.standard_normal((nb, h, k)).astype(np.float32)
    a *= np.exp(3.0 * rng.standard_normal((nb, h, 1))).astype(np.float32)
    b = rng.standard_normal((k, w) if layout == "b broadcast" else (nb, k, w)).astype(np.float32)
    if layout == "a transposed":   # a(t,h,n) as the transpose of a dense (h,t,n) array
        da = C.ndarray.from_numpy(np.ascontiguousarray(a.transpose(0, 2, 1))).xchg(0, 1)
    else:
        da = C.ndarray.from_numpy(a)
    d = C.matmult(da, C.ndarray.from_numpy(b)).numpy().astype(np.float64)
    truth = np.matmul(a.astype(np.float64), b.astype(np.float64))
    err = np.linalg.norm(d - truth, axis=2) / np.linalg.norm(truth, axis=2)
    assert d.shape == truth.shape
    assert err.max() < MM_TOL["split16"], (shape, layout, err.max())


@pytest.mark.parametrize("shape", [(300, 257, 190), (1100, 100, 1030), (1030, 40, 1100), (129, 33, 70)],
                         ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("layout", ["dense", "a transposed"])
def test_matmult_double_default_accuracy(shape, layout):
    """double products in the default mode (FP64 matrix cores once there are
    64 or more 128x128 blocks, the SIMT kernel below): per-row relative error
    against a long-double product"""
    h, k, w = shape
    rng = np.random.default_rng(500 + h + k)
    a = rng.standard_normal((h, k)) * np.exp(3.0 * rng.standard_normal((h, 1)))
    b = rng.standard_normal((k, w))
    if layout == "dense":
        da = C.ndarray.from_numpy(a, O.D)
    else:
        da = C.ndarray.from_numpy(np.ascontiguousarray(a.T), O.D).xchg(0, 1)
    d = C.matmult(da, C.ndarray.from_numpy(b, O.D)).numpy()
    truth = a.astype(np.longdouble) @ b.astype(np.longdouble)
    err = np.linalg.norm((d - truth).astype(np.float64), axis=1) / np.linalg.norm(truth.astype(np.float64), axis=1)
    assert err.max() < 2e-14, (shape, layout, err.max())


def test_matmult_bad_values_exact():
    rng = np.random.default_rng(8)
    for t in (O.F, O.D, O.L):
        a = rand(t, (20, 33), rng, -3, 3)
        b = rand(t, (33, 17), rng, -3, 3)
        if t >= O.F:
            a, b = np.round(a), np.round(b)
        a[3, 5] = O.default_badval(t)
        b[10, 4] = O.default_badval(t)
        (o, d), = run_both("matmult", [pair(a, t, True), pair(b, t)])
        same(o, d, "matmult bad %s" % NAMES[t])


@pytest.mark.parametrize("t", [O.F, O.D, O.L], ids=["float", "double", "long"])
def test_matmult_list_matches_single_products(t):
    """one work-queue launch over uneven products == each product alone (PDL order: bit-exact vs the oracle)"""
    rng = np.random.default_rng(12)
    C.set_matmult_precision("pdlorder")
    try:
        sizes = [(int(rng.integers(1, 90)), int(rng.integers(1, 90)), int(rng.integers(1, 90))) for _ in range(40)]
        As, Bs, oas, obs = [], [], [], []
        for h, tt, w in sizes:
            a = rand(t, (h, tt), rng, -3, 3)
            b = rand(t, (tt, w), rng, -3, 3)
            As.append(C.ndarray.from_numpy(a, t)); Bs.append(C.ndarray.from_numpy(b, t))
            oas.append(O.OPdl.from_numpy(a, t)); obs.append(O.OPdl.from_numpy(b, t))
        for mode in ("queue", "each"):
            outs = C.matmult_list(As, Bs, mode=mode)
            for k in range(len(sizes)):
                oc = O.OPdl.null()
                O.run("matmult", oas[k], obs[k], oc)
                same(oc, outs[k], "list %s item %d %s" % (mode, k, sizes[k]))
    finally:
        C.set_matmult_precision("split16")


def test_prepared_matmult_list_launches_and_graph_replays():
    """a prepared list: queue and per-product launches equal the oracle's
    products (PDL order, bit-exact); a HIP graph of one launch, replayed with
    the inputs changed in between, equals an eager launch"""
    rng = np.random.default_rng(13)
    C.set_matmult_precision("pdlorder")
    try:
        sizes = [(int(rng.integers(1, 140)), int(rng.integers(1, 140)), int(rng.integers(1, 140))) for _ in range(30)]
        As, Bs, ah, bh = [], [], [], []
        for h, tt, w in sizes:
            a = rand(O.F, (h, tt), rng)
            b = rand(O.F, (tt, w), rng)
            ah.append(a); bh.append(b)
            As.append(C.ndarray.from_numpy(a, O.F)); Bs.append(C.ndarray.from_numpy(b, O.F))
        ml = C.MatmultList(As, Bs)
        for mode in ("queue", "each"):
            ml.launch(mode)
            for k in range(len(sizes)):
                oc = O.OPdl.null()
                O.run("matmult", O.OPdl.from_numpy(ah[k], O.F), O.OPdl.from_numpy(bh[k], O.F), oc)
                same(oc, ml.outputs[k], "prepared %s item %d %s" % (mode, k, sizes[k]))
        g = C.Graph(lambda: ml.launch("queue"))
        for rep in range(3):
            C._run("plus", [As[rep], C.topdl(1), As[rep]])
            C.synchronize()
            g.launch()
            C.synchronize()
            got = [o.numpy().copy() for o in ml.outputs]
            ml.launch("queue")
            C.synchronize()
            for k in range(len(sizes)):
                assert got[k].tobytes() == ml.outputs[k].numpy().tobytes(), (rep, k)
        g.destroy()
        ml.free()
    finally:
        C.set_matmult_precision("split16")


# ------------------------------------------------------------------ readdata entry
def test_readdata_entry_runs_a_planned_broadcast():
    """pdlhip_readdata over a broadcast struct built as PDL's redodims would"""
    import ctypes
    from pdl_amd import _lib as L
    rng = np.random.default_rng(9)
    a = rng.standard_normal((8, 16)).astype(np.float32)
    b = rng.standard_normal((16,)).astype(np.float32)
    da, db = C.ndarray.from_numpy(a), C.ndarray.from_numpy(b)
    dc = C.ndarray.new_dense(O.F, [16, 8])
    brc = L.Broadcast()
    brc.npdls, brc.ndims = 3, 2
    brc.dims[0], brc.dims[1] = 16, 8
    incs = [1, 1, 1, 16, 0, 16]          # incs[d*npdls + p]
    for k, v in enumerate(incs):
        brc.incs[k] = v
    cs = [da._to_c(), db._to_c(), dc._to_c()]
    arr = (L.PP * 3)(*[ctypes.pointer(c) for c in cs])
    L.check(L.lib().pdlhip_readdata(b"plus", O.F, 0, ctypes.byref(brc), arr, None))
    np.testing.assert_array_equal(dc.numpy(), a + b[None, :])


def readdata_with_oracle_plan(op, o_ins, d_ins, nout, m_size=None):
    """runs op in the oracle, capturing the broadcast plan PDL's redodims
    built (dims, incs, offsets, core ind_sizes / inc_sizes, trans type,
    bvalflag), then hands that plan to pdlhip_readdata over device copies
    of the same views and outputs of the created dims and types (with the
    bad flag PDL sets on outputs before the loop, pdlapi.c:806-808)"""
    import ctypes
    from pdl_amd import _lib as L
    outs_o = [O.OPdl.null() for _ in range(nout)]
    plan = O.run_capturing_plan(op, *o_ins, *outs_o, m_size=m_size)
    outs_d = []
    for o in outs_o:
        d = C.ndarray.new_dense(o.dtype, o.dims)
        d.badflag(bool(o.state & O.BADVAL))
        outs_d.append(d)
    brc = L.Broadcast()
    brc.npdls, brc.ndims = plan.npdls, plan.ndims
    for d in range(plan.ndims):
        brc.dims[d] = plan.dims[d]
    for k in range(plan.ndims * plan.npdls):
        brc.incs[k] = plan.incs[k]
    for p in range(plan.npdls):
        brc.offs[p] = plan.offs[p]
    for k in range(3):
        brc.ind_sizes[k] = plan.ind_sizes[k]
    brc.ind_sizes[3] = -1
    for k in range(12):
        brc.inc_sizes[k] = plan.inc_sizes[k]
    cs = [x._to_c() for x in list(d_ins) + outs_d]
    arr = (L.PP * len(cs))(*[ctypes.pointer(c) for c in cs])
    other = (ctypes.c_int64 * 1)(m_size or 0)
    L.check(L.lib().pdlhip_readdata(op.encode(), plan.datatype, plan.bvalflag, ctypes.byref(brc), arr, other))
    return outs_o, outs_d


def test_readdata_entry_with_pdl_plans_for_reductions_scans_and_matmult():
    """pdlhip_readdata (the vtable-hook entry the XS shim installs) over the
    plans PDL's own redodims builds: sumover of a transposed int32 view and
    of a bad-valued int16 array, cumusumover of a strided slice,
    minimum_n_ind, minmaximum, and a batched matmult with a transposed
    operand (PDL order): bit-exact with the oracle"""
    rng = np.random.default_rng(19)
    cases = []
    x = rng.integers(-2**20, 2**20, (40, 300), dtype=np.int32)
    cases.append(("sumover", [(O.OPdl.from_numpy(x, O.L).xchg(0, 1), C.ndarray.from_numpy(x, O.L).xchg(0, 1))], 1, None))
    s16 = rng.integers(-300, 300, (30, 700)).astype(np.int16)
    s16.reshape(-1)[::13] = -32768
    s16[4, :] = -32768
    cases.append(("sumover", [(O.OPdl.from_numpy(s16, O.S, badflag=True),
                               C.ndarray.from_numpy(s16, O.S, badflag=True))], 1, None))
    y = rng.integers(-1000, 1000, (20, 500), dtype=np.int32)
    sl = (None, (3, 17, 2))
    cases.append(("cumusumover", [(O.OPdl.from_numpy(y, O.L).slice(*sl), C.ndarray.from_numpy(y, O.L).slice(*sl))],
                  1, None))
    f = rng.standard_normal((25, 400)).astype(np.float32)
    f.reshape(-1)[::97] = np.nan
    cases.append(("minimum_n_ind", [(O.OPdl.from_numpy(f, O.F), C.ndarray.from_numpy(f, O.F))], 1, 3))
    cases.append(("minmaximum", [(O.OPdl.from_numpy(f, O.F), C.ndarray.from_numpy(f, O.F))], 4, None))
    for op, ins, nout, m in cases:
        outs_o, outs_d = readdata_with_oracle_plan(op, [i[0] for i in ins], [i[1] for i in ins], nout, m_size=m)
        for k, (o, d) in enumerate(zip(outs_o, outs_d)):
            same(o, d, "readdata %s output %d" % (op, k))
    C.set_matmult_precision("pdlorder")
    try:
        a = rng.standard_normal((6, 20, 30)).astype(np.float32)     # a(t=30, h=20) x 6
        b = rng.standard_normal((6, 40, 30)).astype(np.float32)     # -> xchg -> b(w=40, t=30)
        ins = [(O.OPdl.from_numpy(a, O.F), C.ndarray.from_numpy(a, O.F)),
               (O.OPdl.from_numpy(b, O.F).xchg(0, 1), C.ndarray.from_numpy(b, O.F).xchg(0, 1))]
        outs_o, outs_d = readdata_with_oracle_plan("matmult", [i[0] for i in ins], [i[1] for i in ins], 1)
        same(outs_o[0], outs_d[0], "readdata matmult")
    finally:
        C.set_matmult_precision("split16")


def test_streaming_kernels_at_256_mib_and_above():
    """the nontemporal (streaming) branches the kernels take for operands of
    256 MiB and more, at the headline workload's shape with 4 k-planes:
    z = x->xchg(0,1) * y->dummy(1) on float32 (4096,4096,4) bit-exact against
    numpy's float32 product; sumover/average of z within the compensated
    bound and minimum/maximum/minimum_ind exact; a 256 MiB flat binop with a
    scalar, a 256 MiB assgn copy, and a 64 Mi-element single-row sumover"""
    rng = np.random.default_rng(2026)
    nx, ny, nk = 4096, 4096, 4
    xn = rng.standard_normal((nk, ny, nx), dtype=np.float32)      # x(i, j, k) = xn[k, j, i]
    yn = rng.standard_normal((nk, nx), dtype=np.float32)          # y(i, k) = yn[k, i]
    x, y = C.ndarray.from_numpy(xn, O.F), C.ndarray.from_numpy(yn, O.F)
    z = C.ndarray.new_dense(O.F, [nx, ny, nk])
    C._run("mult", [x.xchg(0, 1), y.dummy(1), z])
    # z(i, j, k) = x(j, i, k) * y(i, k)  ->  zn[k, j, i] = xn[k, i, j] * yn[k, i]
    want = np.transpose(xn, (0, 2, 1)) * yn[:, None, :]
    got = z.numpy()
    assert got.shape == want.shape
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), "streaming ew_tile"
    del got
    for op in ("sumover", "average"):
        o = C.ndarray.new_dense(O.F, [ny, nk])
        C._run(op, [z, o])
        tr = want.astype(np.float64).sum(axis=2)
        absum = np.abs(want).astype(np.float64).sum(axis=2)
        if op == "average":
            tr, absum = tr / nx, absum / nx
        y64 = o.numpy().astype(np.float64)
        bound = 2 * 2.0 ** -23 * np.abs(tr) + 64 * 2.0 ** -46 * absum + 1e-30
        assert np.all(np.abs(y64 - tr) <= bound), op
    for op, f in (("minimum", np.min), ("maximum", np.max)):
        o = C.ndarray.new_dense(O.F, [ny, nk])
        C._run(op, [z, o])
        assert np.array_equal(o.numpy(), f(want, axis=2)), op
    oi = C.ndarray.new_dense(O.IND, [ny, nk])
    C._run("minimum_ind", [z, oi])
    assert np.array_equal(oi.numpy(), np.argmin(want, axis=2)), "minimum_ind"
    del want
    # flat binop with a scalar (ew_flat streaming) and a 256 MiB copy (k_copy_flat streaming)
    w = C.ndarray.new_dense(O.F, [nx, ny, nk])
    C._run("mult", [x, C.topdl(1.5), w])
    assert np.array_equal(w.numpy().view(np.uint32), (xn * np.float32(1.5)).view(np.uint32)), "streaming ew_flat"
    cp = C.ndarray.new_dense(O.F, [nx, ny, nk])
    C._run("assgn", [x, cp])
    assert np.array_equal(cp.numpy().view(np.uint32), xn.view(np.uint32)), "streaming assgn"
    # one 64 Mi-element row (the split reduction, streaming)
    flat = C.ndarray.from_numpy(xn.reshape(-1), O.F)
    s = C.ndarray.new_dense(O.F, [])
    C._run("sumover", [flat, s])
    tr = float(xn.astype(np.float64).sum())
    absum = float(np.abs(xn).astype(np.float64).sum())
    assert abs(float(s.numpy()) - tr) <= 2 * 2.0 ** -23 * abs(tr) + 64 * 2.0 ** -46 * absum, "split sumover"


@pytest.mark.parametrize("t", [O.F, O.D], ids=["float", "double"])
def test_matmult_list_default_mode_accuracy(t):
    """the benchmarked list kernels (default mode, not PDL order): 96 square
    products, n uniform in [8, 256], queue and per-product launches, each
    row of every product within the float-mode bound of the exact product"""
    rng = np.random.default_rng(31 + t)
    ns = rng.integers(8, 256, size=96, endpoint=True)
    dt = np.float32 if t == O.F else np.float64
    ah = [rng.uniform(-1, 1, (n, n)).astype(dt) for n in ns]
    bh = [rng.uniform(-1, 1, (n, n)).astype(dt) for n in ns]
    As = [C.ndarray.from_numpy(a, t) for a in ah]
    Bs = [C.ndarray.from_numpy(b, t) for b in bh]
    tol = 4e-6 if t == O.F else 2e-14
    ml = C.MatmultList(As, Bs)
    for mode in ("queue", "each"):
        ml.launch(mode)
        for k, n in enumerate(ns):
            truth = ah[k].astype(np.longdouble) @ bh[k].astype(np.longdouble)
            got = ml.outputs[k].numpy().astype(np.longdouble)
            err = (np.linalg.norm((got - truth).astype(np.float64), axis=1) /
                   np.linalg.norm(truth.astype(np.float64), axis=1))
            assert err.max() < tol, (mode, k, int(n), float(err.max()))
    ml.free()


def test_graph_capture_keeps_scratch_and_refuses_what_cannot_be_captured():
    """a float matmult (FP16 packing scratch) and a long-row sumover
    (per-chunk partials) captured in one graph: eager ops of the same shapes
    that run between replays take blocks from the same cache, and every
    replay still equals the eager result bit for bit.  An op whose scratch
    is not cached, a bad-value minimum_n_ind and a download are refused
    while capturing."""
    from pdl_amd._lib import PDLError
    rng = np.random.default_rng(23)
    n = 512
    ah, bh = rng.standard_normal((n, n)).astype(np.float32), rng.standard_normal((n, n)).astype(np.float32)
    xh = rng.standard_normal((3, 300000)).astype(np.float32)
    a, b, x = C.ndarray.from_numpy(ah), C.ndarray.from_numpy(bh), C.ndarray.from_numpy(xh)
    c, s = C.ndarray.new_dense(O.F, [n, n]), C.ndarray.new_dense(O.F, [3])
    C._run("matmult", [a, b, c])           # eager first: the scratch is then cached
    C._run("sumover", [x, s])
    c_eager, s_eager = c.numpy().copy(), s.numpy().copy()

    def body():
        C._run("matmult", [a, b, c])
        C._run("sumover", [x, s])

    g = C.Graph(body)
    zero = C.topdl(0.0)
    for rep in range(3):
        # eager work of the same shapes with other data, reusing cached blocks
        a2 = C.ndarray.from_numpy(rng.standard_normal((n, n)).astype(np.float32))
        c2 = C.ndarray.new_dense(O.F, [n, n])
        C._run("matmult", [a2, b, c2])
        x2 = C.ndarray.from_numpy(rng.standard_normal((3, 300000)).astype(np.float32))
        s2 = C.ndarray.new_dense(O.F, [3])
        C._run("sumover", [x2, s2])
        C._run("mult", [c, zero, c])
        C._run("mult", [s, zero, s])
        g.launch()
        C.synchronize()
        assert c.numpy().tobytes() == c_eager.tobytes(), rep
        assert s.numpy().tobytes() == s_eager.tobytes(), rep
        del a2, c2, x2, s2
    g.destroy()
    # refused while capturing
    big_a = C.ndarray.from_numpy(rng.standard_normal((3000, 3000)).astype(np.float32))
    big_c = C.ndarray.new_dense(O.F, [3000, 3000])
    with pytest.raises(PDLError, match="not cached while a graph is being captured"):
        C.Graph(lambda: C._run("matmult", [big_a, big_a, big_c]))
    sh = rng.integers(-50, 50, (8, 40)).astype(np.int16)
    sh[::3, ::5] = -32768
    sb = C.ndarray.from_numpy(sh, O.S, badflag=True)
    mi = C.ndarray.new_dense(O.IND, [3, 8])
    with pytest.raises(PDLError, match="cannot be captured"):
        C.Graph(lambda: C._run("minimum_n_ind", [sb, mi], m_size=3))
    with pytest.raises(PDLError, match="not allowed while a graph is being captured"):
        C.Graph(lambda: c.numpy())
    # the library still works eagerly afterwards
    C._run("matmult", [a, b, c])
    assert c.numpy().tobytes() == c_eager.tobytes()


def test_device_allocation_failure_is_a_fatal_error_and_leaves_buffers_intact():
    """an output the device cannot hold (a null output broadcast to 2^40
    elements by a dummy dim) fails as PDL_EFATAL with PDL's kind of message
    (pdl.h.PL:370-379; pdlmagic.c:288-296 / pdlutil.c:519-531 report such
    failures as fatal, not user errors); the inputs keep their data and
    bad flags, the library keeps working, and a direct allocation that
    cannot succeed fails the same way"""
    import ctypes
    from pdl_amd import _lib as L
    from pdl_amd._lib import PDLError
    rng = np.random.default_rng(77)
    ah = rng.standard_normal((3, 5))
    a = C.ndarray.from_numpy(ah, O.D)
    b = C.ndarray.from_numpy(ah[0], O.D, badflag=True)
    huge = a.dummy(2, 1 << 40)
    out = C.ndarray()
    with pytest.raises(PDLError) as ei:
        C._run("plus", [huge, b, out])
    assert ei.value.kind == L.EFATAL and "out of device memory" in str(ei.value)
    assert out.isnull()
    np.testing.assert_array_equal(a.numpy(), ah)
    np.testing.assert_array_equal(b.numpy(), ah[0])
    assert b.badflag() and not a.badflag()
    p = ctypes.c_void_p()
    e = L.lib().pdlhip_alloc(1 << 50, ctypes.byref(p))
    assert e.error == L.EFATAL and b"out of device memory" in e.message
    # still working
    np.testing.assert_array_equal((a + a).numpy(), ah + ah)
