"""Float and double complex ndarrays on the device (k_complex.hip) against
the oracle's C99 complex arithmetic, through the C ABI: plus, minus, mult,
eq, ne with broadcasting, views, swapped and inplace operands and real
operands converted to complex; infinities and NaNs through the Annex G
product; divide through libgcc's Smith's method and its recovery; assgn
and conversions to and from real types; matmult (the
reference's known answer, t/primitive-matmult.t:32-34, batches, views,
the PMCode's scalar shortcut and mismatch text).  Bar: bit-exact, every NaN
counted equal.  Sums and products (sumover, prodover and their cumulative
forms) against the oracle in PDL's sequential order.  Refusals: bad values
in complex ndarrays and the ops the device does not take complex for
(power; average computes in long double for a complex input)."""
import numpy as np
import pytest

import pdl_amd
from pdl_amd import core as C
from oracle import pdlcpu as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    C.init(0)
    yield


NPT = {O.CF: np.complex64, O.CD: np.complex128}


def crand(rng, shape, t, special=False):
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(NPT[t])
    if special:
        flat = x.reshape(-1)
        vals = [complex(np.inf, 0), complex(np.inf, np.inf), complex(np.nan, 1), complex(0, np.nan),
                complex(-np.inf, 2), complex(1e38, 1e38), complex(np.nan, np.nan)]
        for k, v in enumerate(vals):
            if k * 7 < flat.size:
                flat[k * 7] = v
    return x


def same(o, d):
    assert o.dtype == d.dtype and list(o.dims) == list(d.dims)
    a, b = o.gather(), d.numpy()
    assert a.dtype == b.dtype
    fl = a.real.dtype
    ar, ai = np.ascontiguousarray(a.real), np.ascontiguousarray(a.imag)
    br, bi = np.ascontiguousarray(b.real), np.ascontiguousarray(b.imag)
    for x, y in ((ar, br), (ai, bi)):
        nx, ny = np.isnan(x), np.isnan(y)
        assert np.array_equal(nx, ny)
        u = {4: np.uint32, 8: np.uint64}[fl.itemsize]
        assert np.array_equal(x[~nx].view(u), y[~ny].view(u))


@pytest.mark.parametrize("t", [O.CF, O.CD], ids=["cfloat", "cdouble"])
@pytest.mark.parametrize("op", ["plus", "minus", "mult", "divide", "eq", "ne"])
def test_elementwise_bit_exact(t, op):
    rng = np.random.default_rng(10 * t + ["plus", "minus", "mult", "divide", "eq", "ne"].index(op))
    x = crand(rng, (6, 50, 33), t, special=True)
    y = crand(rng, (50, 1), t, special=True)              # broadcast along dim 0, size-1 dim
    y[3, 0] = x[0, 3, 0]
    ox, dx = O.OPdl.from_numpy(x, t), C.ndarray.from_numpy(x, t)
    oy, dy = O.OPdl.from_numpy(y, t), C.ndarray.from_numpy(y, t)
    for swap in (0, 1):
        oc, dc = O.OPdl.null(), C.ndarray()
        O.run(op, ox, oy, oc, swap=swap)
        C._run(op, [dx, dy, dc], swap=swap)
        same(oc, dc)
    # a transposed view of x, a real (float) second operand converted to complex
    r = rng.standard_normal((50, 6)).astype(np.float32)
    oc, dc = O.OPdl.null(), C.ndarray()
    O.run(op, ox.xchg(0, 2), O.OPdl.from_numpy(r, O.F), oc)
    C._run(op, [dx.xchg(0, 2), C.ndarray.from_numpy(r, O.F), dc])
    same(oc, dc)


@pytest.mark.parametrize("t", [O.CF, O.CD], ids=["cfloat", "cdouble"])
def test_divide_special_and_wide_operands(t):
    """every pair of a grid of special parts (signed zeros, infinities, NaN,
    the type's extremes) and values of wide exponent range through libgcc's
    Smith's method and its Annex G recovery"""
    rt = np.float32 if t == O.CF else np.float64
    fi = np.finfo(rt)
    parts = np.array([0.0, -0.0, 1.0, -2.5, np.inf, -np.inf, np.nan, fi.max, fi.tiny, fi.tiny * fi.eps], rt)
    vals = (parts[:, None] + 1j * parts[None, :]).reshape(-1).astype(NPT[t])
    x = np.repeat(vals, vals.size)
    y = np.tile(vals, vals.size)
    rng = np.random.default_rng(12 + t)
    span = 60 if t == O.CF else 400
    w = lambda: (rng.standard_normal(x.size) * 2.0 ** rng.integers(-span, span, x.size)).astype(rt)
    x2, y2 = (w() + 1j * w()).astype(NPT[t]), (w() + 1j * w()).astype(NPT[t])
    for xa, ya in ((x, y), (x2, y2)):
        oc, dc = O.OPdl.null(), C.ndarray()
        O.run("divide", O.OPdl.from_numpy(xa, t), O.OPdl.from_numpy(ya, t), oc)
        C._run("divide", [C.ndarray.from_numpy(xa, t), C.ndarray.from_numpy(ya, t), dc])
        same(oc, dc)


def test_inplace_and_integer_operands():
    rng = np.random.default_rng(9)
    x = crand(rng, (4, 100), O.CD)
    k = rng.integers(-50, 50, (100,)).astype(np.int32)
    ox, dx = O.OPdl.from_numpy(x, O.CD), C.ndarray.from_numpy(x, O.CD)
    O.run("mult", ox, O.OPdl.from_numpy(k, O.L), ox)
    C._run("mult", [dx, C.ndarray.from_numpy(k, O.L), dx])
    same(ox, dx)


@pytest.mark.parametrize("src,dst", [(O.CD, O.CF), (O.CF, O.CD), (O.CD, O.LL), (O.CF, O.UL), (O.CD, O.D),
                                     (O.S, O.CF), (O.D, O.CD), (O.CF, O.CF)])
def test_assgn_and_conversions(src, dst):
    rng = np.random.default_rng(src * 16 + dst)
    if src in NPT:
        x = (crand(rng, (3, 40), src) * 100).astype(NPT[src])
        if dst in (O.UL,):
            x = (np.abs(x.real) + 1j * x.imag).astype(NPT[src])
    else:
        x = (rng.standard_normal((3, 40)) * 100).astype(O.NP_TYPES[src])
    onp = NPT.get(dst, O.NP_TYPES.get(dst))
    oo, do = O.OPdl.from_numpy(np.zeros((40, 3), onp), dst), C.ndarray.from_numpy(np.zeros((40, 3), onp), dst)
    O.run("assgn", O.OPdl.from_numpy(x, src).xchg(0, 1), oo)
    C._run("assgn", [C.ndarray.from_numpy(x, src).xchg(0, 1), do])
    if dst in NPT:
        same(oo, do)
    else:
        np.testing.assert_array_equal(oo.gather(), do.numpy())


def test_matmult_known_answer_and_mismatch():
    """t/primitive-matmult.t:32-37"""
    cm1 = C.ndarray.from_numpy(np.array([1, 1 + 1j, 1]), O.CD)
    cm2 = C.ndarray.from_numpy(np.array([2, 3, 1j]), O.CD).transpose()
    got = C.matmult(cm1, cm2)
    assert got.dims == [1, 1] and got.numpy().tolist() == [[5 + 4j]]
    with pytest.raises(pdl_amd.PDLError) as e:
        C.matmult(cm1.transpose(), cm2)
    assert "mismatch" in str(e.value)


@pytest.mark.parametrize("t", [O.CF, O.CD], ids=["cfloat", "cdouble"])
def test_matmult_bit_exact(t):
    rng = np.random.default_rng(11 + t)
    a = crand(rng, (3, 37, 29), t)            # a(t=29, h=37) x 3
    b = crand(rng, (3, 29, 41), t)            # b(w=41, t=29) x 3
    oc, dc = O.OPdl.null(), C.ndarray()
    O.run("matmult", O.OPdl.from_numpy(a, t), O.OPdl.from_numpy(b, t), oc)
    C._run("matmult", [C.ndarray.from_numpy(a, t), C.ndarray.from_numpy(b, t), dc])
    same(oc, dc)
    # transposed views, a broadcast operand, a real operand converted
    bt = np.ascontiguousarray(np.swapaxes(b[0], 0, 1))        # (41, 29)
    r = rng.standard_normal((37, 29)).astype(np.float64)
    oc, dc = O.OPdl.null(), C.ndarray()
    O.run("matmult", O.OPdl.from_numpy(r, O.D), O.OPdl.from_numpy(bt, t).xchg(0, 1), oc)
    C._run("matmult", [C.ndarray.from_numpy(r, O.D), C.ndarray.from_numpy(bt, t).xchg(0, 1), dc])
    same(oc, dc)
    # special values in the products
    a2 = crand(rng, (16, 16), t, special=True)
    b2 = crand(rng, (16, 16), t, special=True)
    oc, dc = O.OPdl.null(), C.ndarray()
    O.run("matmult", O.OPdl.from_numpy(a2, t), O.OPdl.from_numpy(b2, t), oc)
    C._run("matmult", [C.ndarray.from_numpy(a2, t), C.ndarray.from_numpy(b2, t), dc])
    same(oc, dc)
    # the PMCode's scalar shortcut: a 1x1 operand multiplies elementwise
    s = C.ndarray.from_numpy(np.array([[2 - 1j]]), t)
    got = C.matmult(s, C.ndarray.from_numpy(b[0], t))
    want = O.OPdl.null()
    O.run("mult", O.OPdl.from_numpy(np.array([[2 - 1j]]), t), O.OPdl.from_numpy(b[0], t), want)
    same(want, got)


@pytest.mark.parametrize("t", [O.CF, O.CD], ids=["cfloat", "cdouble"])
@pytest.mark.parametrize("op", ["sumover", "prodover", "cumusumover", "cumuprodover"])
def test_sums_and_products_bit_exact(t, op):
    """Ufunc.pd:88-140 in PDL's sequential order: rows of a transposed view,
    a broadcast over three dims, zeros (prodover stops at a zero product),
    infinities and NaNs"""
    rng = np.random.default_rng(40 + t)
    with np.errstate(invalid="ignore"):
        x = (crand(rng, (3, 70, 45), t, special=True) * 0.9).astype(NPT[t])
    x[1, 5, 7] = 0
    x[2, 9, :] = 0
    ox, dx = O.OPdl.from_numpy(x, t), C.ndarray.from_numpy(x, t)
    for v in ((), (("xchg", 0, 1),)):
        o, d = ox, dx
        for step in v:
            o, d = getattr(o, step[0])(*step[1:]), getattr(d, step[0])(*step[1:])
        oc, dc = O.OPdl.null(), C.ndarray()
        O.run(op, o, oc)
        C._run(op, [d, dc])
        same(oc, dc)


def test_refusals():
    x = C.ndarray.from_numpy(np.ones(4, np.complex128), O.CD)
    with pytest.raises(pdl_amd.PDLError) as e:
        C._run("power", [x, x, C.ndarray()])
    assert "not supported on the device" in str(e.value)
    with pytest.raises(pdl_amd.PDLError) as e:
        C._run("average", [x, C.ndarray()])
    assert "not supported on the device" in str(e.value)
    xb = C.ndarray.from_numpy(np.ones(4, np.complex128), O.CD, badflag=True)
    with pytest.raises(pdl_amd.PDLError) as e:
        C._run("plus", [xb, x, C.ndarray()])
    assert "bad values" in str(e.value)


def test_readdata_entry_with_pdl_plans():
    """pdlhip_readdata (the vtable hook of the XS shim) over the plans PDL's
    redodims builds for complex ops: a broadcast cfloat plus with a
    transposed operand and a batched cdouble matmult"""
    from test_gpu_parity import readdata_with_oracle_plan
    rng = np.random.default_rng(77)
    x = crand(rng, (5, 30, 40), O.CF)
    y = crand(rng, (30, 1), O.CF)
    ins = [(O.OPdl.from_numpy(x, O.CF).xchg(0, 1), C.ndarray.from_numpy(x, O.CF).xchg(0, 1)),
           (O.OPdl.from_numpy(y, O.CF).xchg(0, 1), C.ndarray.from_numpy(y, O.CF).xchg(0, 1))]
    outs_o, outs_d = readdata_with_oracle_plan("plus", [i[0] for i in ins], [i[1] for i in ins], 1)
    same(outs_o[0], outs_d[0])
    a = crand(rng, (4, 9, 11), O.CD)
    b = crand(rng, (4, 11, 7), O.CD)
    ins = [(O.OPdl.from_numpy(a, O.CD), C.ndarray.from_numpy(a, O.CD)),
           (O.OPdl.from_numpy(b, O.CD), C.ndarray.from_numpy(b, O.CD))]
    outs_o, outs_d = readdata_with_oracle_plan("matmult", [i[0] for i in ins], [i[1] for i in ins], 1)
    same(outs_o[0], outs_d[0])
