"""Slices.pd index / index1d / index2d on the device (Slices.pd:142-220)
against the oracle, through the C ABI, and the dataflow and host-mirror
behaviour of the Python mirror of PDL's API around them.

Bars: bit-exact values, types and bad flags for the gathers and the
BackCode scatters (repeated indices: the last position in PDL's loop order
wins); PDL's exact $CROAK texts for the first failing position; the
reference's own cases from t/slice.t (the transposed lut, dice sums, a
mutated indexed slice, x .= x->index(idx))."""
import numpy as np
import pytest

from pdl_amd import core as C
from pdl_amd import _lib as L
from oracle import pdlcpu as O

pytestmark = pytest.mark.gpu

NAMES = O.TYPE_NAMES
TYPES = [O.SB, O.B, O.S, O.US, O.L, O.UL, O.IND, O.ULL, O.LL, O.F, O.D]


@pytest.fixture(scope="module", autouse=True)
def device():
    C.init(0)
    yield


def rand(t, shape, rng):
    if t in (O.F, O.D):
        return rng.standard_normal(shape).astype(O.NP_TYPES[t])
    info = np.iinfo(O.NP_TYPES[t])
    return rng.integers(max(info.min, -1000), min(info.max, 1000), shape, endpoint=True).astype(O.NP_TYPES[t])


def pair(arr, t, badflag=False):
    return O.OPdl.from_numpy(arr, t, badflag=badflag), C.ndarray.from_numpy(arr, t, badflag=badflag)


def same(o, d, what=""):
    assert NAMES[o.dtype] == C.TYPE_NAMES[d.dtype], what
    assert list(o.dims) == list(d.dims), what
    a, b = o.gather(), d.numpy()
    np.testing.assert_array_equal(a.view(np.uint8), b.view(np.uint8), err_msg=what)
    assert bool(o.state & O.BADVAL) == bool(d.state & L.BADVAL), what


def both(op, inputs):
    oo, dd = O.OPdl.null(), C.ndarray()
    O.run(op, *[i[0] for i in inputs], oo)
    C._run(op, [i[1] for i in inputs] + [dd])
    return oo, dd


@pytest.mark.parametrize("t", TYPES, ids=[NAMES[t] for t in TYPES])
def test_gathers_bit_exact(t):
    rng = np.random.default_rng(100 + t)
    a = rand(t, (6, 37), rng)                                         # a(37) x 6
    # index: a broadcast index over (6, 4), and through a transposed view
    ind = rng.integers(0, 37, (4, 6)).astype(np.int64)
    same(*both("index", [pair(a, t), pair(ind, O.IND)]), "index")
    at = rng.integers(0, 6, (37,)).astype(np.int64)
    oa, da = pair(a, t)
    same(*both("index", [(oa.xchg(0, 1), da.xchg(0, 1)), pair(at, O.IND)]), "index xchg")
    # index1d with an index list per broadcast row, long (int32) indices converted to indx
    i1 = rng.integers(0, 37, (6, 50)).astype(np.int32)
    same(*both("index1d", [pair(a, t), pair(i1, O.L)]), "index1d")
    # index2d
    ia = rng.integers(0, 37, (5, 3)).astype(np.int64)
    ib = rng.integers(0, 6, (5, 3)).astype(np.int64)
    same(*both("index2d", [pair(a, t), pair(ia, O.IND), pair(ib, O.IND)]), "index2d")


@pytest.mark.parametrize("op", ["index", "index1d", "index2d"])
def test_backcode_bit_exact_with_repeats(op):
    rng = np.random.default_rng(7)
    a = rng.standard_normal((5, 40)).astype(np.float32)               # a(40) x 5
    if op == "index":
        inds = [rng.integers(0, 40, (300, 5)).astype(np.int64)]      # 300 writes per row: repeats
        c = rng.standard_normal((300, 5)).astype(np.float32)
    elif op == "index1d":
        inds = [rng.integers(0, 40, (5, 300)).astype(np.int64)]
        c = rng.standard_normal((5, 300)).astype(np.float32)
    else:
        inds = [rng.integers(0, 40, (700,)).astype(np.int64), rng.integers(0, 5, (700,)).astype(np.int64)]
        c = rng.standard_normal((700,)).astype(np.float32)
    oa, da = pair(a, O.F)
    oi, di = zip(*[pair(i, O.IND) for i in inds])
    oc, dc = pair(c, O.F)
    O.run(op + "_back", oa, *oi, oc)
    cs = [p._to_c() for p in [da, *di, dc]]
    arr = (L.PP * len(cs))(*[C.ctypes.pointer(x) for x in cs])
    L.check(L.lib().pdlhip_writebackdata(op.encode(), arr, len(cs)))
    np.testing.assert_array_equal(oa.gather().view(np.uint32), da.numpy().view(np.uint32))


@pytest.mark.parametrize("op,args,msg", [
    ("index", (np.arange(400.0), np.array([1, 5, 400, 2, 999], np.int64)),
     "Error in index:invalid index 400 (valid range 0..399)"),
    ("index1d", (np.arange(60.0).reshape(3, 20), np.array([[0, 2, 9], [1, 25, -4], [3, 3, 3]], np.int64)),
     "Error in index1d:invalid index 25 at pos 1 (valid range 0..19)"),
    ("index2d", (np.zeros((3, 4)), np.array([1, 4], np.int64), np.array([0, 0], np.int64)),
     "Error in index2d:invalid x-index 4 (valid range 0..3)"),
    ("index2d", (np.zeros((3, 4)), np.array([1, 2], np.int64), np.array([2, 3], np.int64)),
     "Error in index2d:invalid y-index 3 (valid range 0..2)"),
])
def test_croak_texts_first_position(op, args, msg):
    ins = [pair(args[0], O.D)] + [pair(x, O.IND) for x in args[1:]]
    with pytest.raises(O.OracleError) as eo:
        both(op, ins)
    assert str(eo.value) == msg
    with pytest.raises(L.PDLError) as ed:
        C._run(op, [i[1] for i in ins] + [C.ndarray()])
    assert str(ed.value) == msg and ed.value.kind == L.EUSERERROR


def test_bad_indices_like_the_oracle():
    bv = -2 ** 63
    a = np.arange(50.0).reshape(5, 10)
    ind = np.array([[1, bv, 3], [bv, 0, 9], [2, 2, 2], [9, 8, bv], [0, 0, 0]], np.int64)
    same(*both("index1d", [pair(a, O.D), pair(ind, O.IND, badflag=True)]), "index1d bad")
    # index over the same rows: ind() broadcast (5,3); loop order meets ind[1][0] (bad) second
    it = np.ascontiguousarray(ind.T)
    with pytest.raises(L.PDLError) as e:
        C._run("index", [C.ndarray.from_numpy(a), C.ndarray.from_numpy(it, O.IND, badflag=True), C.ndarray()])
    assert str(e.value) == "Error in index:invalid index %d (valid range 0..9)" % bv
    with pytest.raises(O.OracleError) as eo:
        O.run("index", O.OPdl.from_numpy(a, O.D), O.OPdl.from_numpy(it, O.IND, badflag=True), O.OPdl.null())
    assert str(eo.value) == str(e.value)


# ------------------------------------------------------------ dataflow (t/slice.t)
def test_transposed_lut_flows_back():
    """t/slice.t:45-51"""
    lut = C.pdl([[1, 0], [0, 1]])
    im = C.ndarray.from_numpy(np.array([1], np.int64), L.IND)
    inn = lut.transpose().index(im.dummy(0))
    np.testing.assert_array_equal(inn.numpy(), [[0, 1]])
    inn.assign(1)
    np.testing.assert_array_equal(inn.numpy(), [[1, 1]])
    np.testing.assert_array_equal(lut.numpy(), [[1, 0], [1, 1]])


def test_dice_sums():
    """t/slice.t:53-56"""
    x = C.sequence(10, 4)
    assert float(x.dice([1, 2], [0, 3]).sumover().sumover().sclr()) == 66
    assert float(x.dice([0, 1], "X").sumover().sumover().sclr()) == 124


def test_mutating_an_indexed_slice():
    """t/slice.t:393-397: s = indices->index(0); s %= 7 changes only that column"""
    ind = C.ndarray.from_numpy(np.array([[31, 31]], np.int64), L.IND)
    s = ind.index(0)
    s %= 7
    np.testing.assert_array_equal(ind.numpy(), [[3, 31]])


def test_assign_from_own_index():
    """t/slice.t:399-405: x .= x->index(idx) runs without error, and x then
    holds x's old values at idx"""
    x = C.sequence(1000)
    C.srandom(5)
    idx = (C.random(1000) * 1000).convert(L.IND)
    want = np.arange(1000.0)[idx.numpy()]
    x.assign(x.index(idx))
    np.testing.assert_array_equal(x.numpy(), want)
    assert float(x.minimum().sclr()) == want.min()


def test_child_follows_its_parent_and_sever_detaches():
    a = C.ndarray.from_numpy(np.arange(10.0))
    c = a.index1d(C.ndarray.from_numpy(np.array([2, 7], np.int64), L.IND))
    np.testing.assert_array_equal(c.numpy(), [2, 7])
    a += 100                                     # the parent changes: the child is computed again when read
    np.testing.assert_array_equal(c.numpy(), [102, 107])
    c.set(0, -1)                                 # an element written through the mirror flows back
    assert a.at(2) == -1
    c.sever()
    a += 1
    np.testing.assert_array_equal(c.numpy(), [-1, 107])


def test_host_mirror_reads_once_and_writes_before_the_next_op():
    a = C.ndarray.from_numpy(np.arange(12.0).reshape(3, 4))
    b = a * 2
    assert b.at(2, 1) == 12.0 and b.at(3, 0) == 6.0
    b.set(0, 0, 5.0)                             # host write, uploaded before the next device op reads b
    assert b.at(0, 0) == 5.0
    s = b.sumover()
    np.testing.assert_array_equal(s.numpy(), [5 + 2 + 4 + 6, 8 + 10 + 12 + 14, 16 + 18 + 20 + 22])
    b += 1                                       # the device writes b: the mirror is out of date
    assert b.at(0, 0) == 6.0


# ------------------------------------------------------------ range / rangeb / indexND
def source10():
    return (10 * np.arange(10)[None, :] + np.arange(10)[:, None]).astype(np.float64)   # x(i,j) = 10 i + j


IDX = np.array([[[2, 3], [4, 5]], [[6, 7], [8, 9]]], np.int64)


def test_indexND_known_answers():
    """t/slice.t:179-195"""
    s = C.ndarray.from_numpy(source10())
    np.testing.assert_array_equal(s.indexND(IDX).numpy(), [[23, 45], [67, 89]])
    z = np.zeros((2, 10, 10))
    src3 = (100 * np.arange(10)[None, None, :] + 10 * np.arange(10)[None, :, None] + np.arange(2)[:, None, None] + z)
    x = C.ndarray.from_numpy(src3).indexND(IDX)
    np.testing.assert_array_equal(x.numpy(), [[[230, 450], [670, 890]], [[231, 451], [671, 891]]])
    e = C.ndarray.from_numpy(source10()).indexND(C.ndarray.new_dense(L.IND, [2, 0]))
    assert e.dims == [0]


@pytest.mark.parametrize("size,bound,sl,want", [
    (None, None, None, [[23, 45], [67, 89]]),
    (3, "t", (1, 1), [[89, 99, 0], [0, 0, 0], [0, 0, 0]]),
    (3, "tp", (1, 1), [[89, 99, 0], [80, 90, 0], [81, 91, 0]]),
    (3, ["e", "p"], (1, 1), [[89, 99, 99], [80, 90, 90], [81, 91, 91]]),
])
def test_range_known_answers(size, bound, sl, want):
    """t/slice.t:196-233"""
    s = C.ndarray.from_numpy(source10())
    y = s.range(IDX, size, bound)
    assert y.dims == ([2, 2] if size is None else [2, 2, 3, 3])
    got = y.numpy() if sl is None else y.numpy()[:, :, sl[1], sl[0]]
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(s.numpy(), source10())       # the source is not mutated


def test_range_out_of_bounds_and_periodic_writeback():
    s = C.ndarray.from_numpy(source10())
    with pytest.raises(L.PDLError) as e:
        s.range(IDX, 3)
    assert str(e.value) == "Error in rangeb:index out-of-bounds in range (index vector #3)"
    s3 = C.ndarray.from_numpy((10 * np.arange(3)[None, :] + np.arange(3)[:, None]).astype(np.float64))
    y = s3.range([-1, -1], [2, 2], "p")
    np.testing.assert_array_equal(y.numpy(), [[22, 2], [20, 0]])
    y.assign(6)                                                   # flows back through the periodic corners
    np.testing.assert_array_equal(s3.numpy(), [[6, 10, 6], [1, 11, 21], [6, 12, 6]])


BOUNDS = ["f", "t", "e", "p", "m", "tp", "em", "pe"]


@pytest.mark.parametrize("bound", BOUNDS)
@pytest.mark.parametrize("t", [O.S, O.F, O.D], ids=["short", "float", "double"])
def test_range_against_the_oracle(t, bound):
    """random corners (in and out of range), sizes with a zero (a dropped
    dim), source broadcast dims past the index length, bad values; forward
    and BackCode bit-exact with the oracle's rangeb"""
    rng = np.random.default_rng(900 + t + 7 * BOUNDS.index(bound))
    src = rand(t, (3, 7, 9), rng)                                # parent dims (9,7,3): dims past rdim=2 broadcast
    lo, hi = (0, 4) if bound == "f" else (-12, 14)
    ind = rng.integers(lo, hi, (5, 4, 2)).astype(np.int64)        # index dims (2,4,5): 20 vectors
    for size in ([3, 0], [4, 2], [1]):
        for badflag in (False, True):
            s2 = src.copy()
            if badflag:
                s2.reshape(-1)[::11] = O.default_badval(t)
            op_, dp_ = pair(s2, t, badflag)
            oi, di = pair(ind, O.IND)
            oc = O.rangeb(op_, oi, size, bound)
            dc = dp_.rangeb(di, size, bound)
            same(oc, dc, "%s %s %s" % (bound, size, badflag))
            # BackCode: new child values into copies of the sources
            newc = rand(t, tuple(oc.dims[::-1]), rng)
            op2, dp2 = pair(s2, t, badflag)
            oc2 = O.OPdl.from_numpy(newc, t, badflag=badflag)
            O.rangeb(op2, oi, size, bound, back_child=oc2)
            dc2 = C.ndarray.from_numpy(newc, t, badflag=badflag)
            cp, ci, cc = dp2._to_c(), di._to_c(), dc2._to_c()
            arr = np.asarray(size, np.int64)
            L.check(L.lib().pdlhip_writeback_rangeb(C.ctypes.byref(cp), C.ctypes.byref(ci),
                                                     (C.ctypes.c_int64 * arr.size)(*arr.tolist()), arr.size,
                                                     bound.encode(), C.ctypes.byref(cc)))
            np.testing.assert_array_equal(op2.gather().view(np.uint8), dp2.numpy().view(np.uint8))


def test_range_size_and_boundary_errors():
    s = C.ndarray.from_numpy(source10())
    o = O.OPdl.from_numpy(source10(), O.D)
    idx = np.array([[1, 2], [3, 4]], np.int64)
    for size, bound, msg in (([-1], "", "Error in rangeb:Negative range size is not allowed\n"),
                             ([2, -1], "", "Error in rangeb:Negative range sizes are not allowed\n"),
                             ([1, 2, 3], "", "Error in rangeb:Size must match index's 0th dim\n"),
                             ([1], "q", "Error in rangeb:Unknown boundary condition 'q' in range")):
        with pytest.raises(O.OracleError) as eo:
            O.rangeb(o, O.OPdl.from_numpy(idx, O.IND), size, bound)
        assert str(eo.value) == msg
        with pytest.raises(L.PDLError) as ed:
            s.rangeb(C.ndarray.from_numpy(idx, L.IND), size, bound)
        assert str(ed.value) == msg
