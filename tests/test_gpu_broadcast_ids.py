"""Explicit broadcast ids on the device (pdlbroadcast.c:354-467 with
broadcastI views, Slices.pd:1819-1898), through the C ABI, against the
oracle on the same inputs: the reference's known answers (t/thread.t:91-99),
binops, reductions, scans and matmult with explicit dims in every operand
position, an output repeated over an explicit dim (PDL's last write stays),
the last operand's counts, and the refusals.  Bars as for every op here:
bit-exact for integers, IEEE elementwise ops, min/max and double products."""
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


def pair(arr, t):
    return O.OPdl.from_numpy(arr, t), C.ndarray.from_numpy(arr, t)


def bi(p, bid, *dims):
    return p.broadcastI(bid, list(dims))


def check_root(o, d):
    """the whole root buffers agree bit for bit (outputs are views here)"""
    while d._parent is not None:
        d = d._parent
    a, b = o.root_pdl().gather(), d.numpy()
    np.testing.assert_array_equal(a.view(np.uint8), b.view(np.uint8))


def test_thread_t_maximum_minimum():
    pa_np = np.array([[[0, 1, 2], [3, 4, 5], [6, 7, 8]], [[10, 11, 12], [13, 14, 15], [16, 17, 18]]], np.float64)
    _, pa = pair(pa_np, O.D)
    pb = C.ndarray.from_numpy(np.zeros((3, 3)), C.L.D)
    C._run("maximum", [pa.broadcast(0, 1), pb.broadcast(0, 1)])
    got = pb.numpy()
    assert got[0, 0] == 10 and got[1, 1] == 14
    C._run("minimum", [pa.broadcast(0, 1), pb.broadcast(0, 1)])
    got = pb.numpy()
    assert got[0, 0] == 0 and got[1, 1] == 4


CASES = [
    # op, input specs [(type, numpy shape, broadcastI args)], output spec
    ("plus", [(O.L, (2, 4, 3), (1, 1)), (O.L, (4, 2, 3), (1, 2))], (O.L, (2, 4, 3), (1, 1))),
    ("mult", [(O.D, (5, 6, 7), (1, 2, 0)), (O.D, (5, 7), (1, 1, 0))], (O.D, (5, 6, 7), (1, 2, 0))),
    ("minus", [(O.F, (3, 8), (2, 0)), (O.F, (3,), None)], (O.F, (3, 8), (2, 0))),
    ("sumover", [(O.LL, (6, 5), (1, 0))], (O.LL, (5,), (1, 0))),
    ("sumover", [(O.L, (4, 3, 9), (1, 0, 1))], (O.L, (3, 9), (1, 0, 1))),
    ("minimum", [(O.S, (7, 5, 2), (1, 2))], (O.S, (7, 5), (1, 1))),
    ("cumusumover", [(O.L, (3, 10), (1, 0))], (O.L, (3, 10), (1, 0))),
    ("maximum_ind", [(O.D, (4, 6), (1, 0))], (O.IND, (6,), (1, 0))),
    ("matmult", [(O.L, (3, 5, 4), (1, 2)), (O.L, (4, 6), None)], (O.L, (3, 5, 6), (1, 2))),
]


def make(spec, rng):
    t, shape, bargs = spec
    dt = O.NP_TYPES[t]
    if np.dtype(dt).kind == "f":
        arr = rng.standard_normal(shape).astype(dt)
    else:
        arr = rng.integers(-1000, 1000, shape).astype(dt)
    o, d = pair(arr, t)
    if bargs is not None:
        o, d = bi(o, bargs[0], *bargs[1:]), bi(d, bargs[0], *bargs[1:])
    return o, d


@pytest.mark.parametrize("case", CASES, ids=["%s-%d" % (c[0], i) for i, c in enumerate(CASES)])
def test_explicit_ids_match_the_oracle(case):
    op, ins, out = case
    rng = np.random.default_rng(CASES.index(case))
    pairs = [make(s, rng) for s in ins]
    oo, dd = make(out, rng)
    O.run(op, *[p[0] for p in pairs], oo)
    C._run(op, [p[1] for p in pairs] + [dd])
    assert oo.dims == dd.dims
    check_root(oo, dd)


def test_an_output_repeated_over_an_explicit_dim():
    """the output's explicit dim is 1 against the input's 5: PDL writes it
    five times, the last (index 4) stays"""
    rng = np.random.default_rng(4)
    a = rng.standard_normal((5, 3)).astype(np.float32)
    oa, da = pair(a, O.F)
    oc, dc = pair(np.zeros((1, 3), np.float32), O.F)
    one_o, one_d = pair(np.array(1.5, np.float32), O.F)
    O.run("plus", oa.broadcastI(1, [1]), one_o, oc.broadcastI(1, [1]))
    C._run("plus", [da.broadcast(1), one_d, dc.broadcast(1)])
    np.testing.assert_array_equal(dc.numpy(), oc.gather())
    np.testing.assert_array_equal(dc.numpy()[0], a[4] + np.float32(1.5))
    # a dummy output dim of size 5 in the explicit id: the same
    oc2, dc2 = pair(np.zeros(3, np.float32), O.F)
    O.run("plus", oa.broadcastI(1, [1]), one_o, oc2.dummy(1, 5).broadcastI(1, [1]))
    C._run("plus", [da.broadcast(1), one_d, dc2.dummy(1, 5).broadcast(1)])
    np.testing.assert_array_equal(dc2.numpy(), oc2.gather())


def test_the_last_operands_counts():
    """a plain output last: the input's explicit dims are not looped"""
    pa_np = np.arange(18, dtype=np.float64).reshape(2, 3, 3)
    oa, da = pair(pa_np, O.D)
    oc, dc = pair(np.zeros((3, 3)), O.D)
    O.run("maximum", oa.broadcastI(1, [0, 1]), oc)
    C._run("maximum", [da.broadcast(0, 1), dc])
    np.testing.assert_array_equal(dc.numpy(), oc.gather())
    np.testing.assert_array_equal(dc.numpy(), np.full((3, 3), 9.0))


def test_refusals():
    da = C.ndarray.from_numpy(np.zeros((4, 3)), C.L.D)
    with pytest.raises(pdl_amd.PDLError) as e:
        C._run("plus", [da.broadcast(0), C.ndarray.from_numpy(np.zeros(4), C.L.D), C.ndarray()])
    assert str(e.value) == ("plus: trying to create parameter 'c' while explicitly broadcasting.\n"
                            "See the manual for why this is impossible")
    # an active dim that is also broadcast would walk past the end
    a = C.ndarray.from_numpy(np.zeros(3), C.L.D).broadcast(0)
    b = C.ndarray.from_numpy(np.zeros((1, 3)), C.L.D).broadcastI(2, [0])
    c = C.ndarray.from_numpy(np.zeros((3, 2)), C.L.D).broadcast(1)
    with pytest.raises(pdl_amd.PDLError) as e:
        C._run("matmult", [a, b, c])
    assert "is both an active dim and in broadcast id" in str(e.value)
    # an input sharing the buffer of an output that is written repeatedly
    x = C.ndarray.from_numpy(np.zeros((5, 3)), C.L.D)
    with pytest.raises(pdl_amd.PDLError) as e:
        C._run("plus", [x.broadcast(1), x.slice(None, (0, 0, 1)).broadcast(1), x.slice(None, (4, 4, 1)).broadcast(1)])
    assert "shares its buffer" in str(e.value)
