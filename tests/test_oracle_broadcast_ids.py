"""Explicit broadcast ids on the CPU: broadcastI / broadcast / unbroadcast
views (Slices.pd:1819-1925, Core.pm:1995-2058) in the oracle's ndarrays and
in the package's host headers, pinned by the reference's known answers
(t/thread.t:80-107: the info strings, maximum/minimum over two explicit
dims, unbroadcast's dims), and the oracle's broadcast engine with explicit
ids (pdlbroadcast.c:354-467, 490-500) against brute-force loops and the
error texts PDL gives."""
import numpy as np
import pytest

from pdl_amd import core as C
from oracle import pdlcpu as O


def zeroes(*dims):
    return O.OPdl(np.zeros(int(np.prod(dims)) if dims else 1), O.D, list(dims))


INFO = [  # t/thread.t:84-90
    ((3, 3), ("broadcast", 0, 1), "PDL: Double D [] T1 [3,3]"),
    ((3, 3), ("broadcast", 0), "PDL: Double D [3] T1 [3]"),
    ((4, 7, 2, 8), ("broadcast", 2), "PDL: Double D [4,7,8] T1 [2]"),
    ((4, 7, 2, 8), ("broadcast", 2, 1), "PDL: Double D [4,8] T1 [2,7]"),
    ((4, 7, 2, 8, 5, 6), ("broadcast", 2, 4), "PDL: Double D [4,7,8,6] T1 [2,5]"),
]


def device_broadcast(p, bid, *dims):
    return p.broadcastI(bid, dims)


@pytest.mark.parametrize("dims,call,want", INFO)
def test_info_strings(dims, call, want):
    o = zeroes(*dims).broadcastI(1, call[1:])
    assert "PDL: Double " + o.dimstr() == want
    d = C.shell(C.L.D, dims).broadcast(*call[1:])
    assert d.info() == want


def test_info_two_ids():
    """t/thread.t:90: broadcast1(2)->broadcast2(3)"""
    want = "PDL: Double D [4,7,8,6] T1 [2] T2 [5]"
    o = zeroes(4, 7, 2, 8, 5, 6).broadcastI(1, [2]).broadcastI(2, [3])
    assert "PDL: Double " + o.dimstr() == want
    d = C.shell(C.L.D, [4, 7, 2, 8, 5, 6]).broadcast1(2).broadcast2(3)
    assert d.info() == want
    assert d.broadcastids() == o.broadcastids() == [4, 5, 6]


def test_unbroadcast_dims():
    """t/thread.t:102-106"""
    o = zeroes(4, 5, 6).broadcastI(1, [1]).unbroadcast(2)
    assert o.dims == [4, 6, 5]
    d = C.shell(C.L.D, [4, 5, 6]).broadcast(1).unbroadcast(2)
    assert d.dims == [4, 6, 5] and d.broadcastids() == [3]


def test_views_drop_the_ids_and_duplicates_croak():
    d = C.shell(C.L.D, [3, 4]).broadcast(0)
    assert d.xchg(0, 1).broadcastids() == [2]
    with pytest.raises(C.PDLError) as e:
        C.shell(C.L.D, [3, 4]).broadcast(1, 1)
    assert str(e.value) == "Error in broadcastI:duplicate arg 0 1 1"
    with pytest.raises(O.OracleError) as e:
        zeroes(3, 4).broadcastI(1, [1, 1])
    assert str(e.value) == "Error in broadcastI:duplicate arg 0 1 1"


def thread_t_inputs():
    pa = O.OPdl.from_numpy(np.array([[[0, 1, 2], [3, 4, 5], [6, 7, 8]],
                                     [[10, 11, 12], [13, 14, 15], [16, 17, 18]]], np.float64), O.D)
    pb = zeroes(3, 3)
    return pa, pb


def test_maximum_minimum_over_explicit_dims_known_answers():
    """t/thread.t:91-99: maximum($pa->broadcast(0,1), $pb->broadcast(0,1))
    reduces pa's last dim (2) for each of the 3x3 explicit positions"""
    pa, pb = thread_t_inputs()
    O.run("maximum", pa.broadcastI(1, [0, 1]), pb.broadcastI(1, [0, 1]))
    got = pb.gather()
    assert got[0, 0] == 10 and got[1, 1] == 14
    np.testing.assert_array_equal(got, pa.gather().max(axis=0))
    O.run("minimum", pa.broadcastI(1, [0, 1]), pb.broadcastI(1, [0, 1]))
    got = pb.gather()
    assert got[0, 0] == 0 and got[1, 1] == 4
    np.testing.assert_array_equal(got, pa.gather().min(axis=0))


def test_sumover_with_an_explicit_dim_brute_force():
    """sumover(a(4,5)->broadcast(0), c(4)->broadcast(0)): the reduced dim
    is a's dim 1, the explicit dim pairs a's dim 0 with c's"""
    rng = np.random.default_rng(1)
    a = O.OPdl.from_numpy(rng.integers(-50, 50, (5, 4)).astype(np.int32), O.L)
    c = O.OPdl.from_numpy(np.zeros(4, np.int32), O.L)
    O.run("sumover", a.broadcastI(1, [0]), c.broadcastI(1, [0]))
    ah = a.gather()
    want = [sum(int(ah[j, i]) for j in range(5)) for i in range(4)]
    np.testing.assert_array_equal(c.gather(), want)


def test_explicit_dims_loop_after_the_implicit_ones():
    """plus over a(3,4,2)->broadcast(1) and b(3,2,4)->broadcast(2): the
    explicit dims pair a's dim 1 with b's dim 2 whatever their positions"""
    rng = np.random.default_rng(2)
    a = O.OPdl.from_numpy(rng.standard_normal((2, 4, 3)), O.D)      # PDL dims (3,4,2)
    b = O.OPdl.from_numpy(rng.standard_normal((4, 2, 3)), O.D)      # PDL dims (3,2,4)
    c = O.OPdl.from_numpy(np.zeros((2, 4, 3)), O.D)                  # like a
    O.run("plus", a.broadcastI(1, [1]), b.broadcastI(1, [2]), c.broadcastI(1, [1]))
    ah, bh = a.gather(), b.gather()
    want = np.empty((2, 4, 3))
    for i in range(3):
        for j in range(4):
            for k in range(2):
                want[k, j, i] = ah[k, j, i] + bh[j, k, i]
    np.testing.assert_array_equal(c.gather(), want)


def test_an_output_repeated_over_an_explicit_dim_keeps_the_last_write():
    """an output whose explicit dim is 1 is written once per step of that
    dim (no output check there): the last step's values stay"""
    rng = np.random.default_rng(3)
    a = O.OPdl.from_numpy(rng.standard_normal((5, 3)), O.D)         # PDL dims (3,5)
    c = O.OPdl.from_numpy(np.zeros((1, 3)), O.D)                    # PDL dims (3,1)
    O.run("plus", a.broadcastI(1, [1]), O.OPdl.from_numpy(np.array(1.0), O.D), c.broadcastI(1, [1]))
    np.testing.assert_array_equal(c.gather()[0], a.gather()[4] + 1.0)


def test_the_last_operands_counts_are_looped():
    """pdlbroadcast.c:389-392 takes each explicit id's dim count from the
    last ndarray it reads: with a plain output last, the input's explicit
    dims are not looped and stay at index 0"""
    pa, _ = thread_t_inputs()
    pc = zeroes(3, 3)
    O.run("maximum", pa.broadcastI(1, [0, 1]), pc)
    np.testing.assert_array_equal(pc.gather(), np.full((3, 3), 10.0))


def test_creating_an_output_while_explicitly_broadcasting_croaks():
    a = zeroes(3, 4)
    with pytest.raises(O.OracleError) as e:
        O.run("plus", a.broadcastI(1, [0]), zeroes(4), O.OPdl.null())
    assert str(e.value) == ("plus: trying to create parameter 'c' while explicitly broadcasting.\n"
                            "See the manual for why this is impossible")


def test_mismatched_explicit_dims_croak():
    a = zeroes(3, 4)
    c = zeroes(3, 5)
    with pytest.raises(O.OracleError) as e:
        O.run("assgn", a.broadcastI(1, [1]), c.broadcastI(1, [1]))
    assert "Mismatched implicit broadcast dimension 0: should be 4, is 3" in str(e.value)


def test_clump_dims_of_views():
    """_clump_int's dims (Slices.pd:1376-1392): n past ndims clumps all,
    n < -1 counts from the end, -1 stops at broadcastids[0], whose ids move
    by 1 - n; contiguous dims clump into a view"""
    x = C.shell(C.L.D, [4, 5, 6])
    assert x.clump(2).dims == [20, 6] and x.clump(2).incs == [1, 20]
    assert x.clump(-2).dims == [20, 6] and x.clump(0).dims == [1, 4, 5, 6] and x.clump(7).dims == [120]
    b = x.broadcast(2)                                               # dims (4,5,6), ids [2,3]
    v = b.clump(-1)
    assert v.dims == [20, 6] and v.broadcastids() == [1, 2]
