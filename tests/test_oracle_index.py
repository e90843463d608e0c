"""Pins the oracle's Slices.pd index / index1d / index2d (Slices.pd:142-220)
and their BackCode on the CPU: against brute-force loops over the same
inputs in PDL's loop order (dim 0 fastest, m innermost; a scatter's last
writer wins), against the reference's own known answers (t/slice.t:45-51:
the transposed lut indexed by a dummy index), and against the $CROAK texts
the PP code gives (Slices.pd:142-147, 174-177, 199-206; PP.pm:371)."""
import numpy as np
import pytest

from oracle import pdlcpu as O


def run(op, *arrs, types=None, badflags=None):
    types = types or [None] * len(arrs)
    badflags = badflags or [False] * len(arrs)
    ps = [O.OPdl.from_numpy(a, t, badflag=b) for a, t, b in zip(arrs, types, badflags)]
    out = O.OPdl.null()
    O.run(op, *ps, out)
    return out


def test_index_gathers_along_dim0_with_broadcast():
    rng = np.random.default_rng(1)
    a = rng.integers(-100, 100, (5, 7)).astype(np.int32)        # PDL a(7) x 5 broadcast
    ind = rng.integers(0, 7, (3, 5)).astype(np.int64)           # ind() broadcast (5,3)
    c = run("index", a, ind, types=[O.L, O.IND]).gather()
    want = np.empty((3, 5), np.int32)
    for j in range(3):
        for i in range(5):
            want[j, i] = a[i, ind[j, i]]
    np.testing.assert_array_equal(c, want)


def test_index1d_and_index2d_brute_force():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((4, 9))                               # a(9) x 4
    ind = rng.integers(0, 9, (4, 6)).astype(np.int64)             # ind(6) x 4
    c = run("index1d", a, ind, types=[O.D, O.IND]).gather()
    want = np.array([[a[i, ind[i, m]] for m in range(6)] for i in range(4)])
    np.testing.assert_array_equal(c, want)
    b = rng.integers(0, 50, (3, 5, 8)).astype(np.int16)           # a(8,5) x 3
    ia = rng.integers(0, 8, (3,)).astype(np.int64)
    ib = rng.integers(0, 5, (3,)).astype(np.int64)
    c2 = run("index2d", b, ia, ib, types=[O.S, O.IND, O.IND]).gather()
    np.testing.assert_array_equal(c2, [b[k, ib[k], ia[k]] for k in range(3)])


def test_backcode_scatters_and_the_last_writer_wins():
    rng = np.random.default_rng(3)
    a = np.zeros((2, 10))
    ind = rng.integers(0, 10, (2, 25)).astype(np.int64)           # repeats
    c = rng.standard_normal((2, 25))
    pa, pi, pc = O.OPdl.from_numpy(a, O.D), O.OPdl.from_numpy(ind, O.IND), O.OPdl.from_numpy(c, O.D)
    O.run("index1d_back", pa, pi, pc)
    want = a.copy()
    for i in range(2):
        for m in range(25):
            want[i, ind[i, m]] = c[i, m]
    np.testing.assert_array_equal(pa.gather(), want)


def test_transposed_lut_known_answer():
    """t/slice.t:45-51: lut = [[1,0],[0,1]]; in = lut->transpose->index(im->dummy(0)),
    im = indx [1] -> [[0,1]]; in .= 1 flows back: lut = [[1,0],[1,1]]"""
    lut = np.array([[1.0, 0.0], [0.0, 1.0]])
    plut = O.OPdl.from_numpy(lut, O.D)
    a = plut.xchg(0, 1)
    im = O.OPdl.from_numpy(np.array([1], np.int64), O.IND).dummy(0)
    c = O.OPdl.null()
    O.run("index", a, im, c)
    np.testing.assert_array_equal(c.gather(), [[0.0, 1.0]])
    ones = O.OPdl.from_numpy(np.ones((1, 2)), O.D)
    O.run("index_back", a, im, ones)
    np.testing.assert_array_equal(plut.gather(), [[1.0, 0.0], [1.0, 1.0]])


@pytest.mark.parametrize("op,args,msg", [
    ("index", (np.arange(4.0), np.array([1, 5], np.int64)), "Error in index:invalid index 5 (valid range 0..3)"),
    ("index", (np.arange(4.0), np.array([-1], np.int64)), "Error in index:invalid index -1 (valid range 0..3)"),
    ("index1d", (np.arange(6.0), np.array([0, 2, 9], np.int64)),
     "Error in index1d:invalid index 9 at pos 2 (valid range 0..5)"),
    ("index2d", (np.zeros((3, 4)), np.array([4], np.int64), np.array([0], np.int64)),
     "Error in index2d:invalid x-index 4 (valid range 0..3)"),
    ("index2d", (np.zeros((3, 4)), np.array([1], np.int64), np.array([3], np.int64)),
     "Error in index2d:invalid y-index 3 (valid range 0..2)"),
])
def test_croak_texts(op, args, msg):
    types = [O.D] + [O.IND] * (len(args) - 1)
    with pytest.raises(O.OracleError) as e:
        run(op, *args, types=types)
    assert str(e.value) == msg


def test_bad_indices():
    """index croaks on a bad index (in bad mode); index1d sets c bad instead"""
    bv = -2 ** 63
    a = np.arange(5.0)
    ind = np.array([1, bv, 3], np.int64)
    with pytest.raises(O.OracleError) as e:
        run("index", a, ind, types=[O.D, O.IND], badflags=[False, True])
    assert str(e.value) == "Error in index:invalid index %d (valid range 0..4)" % bv
    c = run("index1d", a, ind, types=[O.D, O.IND], badflags=[False, True])
    got = c.gather()
    assert got[0] == 1.0 and got[2] == 3.0 and got[1] == O.default_badval(O.D)
    assert c.state & O.BADVAL
    # without the bad flag the same value is only out of range
    with pytest.raises(O.OracleError) as e:
        run("index1d", a, ind, types=[O.D, O.IND])
    assert "invalid index %d at pos 1" % bv in str(e.value)
