"""clump / flat (Slices.pd:1373-1400 _clump_int, Core.pm flat): the first n
dims as one, PDL's element order kept, a child with dataflow: contiguous dims
clump into a view whose writes land in the parent; others into a dense copy
that follows the parent and writes back into it.  Checked against numpy's
reshape of the same data in PDL order (numpy shape = reversed dims) on the
device."""
import numpy as np
import pytest

from pdl_amd import core as C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    C.init(0)
    yield


def test_clump_of_dense_dims_is_a_view():
    xn = np.arange(120, dtype=np.float64).reshape(6, 5, 4)          # PDL dims (4,5,6)
    x = C.ndarray.from_numpy(xn, C.L.D)
    c = x.clump(2)
    assert c.dims == [20, 6]
    np.testing.assert_array_equal(c.numpy(), xn.reshape(6, 20))
    c.assign(C.ndarray.from_numpy(-np.arange(120, dtype=np.float64).reshape(6, 20), C.L.D))
    np.testing.assert_array_equal(x.numpy(), -xn)
    assert x.clump(-2).dims == [20, 6] and x.clump(0).dims == [1, 4, 5, 6] and x.clump(9).dims == [120]
    f = x.flat()
    assert f.dims == [120]
    f.set(7, 1000.0)
    assert x.at(3, 1, 0) == 1000.0


def test_clump_of_a_transposed_view_copies_with_dataflow():
    xn = np.arange(120, dtype=np.int32).reshape(6, 5, 4)             # PDL dims (4,5,6)
    x = C.ndarray.from_numpy(xn, C.L.L)
    t = x.xchg(0, 1)                                                 # dims (5,4,6): not contiguous
    c = t.clump(2)
    assert c.dims == [20, 6]
    want = np.swapaxes(xn, 1, 2).reshape(6, 20)                      # t in PDL order
    np.testing.assert_array_equal(c.numpy(), want)
    # the parent changes: the child follows
    x.assign(C.ndarray.from_numpy(xn * 2, C.L.L))
    np.testing.assert_array_equal(c.numpy(), want * 2)
    # the child is written: the parent takes the values back
    c.assign(C.ndarray.from_numpy(np.full((6, 20), 9, np.int32), C.L.L))
    np.testing.assert_array_equal(x.numpy(), np.full((6, 5, 4), 9, np.int32))
    c.sever()
    c.assign(C.ndarray.from_numpy(np.zeros((6, 20), np.int32), C.L.L))
    np.testing.assert_array_equal(x.numpy(), np.full((6, 5, 4), 9, np.int32))


def test_flat_follows_broadcast_ids_and_sum_uses_it():
    xn = np.arange(120, dtype=np.float32).reshape(6, 5, 4)
    x = C.ndarray.from_numpy(xn, C.L.F)
    b = x.broadcast(1)                                               # dims (4,6,5), ids [2,3]
    c = b.clump(-1)                                                  # clumps up to broadcastids[0]
    assert c.dims == [24, 5] and c.broadcastids() == [1, 2]
    np.testing.assert_array_equal(c.numpy(), np.swapaxes(xn, 0, 1).reshape(5, 24))
    s = x.xchg(0, 2).sum()
    assert float(s.numpy()) == float(xn.astype(np.float64).sum())
