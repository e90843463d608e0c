"""pdlhip_run_reduce_noalloc (the allocation-free variant of the Ufunc.pd
reductions) against pdlhip_run of the same op, through the C ABI: two calls
in a row, the second on a smaller input, each bit-identical to the
allocating entry point; long rows take the split path whose partials live
in the library's reserved scratch.  Also the refusals: a null output, an
input of another type."""
import ctypes

import numpy as np
import pytest

from pdl_amd import _lib as L
from pdl_amd import core as C
from oracle import pdlcpu as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    C.init(0)
    yield


def noalloc(op, a, b):
    ca, cb = a._to_c(), b._to_c()
    L.check(L.lib().pdlhip_run_reduce_noalloc(op.encode(), ctypes.pointer(ca), ctypes.pointer(cb)))
    b._from_c(cb, False)


CASES = [  # (op, numpy dtype, first shape, second (smaller) shape)
    ("sumover", np.float32, (3, 2_000_003), (2, 300_000)),      # split path: chunk partials in the scratch
    ("average", np.float64, (5, 700_001), (1, 65_536)),
    ("minimum", np.float32, (2, 1_000_000), (7, 90_000)),
    ("maximum_ind", np.int32, (4, 400_000), (3, 70_000)),
    ("sumover", np.int16, (900, 3000), (30, 500)),             # one wave per row: no scratch at all
    ("dsumover", np.uint8, (2, 5_000_000), (1, 100_000)),
]


@pytest.mark.parametrize("op,dt,s1,s2", CASES, ids=["%s-%s" % (c[0], np.dtype(c[1]).name) for c in CASES])
def test_noalloc_equals_the_allocating_entry_point(op, dt, s1, s2):
    rng = np.random.default_rng(2027)
    for shape in (s1, s2):
        if np.dtype(dt).kind == "f":
            x = rng.standard_normal(shape).astype(dt)
            x.reshape(-1)[::97] = np.nan if op in ("minimum", "maximum_ind") else x.reshape(-1)[::97]
        else:
            info = np.iinfo(dt)
            x = rng.integers(max(info.min, -1000), min(info.max, 1000), shape, endpoint=True).astype(dt)
        a = C.ndarray.from_numpy(x)
        ref = C.ndarray()
        C._run(op, [a, ref])                          # the allocating entry point creates the output
        out = C.ndarray.new_dense(ref.dtype, ref.dims)
        noalloc(op, a, out)
        got, want = out.numpy(), ref.numpy()
        assert got.dtype == want.dtype and got.shape == want.shape, (op, shape)
        assert got.tobytes() == want.tobytes(), (op, shape)
        assert out.badflag() == ref.badflag()


def test_noalloc_refusals():
    a = C.ndarray.from_numpy(np.arange(10, dtype=np.int16))
    with pytest.raises(L.PDLError, match="output must exist"):
        noalloc("sumover", a, C.ndarray())
    with pytest.raises(L.PDLError, match="no converted copy"):
        noalloc("sumover", a, C.ndarray.new_dense(O.S, []))      # sumover of short is int+ -> long
    with pytest.raises(L.PDLError, match="is not one of"):
        noalloc("cumusumover", a, C.ndarray.new_dense(O.L, [10]))
