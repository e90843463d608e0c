"""Pins the oracle where the reference's own vectors leave it open:
numpy closed forms (wrapping integer arithmetic, IEEE single/double ops,
sequential float accumulation via cumsum), brute force on small inputs, and
the reference's own RNG source compiled into oracle/_ref."""
import ctypes
import os

import numpy as np
import pytest

from oracle import pdlcpu as O

RNG = np.random.default_rng(1234)
INT_TYPES = [O.SB, O.B, O.S, O.US, O.L, O.UL, O.IND, O.ULL, O.LL]
ALL_TYPES = INT_TYPES + [O.F, O.D]


def rand_arr(t, shape, lo=None, hi=None):
    dt = np.dtype(O.NP_TYPES[t])
    if dt.kind in "iu":
        info = np.iinfo(dt)
        lo = info.min if lo is None else lo
        hi = info.max if hi is None else hi
        return RNG.integers(lo, hi, size=shape, endpoint=True, dtype=dt)
    return RNG.standard_normal(shape).astype(dt)


def run1(op, a, b, t, swap=0):
    pa, pb, pc = O.OPdl.from_numpy(a, t), O.OPdl.from_numpy(b, t), O.OPdl.null()
    O.run(op, pa, pb, pc, swap=swap)
    assert pc.dtype == t
    return pc.numpy()


@pytest.mark.parametrize("t", ALL_TYPES, ids=[O.TYPE_NAMES[t] for t in ALL_TYPES])
@pytest.mark.parametrize("op", ["plus", "minus", "mult"])
def test_arith_wraps_like_numpy(op, t):
    a, b = rand_arr(t, (7, 13)), rand_arr(t, (7, 13))
    with np.errstate(all="ignore"):
        want = {"plus": a + b, "minus": a - b, "mult": a * b}[op]
    got = run1(op, a, b, t)
    if want.dtype.kind == "f":
        np.testing.assert_array_equal(got.view(np.uint8), want.view(np.uint8))
    else:
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("t", ALL_TYPES, ids=[O.TYPE_NAMES[t] for t in ALL_TYPES])
def test_divide_truncates(t):
    a = rand_arr(t, (50,))
    b = rand_arr(t, (50,))
    if np.dtype(O.NP_TYPES[t]).kind in "iu":
        b[b == 0] = 1
        if np.dtype(O.NP_TYPES[t]).kind == "i":
            b[b == -1] = 3          # INT_MIN / -1 is undefined in C
        want = np.array([int(x) // int(y) if (int(x) >= 0) == (int(y) > 0) or int(x) % int(y) == 0
                         else -(abs(int(x)) // abs(int(y))) for x, y in zip(a, b)], dtype=object)
        want = np.array([(abs(int(x)) // abs(int(y))) * (1 if (int(x) >= 0) == (int(y) > 0) else -1)
                         for x, y in zip(a, b)], dtype=object).astype(np.int64).astype(a.dtype)
    else:
        want = a / b
    got = run1("divide", a, b, t)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("t", ALL_TYPES, ids=[O.TYPE_NAMES[t] for t in ALL_TYPES])
@pytest.mark.parametrize("op,fn", [("gt", np.greater), ("lt", np.less), ("le", np.less_equal),
                                   ("ge", np.greater_equal), ("eq", np.equal), ("ne", np.not_equal)])
def test_comparisons_give_0_1_in_trans_type(op, fn, t):
    a = rand_arr(t, (40,), 0, 3) if t in INT_TYPES else np.round(rand_arr(t, (40,)))
    b = rand_arr(t, (40,), 0, 3) if t in INT_TYPES else np.round(rand_arr(t, (40,)))
    np.testing.assert_array_equal(run1(op, a, b, t), fn(a, b).astype(a.dtype))


@pytest.mark.parametrize("t", INT_TYPES, ids=[O.TYPE_NAMES[t] for t in INT_TYPES])
def test_bitwise(t):
    a, b = rand_arr(t, (64,)), rand_arr(t, (64,))
    np.testing.assert_array_equal(run1("or2", a, b, t), a | b)
    np.testing.assert_array_equal(run1("and2", a, b, t), a & b)
    np.testing.assert_array_equal(run1("xor", a, b, t), a ^ b)
    bits = np.dtype(O.NP_TYPES[t]).itemsize * 8
    sh = rand_arr(t, (64,), 0, bits - 1)
    pos = np.abs(a.astype(np.int64)).astype(a.dtype) // 4 if a.dtype.kind == "i" else a
    np.testing.assert_array_equal(run1("shiftright", pos, sh, t), pos >> sh)


def test_modulo_divisor_sign_rule_matches_python():
    """Ops.pd MOD: result takes the divisor's sign, x % 0 == 0 -- Python's %
    on ints is the same rule."""
    a = RNG.integers(-1000, 1000, size=200).astype(np.int32)
    b = RNG.integers(-20, 20, size=200).astype(np.int32)
    want = np.array([0 if y == 0 else int(x) % int(y) for x, y in zip(a, b)], dtype=np.int32)
    np.testing.assert_array_equal(run1("modulo", a, b, O.L), want)


def test_spaceship():
    a = np.array([1, 2, 3, np.nan], dtype=np.float64)
    b = np.array([2, 2, 2, 1], dtype=np.float64)
    np.testing.assert_array_equal(run1("spaceship", a, b, O.D), [-1, 0, 1, 1])


# ---- reductions --------------------------------------------------------------
def reduce1(op, a, t):
    pa, pb = O.OPdl.from_numpy(a, t), O.OPdl.null()
    O.run(op, pa, pb)
    return pb


@pytest.mark.parametrize("t", INT_TYPES, ids=[O.TYPE_NAMES[t] for t in INT_TYPES])
def test_integer_sumover_prodover_wrap(t):
    a = rand_arr(t, (5, 300))
    want_t = np.dtype(O.NP_TYPES[max(t, O.L)]) if t < O.F else a.dtype
    pb = reduce1("sumover", a, t)
    assert pb.dtype == max(t, O.L)
    np.testing.assert_array_equal(pb.numpy(), a.astype(want_t).sum(axis=1, dtype=want_t))
    small = rand_arr(t, (5, 40), 1, 3)
    np.testing.assert_array_equal(reduce1("prodover", small, t).numpy(),
                                  small.astype(want_t).prod(axis=1, dtype=want_t))
    np.testing.assert_array_equal(reduce1("cumusumover", a, t).numpy(),
                                  np.cumsum(a.astype(want_t), axis=1, dtype=want_t))


def test_float_sumover_is_sequential_float32():
    """sumover accumulates in float, in index order: cumsum's last element"""
    a = RNG.standard_normal((6, 5000)).astype(np.float32) * np.float32(1000)
    got = reduce1("sumover", a, O.F).numpy()
    want = np.cumsum(a, axis=1, dtype=np.float32)[:, -1]
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(reduce1("cumusumover", a, O.F).numpy(), np.cumsum(a, axis=1, dtype=np.float32))
    np.testing.assert_array_equal(reduce1("dsumover", a, O.F).numpy(),
                                  np.cumsum(a.astype(np.float64), axis=1)[:, -1])


def test_average_integer_truncates():
    a = np.array([[-7, 0], [7, 0], [-1, 2], [1, 0]], dtype=np.int32)
    np.testing.assert_array_equal(reduce1("average", a, O.L).numpy(), [-3, 3, 0, 0])
    f = RNG.standard_normal((3, 77)).astype(np.float32)
    want = np.cumsum(f, axis=1, dtype=np.float32)[:, -1] / np.float32(77)
    np.testing.assert_array_equal(reduce1("average", f, O.F).numpy(), want)


def test_prodover_stops_at_zero():
    a = np.array([[2, 0, np.inf, np.nan], [2, 3, np.inf, 4]], dtype=np.float64)
    np.testing.assert_array_equal(reduce1("prodover", a, O.D).numpy(), [0, np.inf])


def test_minmax_rules_brute_force():
    vals = np.array([np.nan, 1.0, -0.0, 0.0, 2.0, -1.0])
    for trial in range(300):
        a = RNG.choice(vals, size=RNG.integers(1, 7))
        # sequential scan restated in Python
        def scan(better):
            cur, ind = None, -1
            for i, v in enumerate(a):
                if ind != -1 and not better(v, cur) and not np.isnan(cur):
                    continue
                cur, ind = v, i
            return cur, ind
        for op, better in [("minimum", lambda v, c: v < c), ("maximum", lambda v, c: v > c)]:
            cur, ind = scan(better)
            got = reduce1(op, a, O.D).numpy()
            assert (np.isnan(got) and np.isnan(cur)) or (got == cur and np.signbit(got) == np.signbit(cur))
            assert reduce1(op + "_ind", a, O.D).numpy() == ind


def test_matmult_sequential_float32_order():
    h, t, w = 9, 70, 11
    a = RNG.standard_normal((h, t)).astype(np.float32)
    b = RNG.standard_normal((t, w)).astype(np.float32)
    pa, pb, pc = O.OPdl.from_numpy(a, O.F), O.OPdl.from_numpy(b, O.F), O.OPdl.null()
    O.run("matmult", pa, pb, pc)
    got = pc.numpy()
    prods = a[:, :, None] * b[None, :, :]          # (h, t, w) each product rounded to float32
    want = np.cumsum(prods, axis=1, dtype=np.float32)[:, -1, :]
    np.testing.assert_array_equal(got, want)
    # transposed operand views (Slices.pd xchg) give the same products
    pat = O.OPdl.from_numpy(np.ascontiguousarray(a.T), O.F).xchg(0, 1)
    pc2 = O.OPdl.null()
    O.run("matmult", pat, pb, pc2)
    np.testing.assert_array_equal(pc2.numpy(), want)


def test_matmult_int_wraps_brute_force():
    a = RNG.integers(-2**31, 2**31 - 1, size=(4, 6), dtype=np.int32)
    b = RNG.integers(-2**31, 2**31 - 1, size=(6, 5), dtype=np.int32)
    pa, pb, pc = O.OPdl.from_numpy(a, O.L), O.OPdl.from_numpy(b, O.L), O.OPdl.null()
    O.run("matmult", pa, pb, pc)
    want = (a.astype(np.int64) @ b.astype(np.int64)).astype(np.int32)  # wrapping is a ring hom.
    np.testing.assert_array_equal(pc.numpy(), want)


def test_broadcast_plan_transposed_dummy_view():
    """z = x->xchg(0,1) * y->dummy(1): x' incs (4096-like stride, 1, ...)"""
    x = RNG.standard_normal((3, 5, 4)).astype(np.float32)     # PDL dims (4,5,3)
    y = RNG.standard_normal((3, 5)).astype(np.float32)        # PDL dims (5,3)
    px = O.OPdl.from_numpy(x, O.F).xchg(0, 1)                 # dims (5,4,3)
    py = O.OPdl.from_numpy(y, O.F).dummy(1)                   # dims (5,1,3)
    pz = O.OPdl.null()
    O.run("mult", px, py, pz)
    assert pz.dims == [5, 4, 3]
    want = np.transpose(x, (0, 2, 1)) * y[:, None, :]
    np.testing.assert_array_equal(pz.numpy(), want)


def test_pthread_split_gives_same_results():
    a = RNG.standard_normal((64, 48)).astype(np.float64)
    b = RNG.standard_normal((64, 48)).astype(np.float64)
    outs = []
    for nthr in (0, 5):
        pa, pb, pc = O.OPdl.from_numpy(a), O.OPdl.from_numpy(b), O.OPdl.null()
        O.run("plus", pa, pb, pc, nthreads=nthr, size_mi=0)
        pr = O.OPdl.null()
        O.run("sumover", pc, pr, nthreads=nthr, size_mi=0)
        outs.append(pr.numpy())
    np.testing.assert_array_equal(outs[0], outs[1])
    np.testing.assert_array_equal(outs[0], np.cumsum(a + b, axis=1)[:, -1])


REF = os.path.join(os.path.dirname(O.HERE), "oracle", "_ref", "libxoshiro_ref.so")


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (needs the reference sources)")
def test_rng_matches_reference_source():
    ref = ctypes.CDLL(REF)
    ref.pdl_drand.restype = ctypes.c_double
    ref.pdl_drand.argtypes = [ctypes.c_void_p]
    ref.pdl_srand.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint64, ctypes.c_int]
    sp = ctypes.c_void_p(None)     # the reference reuses its state buffer across calls
    for seed in (0, 1, 42, 2**63 + 5):
        ref.pdl_srand(ctypes.byref(sp), seed, 1)
        want = [ref.pdl_drand(sp) for _ in range(1000)]
        got = O.random_uniform(1000, seed)
        np.testing.assert_array_equal(got, want)
