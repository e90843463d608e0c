"""Ragged lists on the device (pdlhip_ragged_*): a list of 1-D ndarrays of
uneven lengths evaluated in a fixed set of launches, with a chain of Ops.pd
binops fused per element, then a Ufunc reduction or integer scan per item.

The reference has no such entry point: the same result comes from a Perl loop
over the list, one broadcast loop per op per ndarray.  Each test therefore
runs that loop in the oracle (eager binops, then the reduction) and holds the
fused result to the same bars as the single-ndarray ops: bit-exact for
integer scans, for min/max of exactly rounded chains and for bad-value
flags; float sums within the compensated-sum bound against a long-double
truth of the oracle's chain values.  A HIP graph replay must equal an eager
run bit for bit."""
import numpy as np
import pytest

from pdl_amd import core as C
from oracle import pdlcpu as O

pytestmark = pytest.mark.gpu

NAMES = O.TYPE_NAMES
# lengths: empty, single, one below / at / above a 4096-element segment,
# several segments, then a lognormal spread
EDGE_LENS = [0, 1, 2, 63, 64, 65, 4095, 4096, 4097, 8192, 12289, 30000]


@pytest.fixture(scope="module", autouse=True)
def device():
    C.init(0)
    yield


def lengths(rng, n, mean=6.0, sigma=1.5, cap=30000):
    ln = np.clip(np.round(rng.lognormal(mean, sigma, n)), 1, cap).astype(np.int64)
    return list(EDGE_LENS) + [int(x) for x in ln]


def eager_chain(run, slots, chain, null):
    """evaluate the chain as PDL would, one op per ndarray: run(op, x, y, out)"""
    vals = []
    for op, x, y in chain:
        a = vals[x[1]] if isinstance(x, tuple) else slots[x]
        b = vals[y[1]] if isinstance(y, tuple) else slots[y]
        out = null()
        run(op, a, b, out)
        vals.append(out)
    return vals[-1] if vals else slots[0]


def oracle_run(op, a, b, out):
    O.run(op, a, b, out)


def device_run(op, a, b, out):
    C._run(op, [a, b, out])


# (a-b)**2 * w * m : a, b, w of type t, 2 a Perl integer scalar, m a byte mask
CHAIN_POW = [("minus", 0, 1), ("power", ("node", 0), 2), ("mult", ("node", 1), 3), ("mult", ("node", 2), 4)]
# (a-b) * w * m : every step exactly rounded, so values match the oracle bit for bit
CHAIN_EXACT = [("minus", 0, 1), ("mult", ("node", 0), 3), ("mult", ("node", 1), 4)]


def make_items(rng, t, lens, packed):
    """per item: oracle slots and device slots [a, b, 2, w, m]; `packed` puts
    every a and b in one device buffer (views into it), as a caller holding a
    large ragged set would"""
    npt = O.NP_TYPES[t]
    total = int(sum(lens))
    if packed:
        big_a = C.ndarray.from_numpy(np.zeros(max(total, 1), npt), t)
        big_b = C.ndarray.from_numpy(np.zeros(max(total, 1), npt), t)
    two_o = O.OPdl.from_numpy(np.array(2, np.int8), O.SB)
    two_d = C.topdl(2)
    items_o, items_d = [], []
    off = 0
    for n in lens:
        a = rng.standard_normal(n).astype(npt)
        b = rng.standard_normal(n).astype(npt)
        w = rng.random(n).astype(npt)
        m = (rng.random(n) < 0.8).astype(np.uint8)
        so = [O.OPdl.from_numpy(a, t), O.OPdl.from_numpy(b, t), two_o, O.OPdl.from_numpy(w, t),
              O.OPdl.from_numpy(m, O.B)]
        if packed:
            da = C.view_into(big_a, [n], off)
            db = C.view_into(big_b, [n], off)
            if n:
                C._run("assgn", [C.ndarray.from_numpy(a, t), da])
                C._run("assgn", [C.ndarray.from_numpy(b, t), db])
            off += n
        else:
            da, db = C.ndarray.from_numpy(a, t), C.ndarray.from_numpy(b, t)
        sd = [da, db, two_d, C.ndarray.from_numpy(w, t), C.ndarray.from_numpy(m, O.B)]
        items_o.append(so)
        items_d.append(sd)
    return items_o, items_d


# a power node (x ** 2 is the correctly rounded x*x; other exponents the
# device's powf/pow) is within 2 ulp of glibc's value per element (the
# elementwise binop tests check that bound); these relative bounds cover
# 2-3 ulp of the chain's type
POW_TOL = {O.F: 4e-7, O.D: 4e-15}


def sum_bound(truth, absum, out_t, chain_t=None):
    """compensated sum: 2 eps |truth| + 64 eps^2 sum|v| in the output type,
    plus the per-element power tolerance of a chain computed in chain_t"""
    eps = 2.0 ** -23 if out_t == O.F else 2.0 ** -52
    el = POW_TOL[chain_t] * absum if chain_t is not None else 0.0
    return 2 * eps * np.abs(truth) + 64 * eps * eps * absum + el + 1e-300


@pytest.mark.parametrize("t", [O.F, O.D], ids=["float", "double"])
@pytest.mark.parametrize("final", ["sumover", "dsumover", "average", "daverage"])
def test_fused_chain_reductions_match_the_per_ndarray_loop(t, final):
    rng = np.random.default_rng(100 + t)
    lens = lengths(rng, 200)
    items_o, items_d = make_items(rng, t, lens, packed=(final in ("sumover", "average")))
    r = C.Ragged(items_d, CHAIN_POW, final)
    r.run()
    got = [o.numpy().astype(np.float64) for o in r.outputs]
    out_t = O.D if final.startswith("d") else t
    assert all(o.dtype == out_t and o.dims == [] for o in r.outputs)
    for i, (so, sd) in enumerate(zip(items_o, items_d)):
        v = eager_chain(oracle_run, so, CHAIN_POW, O.OPdl.null).gather()
        x = v.astype(np.longdouble)
        s = x.sum()
        absum = np.abs(x).sum()
        if final in ("average", "daverage"):
            s = s / len(v) if len(v) else np.nan
        truth = np.float64(s)
        bound = sum_bound(truth, float(absum) / max(len(v), 1) if final.endswith("average") else float(absum),
                          out_t, t)
        ok = (np.isnan(truth) and np.isnan(got[i])) or abs(got[i] - truth) <= bound
        assert ok, (final, NAMES[t], i, lens[i], truth, got[i])
        if i % 37 == 0:
            # the unfused device path: one launch per op, as PDL would run it
            z = eager_chain(device_run, sd, CHAIN_POW, C.ndarray)
            u = C.ndarray()
            C._run(final, [z, u])
            uv = float(u.numpy())
            assert (np.isnan(truth) and np.isnan(uv)) or abs(uv - truth) <= bound
    r.free()


@pytest.mark.parametrize("t", [O.F, O.D], ids=["float", "double"])
@pytest.mark.parametrize("final", ["minimum", "maximum"])
def test_fused_chain_minmax_bit_exact(t, final):
    rng = np.random.default_rng(200 + t)
    lens = lengths(rng, 150)
    items_o, items_d = make_items(rng, t, lens, packed=True)
    # an all-NaN item and an item with a leading NaN
    npt = O.NP_TYPES[t]
    for k, pat in ((3, "all"), (5, "lead")):
        n = lens[k]
        a = np.full(n, np.nan, npt) if pat == "all" else rng.standard_normal(n).astype(npt)
        if pat == "lead" and n:
            a[0] = np.nan
        items_o[k][0] = O.OPdl.from_numpy(a, t)
        items_d[k][0] = C.ndarray.from_numpy(a, t)
    r = C.Ragged(items_d, CHAIN_EXACT, final)
    r.run()
    flags = r.badflags()
    for i, so in enumerate(items_o):
        v = eager_chain(oracle_run, so, CHAIN_EXACT, O.OPdl.null)
        out = O.OPdl.null()
        O.run(final, v, out)
        a, b = out.gather(), r.outputs[i].numpy()
        assert a.dtype == b.dtype
        assert (np.isnan(a) and np.isnan(b)) or a.tobytes() == b.tobytes(), \
            (final, i, lens[i], a, b)
        assert flags[i] == bool(out.state & O.BADVAL), (final, i, lens[i])
        assert r.outputs[i].badflag() == flags[i]
    r.free()


SCAN_T = [O.SB, O.B, O.S, O.US, O.L, O.UL, O.IND, O.ULL, O.LL]


@pytest.mark.parametrize("t", SCAN_T, ids=[NAMES[t] for t in SCAN_T])
@pytest.mark.parametrize("final", ["cumusumover", "cumuprodover"])
def test_ragged_integer_scans_bit_exact(t, final):
    rng = np.random.default_rng(300 + t)
    lens = lengths(rng, 60, cap=20000)
    dt = np.dtype(O.NP_TYPES[t])
    info = np.iinfo(dt)
    items_o, items_d, strided = [], [], []
    for i, n in enumerate(lens):
        stride = 2 if i % 3 == 1 else 1
        lo, hi = (info.min, info.max) if final == "cumusumover" else (max(info.min, -3), min(info.max, 3))
        base = rng.integers(lo, hi, size=max(n * stride, 1), endpoint=True, dtype=dt)
        if stride == 1:
            base = base[:n]
            items_o.append([O.OPdl.from_numpy(base, t)])
            items_d.append([C.ndarray.from_numpy(base, t)])
        else:
            so = O.OPdl.from_numpy(base, t).slice((0, 2 * n - 1, 2)) if n else O.OPdl.from_numpy(base[:0], t)
            sd = C.ndarray.from_numpy(base, t).slice((0, 2 * n - 1, 2)) if n else C.ndarray.from_numpy(base[:0], t)
            items_o.append([so])
            items_d.append([sd])
        strided.append(stride)
    r = C.Ragged(items_d, [], final)
    r.run()
    for i, so in enumerate(items_o):
        out = O.OPdl.null()
        O.run(final, so[0], out)
        d = r.outputs[i]
        assert NAMES[d.dtype] == NAMES[out.dtype] and list(d.dims) == list(out.dims), (i, lens[i])
        np.testing.assert_array_equal(out.gather(), d.numpy(), err_msg="%s %s item %d len %d stride %d" %
                                      (final, NAMES[t], i, lens[i], strided[i]))
    r.free()


@pytest.mark.parametrize("t,final", [(O.LL, "cumusumover"), (O.B, "cumusumover"), (O.L, "cumuprodover")])
def test_ragged_scans_long_items_span_many_tiles(t, final):
    """items of up to 300000 elements (74 chunks of 4096 carried through one
    block) between empty, one-element and chunk-edge items; unit-stride, misaligned (slice from 1) and strided (step 3)
    layouts; wrapping results bit-exact with the oracle"""
    rng = np.random.default_rng(900 + t)
    lens = [300000, 0, 1, 4096, 4097, 270001, 8191, 12288, 5, 266240, 300000, 3]
    dt = np.dtype(O.NP_TYPES[t])
    info = np.iinfo(dt)
    lo, hi = (info.min, info.max) if final == "cumusumover" else (-3, 3)
    items_o, items_d, how = [], [], []
    for i, n in enumerate(lens):
        kind = ("unit", "misaligned", "strided")[i % 3]
        if kind == "unit":
            base = rng.integers(lo, hi, size=n, endpoint=True, dtype=dt)
            so, sd = O.OPdl.from_numpy(base, t), C.ndarray.from_numpy(base, t)
        else:
            step, start = (1, 1) if kind == "misaligned" else (3, 0)
            base = rng.integers(lo, hi, size=start + step * max(n, 1), endpoint=True, dtype=dt)
            sl = (start, start + step * (n - 1), step)
            so = O.OPdl.from_numpy(base, t).slice(sl) if n else O.OPdl.from_numpy(base[:0], t)
            sd = C.ndarray.from_numpy(base, t).slice(sl) if n else C.ndarray.from_numpy(base[:0], t)
        items_o.append([so])
        items_d.append([sd])
        how.append(kind)
    r = C.Ragged(items_d, [], final)
    for rep in range(2):
        r.run()
        for i, so in enumerate(items_o):
            out = O.OPdl.null()
            O.run(final, so[0], out)
            np.testing.assert_array_equal(out.gather(), r.outputs[i].numpy(), err_msg="%s %s item %d len %d %s rep %d" %
                                          (final, NAMES[t], i, lens[i], how[i], rep))
    r.free()


def test_graph_replays_equal_eager_runs():
    """a fused ragged sumover and an int64 ragged cumusumover captured in one
    HIP graph; the inputs change between replays (eager device ops and
    synchronisations in between), and every replay equals an eager run"""
    rng = np.random.default_rng(7)
    lens = lengths(rng, 300)
    items_o, items_d = make_items(rng, O.F, lens, packed=True)
    ints = [C.ndarray.from_numpy(rng.integers(-2**62, 2**62, size=n, dtype=np.int64), O.LL) for n in lens]
    rs = C.Ragged(items_d, CHAIN_POW, "sumover")
    rc = C.Ragged([[x] for x in ints], [], "cumusumover")

    def both():
        rs.run()
        rc.run()

    g = C.Graph(both)
    for rep in range(4):
        # change the inputs in place on the device
        for k in (0, 17, 101, 250):
            C._run("plus", [items_d[k][0], C.topdl(0.25 * (rep + 1)), items_d[k][0]])
            C._run("mult", [ints[k], C.topdl(3), ints[k]])
        C.synchronize()
        g.launch()
        C.synchronize()
        s_graph = [o.numpy().copy() for o in rs.outputs]
        c_graph = [o.numpy().copy() for o in rc.outputs]
        both()
        C.synchronize()
        for i in range(len(lens)):
            assert s_graph[i].tobytes() == rs.outputs[i].numpy().tobytes() or (
                np.isnan(s_graph[i]) and np.isnan(rs.outputs[i].numpy())), (rep, i)
            np.testing.assert_array_equal(c_graph[i], rc.outputs[i].numpy(), err_msg="rep %d item %d" % (rep, i))
    # the last replay's int64 scans against numpy (wrapping adds)
    for i in range(0, len(lens), 7):
        np.testing.assert_array_equal(np.cumsum(ints[i].numpy(), dtype=np.int64), c_graph[i])
    g.destroy()
    rs.free()
    rc.free()


# ------------------------------------------------------------------ dense chains
def test_chain_c_equals_a_times_b_plus_a_with_sumover_in_one_pass():
    """c = a*b + a and s = sumover(c) (double, 2048 x 300): c bit for bit as
    PDL's two eager ops give it, s within the compensated-sum bound"""
    rng = np.random.default_rng(51)
    a = rng.uniform(-1, 1, (300, 2048))
    b = rng.uniform(-1, 1, (300, 2048))
    ch = C.Chain([C.ndarray.from_numpy(a, O.D), C.ndarray.from_numpy(b, O.D)],
                 [("mult", 0, 1), ("plus", ("node", 0), 0)], "sumover", value=True)
    ch.run()
    oa, ob = O.OPdl.from_numpy(a, O.D), O.OPdl.from_numpy(b, O.D)
    oc = eager_chain(oracle_run, [oa, ob], [("mult", 0, 1), ("plus", ("node", 0), 0)], O.OPdl.null)
    assert ch.value.dims == [2048, 300] and ch.result.dims == [300]
    assert ch.value.numpy().tobytes() == oc.gather().tobytes()
    x = oc.gather().astype(np.longdouble)
    truth = x.sum(axis=-1).astype(np.float64)
    bound = sum_bound(truth, np.abs(x).sum(axis=-1).astype(np.float64), O.D)
    assert np.all(np.abs(ch.result.numpy() - truth) <= bound)


@pytest.mark.parametrize("final", ["minimum", "maximum", "none"])
def test_chain_broadcasts_like_the_eager_ops(final):
    """a (PDL dims 64,50,3, an xchg view) - b (64,1,3: a size-1 dim) times a
    0-dim Perl integer 2: values and min/max bit for bit with the oracle's
    eager ops"""
    rng = np.random.default_rng(52)
    a = rng.standard_normal((3, 64, 50)).astype(np.float32)      # PDL dims (50,64,3) -> xchg(0,1) -> (64,50,3)
    b = rng.standard_normal((3, 1, 64)).astype(np.float32)       # PDL dims (64,1,3)
    a[1, 5, :] = np.nan                                          # a NaN column of the view
    w = np.array(2, np.int8)
    chain = [("minus", 0, 1), ("mult", ("node", 0), 2)]
    da = C.ndarray.from_numpy(a, O.F).xchg(0, 1)
    db = C.ndarray.from_numpy(b, O.F)
    dw = C.topdl(2)
    ch = C.Chain([da, db, dw], chain, final, value=True)
    ch.run()
    flag = ch.badflags()
    oa = O.OPdl.from_numpy(a, O.F).xchg(0, 1)
    ob = O.OPdl.from_numpy(b, O.F)
    ow = O.OPdl.from_numpy(w, O.SB)
    ov = eager_chain(oracle_run, [oa, ob, ow], chain, O.OPdl.null)
    assert ch.value.dims == list(ov.dims) and ch.value.dtype == ov.dtype
    np.testing.assert_array_equal(ch.value.numpy(), ov.gather())
    if final != "none":
        out = O.OPdl.null()
        O.run(final, ov, out)
        np.testing.assert_array_equal(ch.result.numpy(), out.gather())
        assert flag == bool(out.state & O.BADVAL)


def test_chain_integer_scan_rows():
    rng = np.random.default_rng(53)
    x = rng.integers(-2**31, 2**31 - 1, size=(37, 5000), dtype=np.int32)
    ch = C.Chain([C.ndarray.from_numpy(x, O.L)], [], "cumusumover")
    ch.run()
    out = O.OPdl.null()
    O.run("cumusumover", O.OPdl.from_numpy(x, O.L), out)
    np.testing.assert_array_equal(ch.result.numpy(), out.gather())


def test_chain_graph_replays_equal_eager_runs():
    rng = np.random.default_rng(54)
    a = C.ndarray.from_numpy(rng.uniform(-1, 1, (64, 3000)), O.D)
    b = C.ndarray.from_numpy(rng.uniform(-1, 1, (64, 3000)), O.D)
    ch = C.Chain([a, b], [("mult", 0, 1), ("plus", ("node", 0), 0)], "sumover", value=True)
    g = C.Graph(ch.run)
    for rep in range(4):
        C._run("plus", [a, C.topdl(0.5), a])
        C.synchronize()
        g.launch()
        C.synchronize()
        v, s = ch.value.numpy().copy(), ch.result.numpy().copy()
        ch.run()
        C.synchronize()
        assert v.tobytes() == ch.value.numpy().tobytes(), rep
        assert s.tobytes() == ch.result.numpy().tobytes(), rep
    g.destroy()


def test_general_chain_program_matches_eager_ops():
    """a program that is not left-deep ((a-b) + (a*w), then times the mask)
    runs in the general per-element kernel: values bit for bit with the
    oracle's eager ops, minimum bit-exact, sums within the bound"""
    rng = np.random.default_rng(55)
    lens = lengths(rng, 40)
    items_o, items_d = make_items(rng, O.F, lens, packed=False)
    chain = [("minus", 0, 1), ("mult", 0, 3), ("plus", ("node", 0), ("node", 1)), ("mult", ("node", 2), 4)]
    rv = C.Ragged(items_d, chain, "minimum", values=True)
    rv.run()
    rs = C.Ragged(items_d, chain, "sumover")
    rs.run()
    for i, so in enumerate(items_o):
        v = eager_chain(oracle_run, so, chain, O.OPdl.null)
        np.testing.assert_array_equal(rv.values[i].numpy(), v.gather(), err_msg="item %d" % i)
        out = O.OPdl.null()
        O.run("minimum", v, out)
        assert out.gather().tobytes() == rv.outputs[i].numpy().tobytes(), i
        x = v.gather().astype(np.longdouble)
        truth = float(x.sum())
        assert abs(float(rs.outputs[i].numpy()) - truth) <= sum_bound(truth, float(np.abs(x).sum()), O.F), i


# ------------------------------------------------------------------ mixed-type chains
MIXED = [
    # (slot types, chain): each node computes in PDL's trans type of its op
    ([O.L, O.L, O.F], [("mult", 0, 1), ("plus", ("node", 0), 2)]),                # long (wraps) -> float
    ([O.S, O.US, O.F, O.D], [("minus", 0, 1), ("mult", ("node", 0), 2), ("divide", ("node", 1), 3)]),
    ([O.F, O.F, O.D], [("or2", 0, 1), ("plus", ("node", 0), 2)]),                 # floats -> longlong -> double
    ([O.L, O.L, O.D], [("gt", 0, 1), ("mult", ("node", 0), 2)]),                  # a comparison in long, then double
    ([O.B, O.B, O.F, O.F], [("plus", 0, 1), ("minus", 2, 3), ("mult", ("node", 0), ("node", 1))]),  # byte wraps
]


def mixed_slots(rng, types, n):
    out_o, out_d = [], []
    for t in types:
        dt = O.NP_TYPES[t]
        if np.dtype(dt).kind == "f":
            x = (rng.standard_normal(n) * 1000).astype(dt)
        else:
            info = np.iinfo(dt)
            x = rng.integers(max(info.min, -2 ** 40), min(info.max, 2 ** 40), n, endpoint=True).astype(dt)
        out_o.append(O.OPdl.from_numpy(x, t))
        out_d.append(C.ndarray.from_numpy(x, t))
    return out_o, out_d


@pytest.mark.parametrize("k", range(len(MIXED)))
def test_mixed_type_chains_match_the_eager_ops(k):
    """every node in its own trans type with its operands converted
    (pdlconv.c:163-174): the chain's values bit for bit with the oracle's
    eager ops (integer wraps, float-to-integer truncation, comparisons),
    minimum bit-exact, sums within the compensated-sum bound"""
    types, chain = MIXED[k]
    rng = np.random.default_rng(60 + k)
    lens = lengths(rng, 30)
    items = [mixed_slots(rng, types, n) for n in lens]
    rv = C.Ragged([d for _, d in items], chain, "minimum", values=True)
    rv.run()
    rs = C.Ragged([d for _, d in items], chain, "sumover")
    rs.run()
    for i, (so, _) in enumerate(items):
        v = eager_chain(oracle_run, so, chain, O.OPdl.null)
        assert rv.values[i].dtype == v.dtype
        np.testing.assert_array_equal(rv.values[i].numpy().view(np.uint8), v.gather().view(np.uint8),
                                      err_msg="item %d" % i)
        out = O.OPdl.null()
        O.run("minimum", v, out)
        assert out.gather().tobytes() == rv.outputs[i].numpy().tobytes(), i
        x = v.gather().astype(np.longdouble)
        truth = float(x.sum())
        assert abs(float(rs.outputs[i].numpy()) - truth) <= sum_bound(truth, float(np.abs(x).sum()), v.dtype), i


def test_mixed_type_dense_chain_broadcasts_like_the_eager_ops():
    """short * short (wrapping in short), plus a broadcast float row, over
    (300, 257): values bit for bit, maximum per row bit-exact"""
    rng = np.random.default_rng(70)
    a = rng.integers(-30000, 30000, (300, 257)).astype(np.int16)
    b = rng.integers(-30000, 30000, (300, 257)).astype(np.int16)
    c = rng.standard_normal((1, 257)).astype(np.float32)
    chain = [("mult", 0, 1), ("plus", ("node", 0), 2)]
    ch = C.Chain([C.ndarray.from_numpy(a, O.S), C.ndarray.from_numpy(b, O.S), C.ndarray.from_numpy(c, O.F)],
                 chain, "maximum", value=True)
    ch.run()
    so = [O.OPdl.from_numpy(a, O.S), O.OPdl.from_numpy(b, O.S), O.OPdl.from_numpy(c, O.F)]
    ov = eager_chain(oracle_run, so, chain, O.OPdl.null)
    assert ch.value.dtype == ov.dtype == O.F
    np.testing.assert_array_equal(ch.value.numpy(), ov.gather())
    out = O.OPdl.null()
    O.run("maximum", ov, out)
    np.testing.assert_array_equal(ch.result.numpy(), out.gather())
