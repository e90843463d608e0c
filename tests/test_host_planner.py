"""Host logic without a GPU: the C ABI loads and exports every symbol
include/pdlhip.h declares, and its planner (dry-run mode) agrees with the
oracle's restatement of PDL's type selection and broadcast engine on output
types, dims, bad flags and error texts."""
import os
import re
import subprocess

import numpy as np
import pytest

import pdl_amd
from pdl_amd import _lib as L
from pdl_amd import core as C
from oracle import pdlcpu as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pdlhip.h")).read()
    return sorted(set(re.findall(r"\b(pdlhip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIBPATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (pdlhip_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    assert len(declared_symbols()) >= 50


def test_init_without_gpu_fails_loudly_or_succeeds():
    e = L.lib().pdlhip_abi_version()
    assert e == 2


@pytest.fixture
def plan_only():
    C.set_plan_only(True)
    yield
    C.set_plan_only(False)


RNG = np.random.default_rng(7)
OPS3 = ["plus", "mult", "minus", "divide", "gt", "eq", "shiftleft", "or2", "power", "atan2", "modulo",
        "spaceship"]
OPS2 = ["sumover", "prodover", "dsumover", "cumusumover", "average", "daverage", "minimum", "maximum_ind",
        "assgn"]


def rand_dims():
    nd = RNG.integers(0, 4)
    return [int(RNG.choice([1, 1, 2, 3, 4])) for _ in range(nd)]


def both(op, specs, swap=0, m_size=None):
    """run `op` on matching oracle / device headers; return (result, error)"""
    res = []
    for backend in ("oracle", "device"):
        objs = []
        for sp in specs:
            if sp is None:
                objs.append(O.OPdl.null() if backend == "oracle" else C.ndarray())
                continue
            t, dims, state, view = sp
            if backend == "oracle":
                p = O.OPdl(np.ones(max(int(np.prod(dims)) if dims else 1, 1), O.NP_TYPES[t]), t, dims,
                           state=state)
            else:
                p = C.shell(t, dims, state)
            for v in view:
                p = getattr(p, v[0])(*v[1:])
            objs.append(p)
        try:
            if backend == "oracle":
                O.run(op, *objs, swap=swap, m_size=m_size)
            else:
                C._run(op, objs, swap=swap, m_size=m_size)
            res.append(("ok", [(p.dims, p.dtype, bool(p.state & 2)) for p in objs]))
        except (O.OracleError, pdl_amd.PDLError) as e:
            res.append(("err", str(e)))
    return res


def test_planner_matches_oracle_random(plan_only):
    types = [O.SB, O.B, O.S, O.US, O.L, O.UL, O.IND, O.ULL, O.LL, O.F, O.D]
    n = 0
    for trial in range(3000):
        op = RNG.choice(OPS3 + OPS2 + ["matmult"])
        nin = 2 if op in OPS3 or op == "matmult" else 1
        specs = []
        for k in range(nin):
            dims = rand_dims()
            if op == "matmult":
                dims = [int(RNG.integers(1, 4)), int(RNG.integers(1, 4))] + dims[:1]
            view = []
            if len(dims) >= 2 and RNG.random() < 0.3:
                view.append(("xchg", 0, 1))
            if RNG.random() < 0.2:
                view.append(("dummy", int(RNG.integers(0, len(dims) + 1)), int(RNG.choice([1, 2, 3]))))
            specs.append((int(RNG.choice(types)), dims, 2 if RNG.random() < 0.2 else 0, view))
        nout = 1
        out = None
        if RNG.random() < 0.3:
            dims = rand_dims()
            out = (int(RNG.choice(types)), dims, 0, [])
        specs += [out] * nout
        swap = int(RNG.random() < 0.3) if op in OPS3 else 0
        o, d = both(op, specs, swap=swap)
        assert o == d, (op, specs, swap, o, d)
        n += o[0] == "ok"
    assert n > 300


def test_planner_matches_oracle_explicit_ids(plan_only):
    """operands (outputs included) made with broadcastI over random dims:
    the explicit-id pass (pdlbroadcast.c:437-467), the last operand's
    counts, the mismatch and create-while-broadcasting errors"""
    types = [O.S, O.L, O.F, O.D]
    n_ok = n_err = 0
    for trial in range(3000):
        op = RNG.choice(OPS3[:6] + ["sumover", "minimum", "assgn", "matmult"])
        nin = 2 if op in OPS3 or op == "matmult" else 1
        nout_given = RNG.random() < 0.85
        specs = []
        for k in range(nin + 1):
            if k == nin and not nout_given:
                specs.append(None)
                continue
            dims = [int(RNG.choice([1, 2, 3])) for _ in range(int(RNG.integers(1, 5)))]
            view = []
            if RNG.random() < 0.7:
                nd = len(dims)
                wd = [int(x) for x in RNG.permutation(nd)[:int(RNG.integers(1, nd + 1))]]
                view.append(("broadcastI", int(RNG.choice([1, 1, 2])), wd))
            specs.append((int(RNG.choice(types)), dims, 0, view))
        o, d = both(op, specs)
        assert o == d, (op, specs, o, d)
        if o[0] == "ok":
            n_ok += 1
        else:
            n_err += 1
    assert n_ok > 200 and n_err > 200, (n_ok, n_err)


def test_planner_golden_errors(plan_only):
    """error rows of the reference's dim-compatibility table, on this
    engine's own signatures"""
    F64 = O.D
    cases = [
        ("plus", [(F64, [1, 3], 0, []), (F64, [1], 0, []), (F64, [2, 2], 0, [])], "Mismatch"),
        ("plus", [(F64, [2, 2], 0, []), (F64, [2, 2], 0, []), (F64, [2], 0, [])], "implicit dim"),
        ("plus", [(F64, [2, 2], 0, []), (F64, [2, 2], 0, []), (F64, [2], 0, [("dummy", 1, 2)])], "dim is dummy"),
        ("sumover", [(F64, [3, 2], 0, []), (F64, [2], 0, [("dummy", 0, 2)])], "implicit dim"),
        ("cumusumover", [(F64, [3, 2], 0, []), (O.L, [2, 2], 0, [])], "dim has size 2"),
        ("cumusumover", [(F64, [3, 2], 0, []), (F64, [2], 0, [("dummy", 0, 3)])], "over dummy dim"),
        ("matmult", [(F64, [3, 2], 0, []), (F64, [2, 4], 0, []), None], "dim has size 4"),
    ]
    for op, specs, pat in cases:
        o, d = both(op, specs)
        assert o == d, (op, o, d)
        assert o[0] == "err" and re.search(pat, o[1]), o


def test_n_ind_m_size_check(plan_only):
    o, d = both("minimum_n_ind", [(O.D, [3], 0, []), None], m_size=5)
    assert o == d and "m_size > n_size" in o[1]
