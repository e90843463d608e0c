"""Runs the golden cases (tests/golden/reference_vectors.json) on a backend.

A backend provides:  make(spec) -> ndarray, null(), run(op, pdls, swap, m_size),
read(p) -> (type_name, dims, flat numpy data, badflag), plus the views
xchg/mv/dummy/slice as ndarray methods.  Two backends exist: the oracle
(oracle/pdlcpu.py, CPU) and the device (pdl_amd, HIP through the C ABI).

PMCode restatements (Perl wrapper code the reference runs before the
compiled op):
  matmult  lib/PDL/Primitive.pd:197-209
  append   lib/PDL/Primitive.pd:1355-1368
"""
import json
import math
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TYPES = ["sbyte", "byte", "short", "ushort", "long", "ulong", "indx", "ulonglong", "longlong",
         "float", "double"]
NP = {"sbyte": np.int8, "byte": np.uint8, "short": np.int16, "ushort": np.uint16, "long": np.int32,
      "ulong": np.uint32, "indx": np.int64, "ulonglong": np.uint64, "longlong": np.int64,
      "float": np.float32, "double": np.float64}


def load_cases():
    with open(os.path.join(HERE, "golden", "reference_vectors.json")) as f:
        return json.load(f)


def spec_array(spec):
    t = spec["type"]
    data = spec["data"]
    if t in ("ulonglong", "longlong", "indx"):
        a = np.array([int(v) for v in data], dtype=object).astype(NP[t]) if data else np.zeros(0, NP[t])
    else:
        a = np.array(data, dtype=np.float64).astype(NP[t]) if data else np.zeros(0, NP[t])
    return a.reshape(spec["dims"][::-1]) if spec["dims"] else a.reshape(())


def apply_view(p, view):
    for v in view or []:
        if v[0] == "xchg":
            p = p.xchg(v[1], v[2])
        elif v[0] == "mv":
            p = p.mv(v[1], v[2])
        elif v[0] == "dummy":
            p = p.dummy(v[1], v[2])
        elif v[0] == "slice":
            p = p.slice(*v[1:])
        else:
            raise ValueError(v)
    return p


def build_args(B, case):
    objs = []
    for a in case["args"]:
        if a == "null":
            objs.append(B.null())
        elif isinstance(a, str) and a.startswith("inplace:"):
            objs.append(objs[int(a.split(":")[1])])
        else:
            p = B.make(a["type"], spec_array(a), badflag=bool(a.get("badflag")))
            p = apply_view(p, a.get("view"))
            objs.append(p)
    return objs


def pm_matmult(B, x, y, c):
    """lib/PDL/Primitive.pd:197-209"""
    while len(B.dims(x)) < 2:
        x = x.dummy(-1)
    while len(B.dims(y)) < 2:
        y = y.dummy(-1)
    xd, yd = B.dims(x), B.dims(y)
    if (xd[0] == 1 and xd[1] == 1) or (yd[0] == 1 and yd[1] == 1):
        t = B.null()
        B.run("mult", [x, y, t])
        B.run("assgn", [t, c])
        return c
    if yd[1] != xd[0]:
        raise B.Error("Dim mismatch in matmult of [%dx%d] x [%dx%d]: %d != %d"
                      % (xd[0], xd[1], yd[0], yd[1], xd[0], yd[1]))
    B.run("matmult", [x, y, c])
    return c


def pm_append(B, i1, i2, o):
    """lib/PDL/Primitive.pd:1355-1368"""
    nempty = sum(1 for p in (i1, i2) if B.nvals(p) == 0)
    if nempty == 2:
        raise NotImplementedError
    if nempty == 1:
        B.run("assgn", [i2 if B.nvals(i1) == 0 else i1, o])
        return o
    B.run("append", [i1, i2, o])
    return o


def run_case(B, case):
    objs = build_args(B, case)
    pm = case.get("pmcode")
    if pm == "matmult":
        objs[2] = pm_matmult(B, objs[0], objs[1], objs[2])
    elif pm == "append":
        objs[2] = pm_append(B, objs[0], objs[1], objs[2])
    else:
        B.run(case["op"], objs, swap=case.get("swap", 0), m_size=case.get("m_size"), dpar=case.get("dpar"))
    return objs


def check_case(B, case):
    """Runs one case; returns None on success or a description of the failure."""
    exp = case["expect"]
    if isinstance(exp, dict) and "error" in exp:
        try:
            run_case(B, case)
        except B.Error as e:
            if re.search(exp["error"], str(e)):
                return None
            return "wrong error: %s" % e
        return "expected error /%s/" % exp["error"]
    objs = run_case(B, case)
    exps = exp if isinstance(exp, list) else [exp]
    for e in exps:
        p = objs[e["arg"]]
        if case.get("check_root"):
            p = B.root_of(p)
        tname, dims, data, badflag = B.read(p)
        if tname != e["type"]:
            return "arg %d: type %s != %s" % (e["arg"], tname, e["type"])
        if list(dims) != list(e["dims"]):
            return "arg %d: dims %s != %s" % (e["arg"], dims, e["dims"])
        want = spec_array({"type": e["type"], "dims": [len(e["data"])], "data": e["data"]}).reshape(-1)
        got = np.asarray(data).reshape(-1)
        if want.dtype.kind == "f":
            tol = e.get("tol", 0.0)
            ok = (np.isnan(want) & np.isnan(got)) | (np.abs(got.astype(np.float64) - want) <= tol)
            if not ok.all():
                return "arg %d: data %s != %s" % (e["arg"], got, want)
        elif not np.array_equal(got, want):
            return "arg %d: data %s != %s" % (e["arg"], got, want)
        if "badflag" in e and bool(badflag) != bool(e["badflag"]):
            return "arg %d: badflag %s != %s" % (e["arg"], badflag, e["badflag"])
    return None
