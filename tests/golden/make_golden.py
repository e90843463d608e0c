"""Writes tests/golden/reference_vectors.json: known-answer vectors transcribed
from the reference's own test files (inline literals in t/*.t; PDL 2.106).

Each case names the .t file:line it comes from.  Inputs are given in PDL
layout (dims with dim 0 fastest, data flat in that order); "bad" lists flat
indices that hold the type's bad value.  Cases that exercise Perl-level
wrapper code (PMCode: matmult's dummy promotion / scalar shortcut / mismatch
message, append's empty handling) carry "pmcode": the harness restates that
wrapper logic.

Run:  python tests/golden/make_golden.py
"""
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def arr(t, dims, data, bad=None, badflag=None, view=None):
    d = {"type": t, "dims": list(dims), "data": list(data)}
    if bad:
        d["bad"] = list(bad)
    if badflag or bad:
        d["badflag"] = True
    if view:
        d["view"] = view
    return d


def scalar(t, v):
    return arr(t, [], [v])


def xvals(t, w, h):
    return arr(t, [w, h], [x for y in range(h) for x in range(w)])


def yvals(t, w, h):
    return arr(t, [w, h], [y for y in range(h) for x in range(w)])


def seq(t, *dims):
    n = 1
    for d in dims:
        n *= d
    return arr(t, dims, list(range(n)))


def zeroes(t, *dims):
    n = 1
    for d in dims:
        n *= d
    return arr(t, dims, [0] * n)


NULL = "null"
cases = []


def case(cid, src, op, args, expect, swap=0, **kw):
    c = {"id": cid, "src": src, "op": op, "args": args, "swap": swap, "expect": expect}
    c.update(kw)
    cases.append(c)


def out(arg, t, dims, data, bad=None, badflag=None, tol=0.0):
    e = {"arg": arg, "type": t, "dims": list(dims), "data": list(data)}
    if bad is not None:
        e["bad"] = list(bad)
    if badflag is not None:
        e["badflag"] = badflag
    if tol:
        e["tol"] = tol
    return e


# ---------------------------------------------------------------- t/ops.t
W, H = 3, 5
mx = [x - y for y in range(H) for x in range(W)]
case("ops_minus", "t/ops.t:16-30", "minus", [xvals("double", W, H), yvals("double", W, H), NULL],
     out(2, "double", [W, H], mx))
case("ops_minus_swap", "t/ops.t:16-30", "minus", [xvals("double", W, H), yvals("double", W, H), NULL],
     out(2, "double", [W, H], [-v for v in mx]), swap=1)
case("ops_power_defaults_double", "t/ops.t:33", "power", [scalar("sbyte", 10), scalar("sbyte", 2), NULL],
     out(2, "double", [], [100.0]))
case("ops_minus_scalar", "t/ops.t:37", "minus", [arr("double", [2], [5, 6]), scalar("sbyte", 1), NULL],
     out(2, "double", [2], [4, 5]))
case("ops_minus_scalar_swap", "t/ops.t:38", "minus", [arr("double", [2], [5, 6]), scalar("sbyte", 1), NULL],
     out(2, "double", [2], [-4, -5]), swap=1)
case("ops_gt", "t/ops.t:67", "gt", [arr("double", [3], [0, 1, 2]), arr("double", [1], [1.5]), NULL],
     out(2, "double", [3], [0, 0, 1]))
case("ops_shiftleft_byte", "t/ops.t:70-72", "shiftleft", [arr("byte", [3], [0, 1, 3]), scalar("sbyte", 2), NULL],
     out(2, "byte", [3], [0, 4, 12]))
case("ops_modulo_scalar", "t/ops.t:97", "modulo", [arr("double", [6], [12, 13, 14, 15, 16, 17]), scalar("sbyte", 3), NULL],
     out(2, "double", [6], [0, 1, 2, 0, 1, 2]))
case("ops_and2_floats_to_longlong", "t/ops.t:117", "and2", [arr("double", [3], [1, 0, 1]), arr("double", [3], [1, 1, 0]), NULL],
     out(2, "longlong", [3], [1, 0, 0]))
case("ops_or2_floats_to_longlong", "t/ops.t:118", "or2", [arr("double", [3], [1, 0, 1]), arr("double", [3], [1, 1, 0]), NULL],
     out(2, "longlong", [3], [1, 1, 1]))
case("ops_atan2", "t/ops.t:126-127", "atan2", [arr("double", [2], [1, 1]), arr("double", [2], [1, 1]), NULL],
     out(2, "double", [2], [math.atan2(1, 1)] * 2, tol=1e-15))
sa = list(range(12))
case("ops_or2_sequence", "t/ops.t:130-133", "or2", [seq("double", 3, 4), arr("double", [3, 4], [v + 1 for v in sa]), NULL],
     out(2, "longlong", [3, 4], [v | (v + 1) for v in sa]))
case("ops_and2_sequence", "t/ops.t:130-134", "and2", [seq("double", 3, 4), arr("double", [3, 4], [v + 1 for v in sa]), NULL],
     out(2, "longlong", [3, 4], [v & (v + 1) for v in sa]))
case("ops_minus_sequence", "t/ops.t:135-136", "minus", [arr("double", [3, 4], [v + 1 for v in sa]), seq("double", 3, 4), NULL],
     out(2, "double", [3, 4], [1] * 12))
case("ops_inplace_plus", "t/ops.t:143-145", "plus", [scalar("double", 1), scalar("sbyte", 1), "inplace:0"],
     out(0, "double", [], [2]))
case("ops_and_assign", "t/ops.t:189-192", "and2", [arr("double", [5], [1] * 5), scalar("sbyte", 0), "inplace:0"],
     out(0, "double", [5], [0] * 5))
case("ops_or_assign", "t/ops.t:193-194", "or2", [arr("double", [5], [0] * 5), scalar("sbyte", 1), "inplace:0"],
     out(0, "double", [5], [1] * 5))
case("ops_eq", "t/ops.t:195", "eq", [arr("double", [5], [1] * 5), arr("double", [5], [1] * 5), NULL],
     out(2, "double", [5], [1] * 5))
# modulus sign tables (t/ops.t:203-216): a = -7..7, b = (-3,0,3) transposed
ma = list(range(-7, 8))
mexp = [int(v) for v in ("-1 0 -2 " * 5).split()] + [0] * 15 + [int(v) for v in ("2 0 1 " * 5).split()]
for t in ["short", "long", "indx", "longlong", "float", "double"]:
    case("ops_modulus_" + t, "t/ops.t:203-216", "modulo", [arr(t, [15], ma), arr(t, [1, 3], [-3, 0, 3]), NULL],
         out(2, t, [15, 3], mexp))
uexp = [0] * 15 + [int(v) for v in ("0 1 2 " * 5).split()]
for t in ["byte", "ushort"]:
    case("ops_unsigned_modulus_" + t, "t/ops.t:218-227", "modulo", [arr(t, [15], range(15)), arr(t, [1, 2], [0, 3]), NULL],
         out(2, t, [15, 2], uexp))
case("ops_big_byte_modulus", "t/ops.t:233", "modulo", [scalar("byte", 255), scalar("sbyte", 1), NULL], out(2, "byte", [], [0]))
case("ops_big_ushort_modulus", "t/ops.t:234", "modulo", [scalar("ushort", 65535), scalar("sbyte", 1), NULL], out(2, "ushort", [], [0]))
for i, (x, y, op, e) in enumerate([
        (10555000100001145, 10555000100001144, "minus", 1),
        (9000000000000000002, 9000000000000000001, "minus", 1),
        (-8999999999999999998, 8999999999999999999, "plus", 1),
        (1000000000000000001, 1000000000000000000, "minus", 1),
        (9223372036854775807, 9223372036854775806, "minus", 1),
        (9223372036854775807, -9223372036854775808, "plus", -1)]):
    case("ops_longlong_precision_%d" % (i + 1), "t/ops.t:242-247", op, [scalar("longlong", x), scalar("longlong", y), NULL],
         out(2, "longlong", [], [e]))
INT_MAX = 2147483647
case("ops_big_long_modulus", "t/ops.t:256", "modulo", [scalar("long", INT_MAX), scalar("sbyte", 1), NULL], out(2, "long", [], [0]))
for t in ["indx", "longlong", "ulonglong", "double"]:
    case("ops_big_%s_modulus" % t, "t/ops.t:258-263", "modulo", [scalar(t, INT_MAX * 4), scalar("sbyte", 2), NULL],
         out(2, t, [], [0]))
case("ops_xor", "t/ops.t:273", "xor", [arr("double", [3], [1, 2, 3]), arr("double", [3], [4, 5, 6]), NULL],
     out(2, "longlong", [3], [5, 7, 5]))
case("ops_empty_dims", "t/ops.t:283", "and2", [zeroes("double", 1, 1, 0), zeroes("double", 1, 1, 0), NULL],
     out(2, "longlong", [1, 1, 0], []))

# ---------------------------------------------------------------- t/ufunc.t
nan = float("nan")
xn = arr("double", [4], [nan, 0, 1, 2])
yn = arr("double", [4], [0, 1, 2, nan])
case("ufunc_min_leading_nan", "t/ufunc.t:90", "minimum", [xn, NULL], out(1, "double", [], [0]))
case("ufunc_min_trailing_nan", "t/ufunc.t:91,97", "minimum", [yn, NULL], out(1, "double", [], [0]))
case("ufunc_max_leading_nan", "t/ufunc.t:92", "maximum", [xn, NULL], out(1, "double", [], [2]))
case("ufunc_max_trailing_nan", "t/ufunc.t:93", "maximum", [yn, NULL], out(1, "double", [], [2]))
case("ufunc_minmax_leading_nan", "t/ufunc.t:94", "minmaximum", [xn, NULL, NULL, NULL, NULL],
     [out(1, "double", [], [0]), out(2, "double", [], [2]), out(3, "indx", [], [1]), out(4, "indx", [], [3])])
case("ufunc_minmax_trailing_nan", "t/ufunc.t:95", "minmaximum", [yn, NULL, NULL, NULL, NULL],
     [out(1, "double", [], [0]), out(2, "double", [], [2]), out(3, "indx", [], [0]), out(4, "indx", [], [2])])
case("ufunc_max_of_empty_is_bad", "t/ufunc.t:100", "maximum", [arr("sbyte", [0], []), NULL],
     out(1, "sbyte", [], [-128], bad=[0], badflag=True))
case("ufunc_max_of_empty_badflag", "t/ufunc.t:103", "maximum", [arr("sbyte", [0], [], badflag=True), NULL],
     out(1, "sbyte", [], [-128], bad=[0], badflag=True))
bad0 = arr("double", [1], [-1.7976931348623157e308], bad=[0])
case("ufunc_min_of_bad", "t/ufunc.t:180-181", "minimum", [bad0, NULL], out(1, "double", [], [-1.7976931348623157e308], bad=[0], badflag=True))
case("ufunc_minmaximum_of_bad", "t/ufunc.t:183", "minmaximum", [bad0, NULL, NULL, NULL, NULL],
     [out(1, "double", [], [-1.7976931348623157e308], bad=[0], badflag=True)])
case("ufunc_ushort_max_not_bad", "t/ufunc.t:186", "maximum", [arr("ushort", [1], [65535]), NULL],
     out(1, "ushort", [], [65535], badflag=False))
case("ufunc_avg_long_empty", "t/ufunc.t:189", "average", [arr("long", [0], []), NULL], out(1, "long", [], [0]))
case("ufunc_average_double_empty", "t/ufunc.t:190", "average", [arr("double", [0], []), NULL], out(1, "double", [], [nan]))
X = arr("double", [3, 2], [5, 4, 3, 2, 3, 1.5])
case("ufunc_average", "t/ufunc.t:194", "average", [X, NULL], out(1, "double", [2], [4, 13.0 / 6], tol=1e-15))
case("ufunc_sumover", "t/ufunc.t:195", "sumover", [X, NULL], out(1, "double", [2], [12, 6.5]))
case("ufunc_prodover", "t/ufunc.t:196", "prodover", [X, NULL], out(1, "double", [2], [60, 9]))
IM = [1, 2, 3, 3, 5, 2, 3, 4, 5, 6, 13, 13, 13, 13, 13, 1, 3, 1, 3, 1, 10, 10, 2, 2, 2]
case("ufunc_minmax_IM", "t/ufunc.t:213-217", "minmaximum", [arr("double", [25], IM), NULL, NULL, NULL, NULL],
     [out(1, "double", [], [1]), out(2, "double", [], [13])])
case("ufunc_dsumover_bytes", "t/ufunc.t:219", "dsumover", [arr("byte", [3000], [1] * 3000), NULL], out(1, "double", [], [3000]))
P11 = [1, 2, 3, 4, 7, 9, 1, 1, 6, 2, 5]
case("ufunc_minimum_n_ind", "t/ufunc.t:224-231", "minimum_n_ind", [arr("double", [11], P11), NULL],
     out(1, "indx", [5], [0, 6, 7, 1, 9]), m_size=5)
BAD = -1.7976931348623157e308
case("ufunc_minimum_n_ind_bad", "t/ufunc.t:234-238", "minimum_n_ind",
     [arr("double", [11], [1, BAD, 3, 4, 7, 9, 1, 1, 6, 2, 5], bad=[1]), NULL],
     out(1, "indx", [5], [0, 6, 7, 9, 2]), m_size=5)
case("ufunc_minimum_n_ind_insufficient", "t/ufunc.t:240-244", "minimum_n_ind",
     [arr("double", [6], [1, BAD, 3, 4, BAD, BAD], bad=[1, 4, 5]), NULL],
     out(1, "indx", [5], [0, 2, 3, -2**63, -2**63], bad=[3, 4], badflag=True), m_size=5)
case("ufunc_minimum_n_ind_bad_good", "t/ufunc.t:246-250", "minimum_n_ind",
     [arr("double", [11], [1, BAD, 3, 4, BAD, BAD, 3, 1, 5, 8, 9], bad=[1, 4, 5]), NULL],
     out(1, "indx", [5], [0, 7, 2, 6, 3]), m_size=5)
case("core_maximum_n_ind", "t/core.t:242", "maximum_n_ind", [arr("double", [5], [0, 5, 10, 15, 20]), NULL],
     out(1, "indx", [3], [4, 3, 2]), m_size=3)

# ---------------------------------------------------------------- t/primitive-matmult.t
IMM = arr("double", [5, 5], IM)
case("matmult_IM", "t/primitive-matmult.t:21-27", "matmult", [IMM, IMM, NULL],
     out(2, "double", [5, 5], [97, 106, 63, 71, 69, 125, 140, 87, 100, 97, 351, 403, 299, 338, 351,
                               33, 43, 33, 42, 41, 78, 102, 102, 116, 142]), pmcode="matmult")
PA = arr("double", [4, 3], [1, 2, 3, 0, 1, -1, 2, 7, 1, 0, 0, 1])
PB = arr("double", [2, 4], [1, 1, 0, 2, 0, 2, 1, 1])
PC = [1, 11, 8, 10, 2, 2]
case("matmult_PA_PB", "t/primitive-matmult.t:40-51", "matmult", [PA, PB, NULL], out(2, "double", [2, 3], PC), pmcode="matmult")
# PA->dummy(0,3)->dummy(-1,3)->make_physical->slice('(1),,,(1)'): the
# physical (3,4,3,3) copy holds PA[t,h] at 1 + 3*t + 12*h (+36*k)
phys = [0.0] * (3 * 4 * 3 * 3)
for k in range(3):
    for h in range(3):
        for t in range(4):
            for j in range(3):
                phys[j + 3 * t + 12 * h + 36 * k] = PA["data"][t + 4 * h]
case("matmult_sliced_input", "t/primitive-matmult.t:53-57", "matmult",
     [arr("double", [3, 4, 3, 3], phys, view=[["slice", 1, None, None, 1]]), PB, NULL],
     out(2, "double", [2, 3], PC), pmcode="matmult")
case("matmult_output_zeroes", "t/primitive-matmult.t:59-63", "matmult", [PA, PB, zeroes("double", 2, 3)],
     out(2, "double", [2, 3], PC), pmcode="matmult")
case("matmult_output_ones", "t/primitive-matmult.t:65-69", "matmult", [PA, PB, arr("double", [2, 3], [1] * 6)],
     out(2, "double", [2, 3], PC), pmcode="matmult")
EQ = arr("float", [4, 1], [1, 1, 1, 1])
case("matmult_collapse", "t/primitive-matmult.t:72", "matmult", [EQ, PB, NULL], out(2, "double", [2, 1], [2, 6]), pmcode="matmult")
case("matmult_mismatch", "t/primitive-matmult.t:75-79", "matmult", [PB, EQ, NULL], {"error": "mismatch in matmult"}, pmcode="matmult")
case("matmult_scalar_rhs", "t/primitive-matmult.t:81", "matmult", [PB, scalar("sbyte", 2), NULL],
     out(2, "double", [2, 4], [2 * v for v in PB["data"]]), pmcode="matmult")
case("matmult_scalar_lhs", "t/primitive-matmult.t:82", "matmult", [arr("double", [], [3]), PB, NULL],
     out(2, "double", [2, 4], [3 * v for v in PB["data"]]), pmcode="matmult")
case("matmult_nan_row", "t/primitive-matmult.t:84-91", "matmult",
     [arr("double", [3, 3], [1, nan, 0, 0, 1, 0, 0, 0, 1]), seq("double", 2, 3), NULL],
     out(2, "double", [2, 3], [nan, nan, 2, 3, 4, 5]), pmcode="matmult")
case("matmult_bad_row", "t/primitive-matmult.t:93-99", "matmult",
     [arr("double", [3, 3], [1, BAD, 0, 0, 1, 0, 0, 0, 1], bad=[1]), seq("double", 2, 3), NULL],
     out(2, "double", [2, 3], [BAD, BAD, 2, 3, 4, 5], bad=[0, 1], badflag=True), pmcode="matmult")

# ---------------------------------------------------------------- t/bad.t
case("bad_byte_addition", "t/bad.t:106-111", "plus",
     [arr("byte", [3], [1, 2, 3], badflag=True), arr("byte", [3], [1, 255, 3], bad=[1]), NULL],
     out(2, "byte", [3], [2, 255, 6], bad=[1], badflag=True))
case("bad_float_sum", "t/bad.t:114-116", "sumover", [arr("float", [3], [1, -3.4028234663852886e38, 3], bad=[1]), NULL],
     out(1, "float", [], [4], badflag=True))
XB = arr("byte", [5, 2], [1, 2, 255, 4, 5, 255, 0, 1, 2, 255], bad=[2, 5, 9])
case("bad_sum_of_slice", "t/bad.t:132-134", "sumover", [dict(XB, view=[["slice", None, 1]]), NULL],
     out(1, "long", [], [3], badflag=True))
case("bad_inplace_slice_increment", "t/bad.t:132-136", "plus", [dict(XB, view=[["slice", None, 1]]), scalar("sbyte", 1), "inplace:0"],
     out(0, "byte", [5, 2], [1, 2, 255, 4, 5, 255, 1, 2, 3, 255], bad=[2, 5, 9], badflag=True), check_root=True)
Y9 = arr("double", [3, 3], [0, BAD, 2, BAD, 4, BAD, 6, BAD, 8], bad=[1, 3, 5, 7])
case("bad_inplace_biop_plus", "t/bad.t:289-293", "plus", [dict(seq("double", 3, 3)), Y9, "inplace:0"],
     out(0, "double", [3, 3], [0, BAD, 4, BAD, 8, BAD, 12, BAD, 16], bad=[1, 3, 5, 7], badflag=True))
Y3 = arr("double", [3, 3], [0, BAD, BAD, 3, BAD, BAD, 6, BAD, BAD], bad=[1, 2, 4, 5, 7, 8])
case("bad_inplace_bifunc_power", "t/bad.t:296-300", "power", [dict(seq("double", 3, 3)), Y3, "inplace:0"],
     out(0, "double", [3, 3], [1, BAD, BAD, 27, BAD, BAD, 6.0 ** 6, BAD, BAD], bad=[1, 2, 4, 5, 7, 8], badflag=True))

# Bad.pd elementwise ops (t/bad.t)
DMAX = 1.7976931348623157e308       # double's default bad value is -DBL_MAX (Types.pm defbval)
FMAX = 3.4028234663852886e38
case("bad_isbad", "t/bad.t:118-119", "isbad", [arr("byte", [9], [1, 2, 255, 255, 5, 6, 255, 8, 9], bad=[2, 3, 6]), NULL],
     out(1, "long", [9], [0, 0, 1, 1, 0, 0, 1, 0, 0]))
case("bad_isgood", "t/bad.t:118-120", "isgood", [arr("byte", [9], [1, 2, 255, 255, 5, 6, 255, 8, 9], bad=[2, 3, 6]), NULL],
     out(1, "long", [9], [1, 1, 0, 0, 1, 1, 0, 1, 1]))
X20 = [42, 47, 98, 13, 22, 96, 74, 41, 79, 76, 96, 3, 32, 76, 25, 59, 5, 96, 32, 6]
Y20 = [(-DMAX if v < 20 else v) for v in X20]
case("bad_setbadif_lt20", "t/bad.t:173-182", "setbadif",
     [arr("double", [20], X20), arr("double", [20], [1.0 if v < 20 else 0.0 for v in X20]), NULL],
     out(2, "double", [20], Y20, badflag=True))
case("bad_setbadtoval_20", "t/bad.t:175-176", "setbadtoval",
     [arr("double", [20], Y20, bad=[i for i, v in enumerate(X20) if v < 20]), NULL],
     out(1, "double", [20], [20 if v < 20 else v for v in X20], badflag=False), dpar=20.0)
case("bad_setvaltobad_seq_mod4", "t/bad.t:318-321", "setvaltobad", [arr("double", [10], [i % 4 for i in range(10)]), NULL],
     out(1, "double", [10], [0, -DMAX, 2, 3, 0, -DMAX, 2, 3, 0, -DMAX], badflag=True), dpar=1.0)
case("bad_setbadtonan", "t/bad.t:323-324", "setbadtonan",
     [arr("double", [10], [0, -DMAX, 2, 3, 0, -DMAX, 2, 3, 0, -DMAX], bad=[1, 5, 9]), NULL],
     out(1, "double", [10], [0, math.nan, 2, 3, 0, math.nan, 2, 3, 0, math.nan], badflag=False))
case("bad_setvaltobad_float", "t/bad.t:327", "setvaltobad", [arr("float", [4], [1, 2, 3, 4]), NULL],
     out(1, "float", [4], [1, -FMAX, 3, 4], badflag=True), dpar=2.0)
case("bad_setvaltobad_double", "t/bad.t:328", "setvaltobad", [arr("double", [4], [1, 2, 3, 4]), NULL],
     out(1, "double", [4], [1, -DMAX, 3, 4], badflag=True), dpar=2.0)
case("bad_setinftobad", "t/bad.t:330-332", "setinftobad", [arr("double", [3], [0, math.inf, math.nan]), NULL],
     out(1, "double", [3], [0, -DMAX, math.nan], badflag=True))
case("bad_setnonfinitetobad", "t/bad.t:334-336", "setnonfinitetobad",
     [arr("double", [10], [0, math.inf, 2, 3, 0, math.nan, 2, 3, 0, math.nan]), NULL],
     out(1, "double", [10], [0, -DMAX, 2, 3, 0, -DMAX, 2, 3, 0, -DMAX], badflag=True))
case("bad_setnantobad", "t/bad.t:323-343", "setnantobad",
     [arr("double", [10], [0, math.nan, 2, 3, 0, math.nan, 2, 3, 0, math.nan]), NULL],
     out(1, "double", [10], [0, -DMAX, 2, 3, 0, -DMAX, 2, 3, 0, -DMAX], badflag=True))

# Ops.pd ufuncs (t/ops.t)
case("ops_exp", "t/ops.t:111-112", "exp", [arr("double", [2], [2, 3]), NULL],
     out(1, "double", [2], [7.3891, 20.0855], tol=1e-2))
case("ops_sqrt_approx", "t/ops.t:111-113", "sqrt", [arr("double", [2], [2, 3]), NULL],
     out(1, "double", [2], [1.4142, 1.7321], tol=1e-2))
case("ops_sqrt_exact", "t/ops.t:76-78", "sqrt", [arr("double", [3], [16, 64, 9]), NULL],
     out(1, "double", [3], [4, 8, 3]))
case("ops_log10", "t/ops.t:178-180", "log10", [arr("double", [2], [110, 23]), NULL],
     out(1, "double", [2], [math.log(110) / math.log(10), math.log(23) / math.log(10)], tol=1e-12))
case("ops_log_float_type", "t/ops.t:173-174", "log", [arr("float", [5, 4], [i + 2 for i in range(20)]), NULL],
     out(1, "float", [5, 4], [math.log(i + 2) for i in range(20)], tol=1e-6))
case("ops_bitnot_double_to_longlong", "t/ops.t:272", "bitnot", [arr("double", [3], [1, 2, 3]), NULL],
     out(1, "longlong", [3], [-2, -3, -4]))
case("ops_bitnot_ulonglong", "t/ops.t:267-269", "bitnot", [arr("ulonglong", [2], [0, 1]), NULL],
     out(1, "ulonglong", [2], [18446744073709551615, 18446744073709551614]))

# ---------------------------------------------------------------- t/core.t dim compatibility (append)
def z(*dims):
    return zeroes("double", *dims)


def one(v):
    return arr("double", [], [v])


APP = [
    ([z(1), z(1), z(1)], "dim has size 1"),
    ([one(1), one(2), NULL], ("double", [2], [1, 2])),
    ([one(1), one(2), z(2)], ("double", [2], [1, 2])),
    ([z(1), z(1), z(3)], "dim has size 3"),
    ([z(1), z(0), z()], ("double", [], [0])),
    ([z(1), z(1), z()], "can't broadcast"),
    ([z(1), z(1), z(1, 1)], "dim has size 1"),
    ([one(1), one(2), z(2, 1)], ("double", [2, 1], [1, 2])),
    ([z(1), z(1), z(3, 1)], "dim has size 3"),
    ([z(1), z(1), z(1, 2)], "dim has size 1"),
    ([z(1), z(1), z(2, 2)], ("double", [2, 2], [0, 0, 0, 0])),
    ([z(1, 2), z(1), z(2, 2)], ("double", [2, 2], [0, 0, 0, 0])),
    ([z(1, 3), z(1), z(2, 2)], "Mismatch"),
    ([z(1, 2), z(1), z(2, 1)], "implicit dim"),
    ([z(1, 2), z(1), z(2)], "implicit dim"),
    ([z(1, 2), z(1, 2), z(2, 1)], "implicit dim"),
    ([z(1, 2), z(1, 2), dict(z(2), view=[["dummy", 1, 2]])], "implicit dim"),
    ([z(1, 2), z(1, 2), dict(z(2), view=[["dummy", 1, 2]])], "implicit dim"),
    ([z(1, 2), z(1, 2), dict(z(2), view=[["dummy", 0, 2]])], "over dummy dim"),
    ([z(1, 2), z(1, 2), dict(z(2), view=[["dummy", 0, 2]])], "over dummy dim"),
]
for i, (args, exp) in enumerate(APP):
    if isinstance(exp, str):
        e = {"error": exp}
    else:
        e = out(2, *exp)
    case("core_dim_compat_%02d" % (i + 1), "t/core.t:%d" % (252 + i), "append", args, e, pmcode="append")

# ---------------------------------------------------------------- t/pthread.t
pthread_rows = [
    [6, [6], 6, 0], [6, [5], 5, 0], [6, [4], 4, 0], [6, [7], 6, 0], [6, [7, 12], 6, 1], [6, [5, 12], 6, 1],
    [6, [12, 7], 6, 0], [6, [12, 5], 6, 0], [6, [7, 5], 6, 0], [6, [4, 5], 5, 1], [6, [5, 4], 5, 0],
    [6, [4, 5, 12], 6, 2], [4, [9, 6, 2], 4, 1], [4, [6, 9, 2], 4, 0]]

doc = {"source": "PDL 2.106 t/*.t inline literals", "cases": cases,
       "pthread": {"src": "t/pthread.t:12-33", "rows": pthread_rows}}
with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
    json.dump(doc, f, indent=1, allow_nan=True)
print("wrote %d cases" % len(cases))
