"""pdl_amd: PDL's broadcast-loop compute path (PDL::Ops binops, PDL::Ufunc
reductions and scans, PDL::Primitive matmult) on MI355X.

Host side: a mirror of PDL's ndarray API (core.py) over the C ABI of
include/pdlhip.h (libpdlhip.so, built by build.py); compute: hand-written
HIP kernels for gfx950.
"""
from ._lib import PDLError  # noqa: F401
from .core import (TYPE_NAMES, Chain, Event, Graph, MatmultList, Ragged, Timer, matmult_list, view_into, default_badvalue, empty, init, matmult, ndarray, null, ones,  # noqa: F401
                   grandom, pdl, random, randsym, run, sequence, srandom, set_matmult_precision, synchronize,
                   topdl, typecode, zeroes)
from . import core as _core

for _n in list(_core.L.BIOPS) + list(_core.L.UNARY):
    globals()[_n] = getattr(_core, _n)

SB, B, S, US, L, UL, IND, ULL, LL, F, D = range(11)
