"""ctypes binding of libpdlhip.so (include/pdlhip.h).

Loads the in-tree library built by pdl_amd/build.py and raises if it is
missing: there is no CPU fallback for any op.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libpdlhip.so")

SB, B, S, US, L, UL, IND, ULL, LL, F, D, LD, CF, CD, CLD = range(15)
NOMYDIMS, BADVAL, VAFFINE = 1, 2, 4
MAXDIMS = 16
MAXBIDS = 8
ENONE, EUSERERROR, EFATAL = 0, 1, 2
MM_SPLIT16, MM_FP32, MM_PDLORDER, MM_FP8 = 0, 1, 2, 3


class Value(ctypes.Union):
    _fields_ = [("A", ctypes.c_int8), ("B", ctypes.c_uint8), ("S", ctypes.c_int16),
                ("U", ctypes.c_uint16), ("L", ctypes.c_int32), ("K", ctypes.c_uint32),
                ("N", ctypes.c_int64), ("P", ctypes.c_uint64), ("Q", ctypes.c_int64),
                ("F", ctypes.c_float), ("D", ctypes.c_double)]


class Anyval(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("pad_", ctypes.c_int32), ("value", Value)]


class Pdl(ctypes.Structure):
    """pdlhip_pdl"""
    _fields_ = [("datatype", ctypes.c_int32), ("state", ctypes.c_int32), ("ndims", ctypes.c_int32),
                ("has_badvalue", ctypes.c_int32), ("dims", ctypes.c_int64 * MAXDIMS),
                ("incs", ctypes.c_int64 * MAXDIMS), ("offs", ctypes.c_int64),
                ("data", ctypes.c_void_p), ("badvalue", Anyval),
                ("nbroadcastids", ctypes.c_int32), ("pad2_", ctypes.c_int32),
                ("broadcastids", ctypes.c_int64 * MAXBIDS)]


class Error(ctypes.Structure):
    _fields_ = [("error", ctypes.c_int32), ("message", ctypes.c_char_p)]


class Broadcast(ctypes.Structure):
    """pdlhip_broadcast"""
    _fields_ = [("npdls", ctypes.c_int32), ("ndims", ctypes.c_int32),
                ("dims", ctypes.c_int64 * (MAXDIMS + 2)), ("incs", ctypes.c_int64 * ((MAXDIMS + 2) * 6)),
                ("offs", ctypes.c_int64 * 6), ("ind_sizes", ctypes.c_int64 * 4),
                ("inc_sizes", ctypes.c_int64 * 12)]


class Node(ctypes.Structure):
    """pdlhip_node: one binop of a fused ragged chain"""
    _fields_ = [("op", ctypes.c_int32), ("x", ctypes.c_int32), ("y", ctypes.c_int32)]


class PDLError(RuntimeError):
    """an error returned through the C ABI (PDL's barf)"""

    def __init__(self, kind, message=None):
        if message is None:
            kind, message = EUSERERROR, kind
        super().__init__(message)
        self.kind = kind


PP = ctypes.POINTER(Pdl)

EXPORTS = {
    "pdlhip_init": ([ctypes.c_int], Error),
    "pdlhip_shutdown": ([], Error),
    "pdlhip_abi_version": ([], ctypes.c_int),
    "pdlhip_alloc": ([ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)], Error),
    "pdlhip_free": ([ctypes.c_void_p], Error),
    "pdlhip_upload": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t], Error),
    "pdlhip_download": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t], Error),
    "pdlhip_synchronize": ([], Error),
    "pdlhip_stream": ([], ctypes.c_void_p),
    "pdlhip_run": ([ctypes.c_char_p, ctypes.POINTER(PP), ctypes.c_int, ctypes.POINTER(ctypes.c_int64)], Error),
    "pdlhip_set_matmult_precision": ([ctypes.c_int], Error),
    "pdlhip_get_matmult_precision": ([], ctypes.c_int),
    "pdlhip_readdata": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Broadcast),
                         ctypes.POINTER(PP), ctypes.POINTER(ctypes.c_int64)], Error),
    "pdlhip_set_plan_only": ([ctypes.c_int], Error),
    "pdlhip_timer_start": ([], Error),
    "pdlhip_timer_stop": ([ctypes.POINTER(ctypes.c_float)], Error),
    "pdlhip_event_create": ([ctypes.POINTER(ctypes.c_void_p)], Error),
    "pdlhip_event_record": ([ctypes.c_void_p], Error),
    "pdlhip_event_elapsed": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)], Error),
    "pdlhip_event_destroy": ([ctypes.c_void_p], Error),
    "pdlhip_fill_uniform": ([PP, ctypes.c_uint64, ctypes.c_double, ctypes.c_double], Error),
    "pdlhip_fill_normal": ([PP, ctypes.c_uint64, ctypes.c_double, ctypes.c_double], Error),
}
BIOPS = ["plus", "mult", "minus", "divide", "gt", "lt", "le", "ge", "eq", "ne", "shiftleft", "shiftright",
         "or2", "and2", "xor", "power", "atan2", "modulo", "spaceship"]
UNARY = ["assgn", "sumover", "prodover", "dsumover", "dprodover", "cumusumover", "cumuprodover",
         "dcumusumover", "dcumuprodover", "average", "daverage", "minimum", "maximum", "minimum_ind",
         "maximum_ind"]
for _n in BIOPS:
    EXPORTS["pdlhip_run_" + _n] = ([PP, PP, PP, ctypes.c_int], Error)
for _n in UNARY:
    EXPORTS["pdlhip_run_" + _n] = ([PP, PP], Error)
EXPORTS["pdlhip_run_minmaximum"] = ([PP] * 5, Error)
EXPORTS["pdlhip_run_minimum_n_ind"] = ([PP, PP, ctypes.c_int64], Error)
EXPORTS["pdlhip_run_maximum_n_ind"] = ([PP, PP, ctypes.c_int64], Error)
EXPORTS["pdlhip_run_matmult"] = ([PP, PP, PP], Error)
UFUNCS = ["bitnot", "sqrt", "sin", "cos", "not", "exp", "log", "log10", "abs"]
BADOPS = ["isbad", "isgood", "setnantobad", "setinftobad", "setnonfinitetobad", "setbadtonan"]
for _n in UFUNCS + BADOPS:
    EXPORTS["pdlhip_run_" + _n] = ([PP, PP], Error)
EXPORTS["pdlhip_run_setbadif"] = ([PP, PP, PP], Error)
EXPORTS["pdlhip_run_reduce_noalloc"] = ([ctypes.c_char_p, PP, PP], Error)
EXPORTS["pdlhip_tune_stats"] = ([ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)], Error)
EXPORTS["pdlhip_op_info"] = ([ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)], Error)
# Slices.pd index family and the dataflow writeback
EXPORTS["pdlhip_run_index"] = ([PP, PP, PP], Error)
EXPORTS["pdlhip_run_index1d"] = ([PP, PP, PP], Error)
EXPORTS["pdlhip_run_index2d"] = ([PP, PP, PP, PP], Error)
EXPORTS["pdlhip_writebackdata"] = ([ctypes.c_char_p, ctypes.POINTER(PP), ctypes.c_int], Error)
for _n in ("pdlhip_run_rangeb", "pdlhip_writeback_rangeb"):
    EXPORTS[_n] = ([PP, PP, ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, ctypes.c_char_p, PP], Error)
# PDL's random-number stream
EXPORTS["pdlhip_srandom"] = ([ctypes.c_uint64], Error)
EXPORTS["pdlhip_rand_skip"] = ([ctypes.c_uint64], Error)
EXPORTS["pdlhip_rand_state"] = ([ctypes.POINTER(ctypes.c_uint64)], Error)
for _n in ("pdlhip_run_random", "pdlhip_run_randsym", "pdlhip_grandom"):
    EXPORTS[_n] = ([PP], Error)
EXPORTS["pdlhip_run_setvaltobad"] = ([PP, PP, ctypes.c_double], Error)
EXPORTS["pdlhip_run_setbadtoval"] = ([PP, PP, ctypes.c_double], Error)
EXPORTS["pdlhip_ragged_prepare"] = ([ctypes.c_int, ctypes.c_int, ctypes.POINTER(PP), ctypes.c_int,
                                     ctypes.POINTER(Node), ctypes.c_char_p, ctypes.POINTER(PP), ctypes.POINTER(PP),
                                     ctypes.POINTER(ctypes.c_void_p)], Error)
EXPORTS["pdlhip_chain_prepare"] = ([ctypes.c_int, ctypes.POINTER(PP), ctypes.c_int, ctypes.POINTER(Node),
                                    ctypes.c_char_p, PP, PP, ctypes.POINTER(ctypes.c_void_p)], Error)
EXPORTS["pdlhip_ragged_count"] = ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)], Error)
EXPORTS["pdlhip_ragged_run"] = ([ctypes.c_void_p], Error)
EXPORTS["pdlhip_ragged_badflags"] = ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)], Error)
EXPORTS["pdlhip_ragged_free"] = ([ctypes.c_void_p], Error)
EXPORTS["pdlhip_graph_begin"] = ([], Error)
EXPORTS["pdlhip_graph_end"] = ([ctypes.POINTER(ctypes.c_void_p)], Error)
EXPORTS["pdlhip_graph_launch"] = ([ctypes.c_void_p], Error)
EXPORTS["pdlhip_graph_destroy"] = ([ctypes.c_void_p], Error)
EXPORTS["pdlhip_matmult_list_prepare"] = ([ctypes.c_int, ctypes.POINTER(PP), ctypes.POINTER(PP), ctypes.POINTER(PP),
                                           ctypes.POINTER(ctypes.c_void_p)], Error)
EXPORTS["pdlhip_matmult_list_launch"] = ([ctypes.c_void_p, ctypes.c_int], Error)
EXPORTS["pdlhip_matmult_list_free"] = ([ctypes.c_void_p], Error)
EXPORTS["pdlhip_run_matmult_list"] = ([ctypes.c_int, ctypes.POINTER(PP), ctypes.POINTER(PP), ctypes.POINTER(PP),
                                      ctypes.c_int], Error)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("pdl_amd: %s is not built (run `python pdl_amd/build.py`); "
                               "there is no CPU fallback" % LIBPATH)
        l = ctypes.CDLL(LIBPATH)
        for name, (args, res) in EXPORTS.items():
            f = getattr(l, name)
            f.argtypes = args
            f.restype = res
        _lib = l
    return _lib


def check(e):
    if e.error:
        raise PDLError(e.error, e.message.decode() if e.message else "pdlhip error %d" % e.error)
