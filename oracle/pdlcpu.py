"""ctypes binding of the ORACLE (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product package (pdl_amd) never imports it.

`OPdl` is a minimal host-side ndarray in PDL's layout (dim 0 fastest, a root
buffer plus composed virtual-affine dims/incs/offs, PDL's bad-value state)
with the zero-copy views of Slices.pd (xchg, mv, dummy, slice) and
`run(op, ...)`, which calls the C restatement of pdl_run_<op>.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")

SB, B, S, US, L, UL, IND, ULL, LL, F, D, LD, CF, CD, CLD = range(15)
TYPE_NAMES = ["sbyte", "byte", "short", "ushort", "long", "ulong", "indx", "ulonglong",
              "longlong", "float", "double", "ldouble", "cfloat", "cdouble", "cldouble"]
NP_TYPES = {SB: np.int8, B: np.uint8, S: np.int16, US: np.uint16, L: np.int32, UL: np.uint32,
            IND: np.int64, ULL: np.uint64, LL: np.int64, F: np.float32, D: np.float64,
            CF: np.complex64, CD: np.complex128}
NOMYDIMS, BADVAL, VAFFINE = 1, 2, 4
MAXDIMS = 16


class Value(ctypes.Union):
    _fields_ = [("A", ctypes.c_int8), ("B", ctypes.c_uint8), ("S", ctypes.c_int16),
                ("U", ctypes.c_uint16), ("L", ctypes.c_int32), ("K", ctypes.c_uint32),
                ("N", ctypes.c_int64), ("P", ctypes.c_uint64), ("Q", ctypes.c_int64),
                ("F", ctypes.c_float), ("D", ctypes.c_double)]


class Anyval(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("pad_", ctypes.c_int32), ("value", Value)]


class CPdl(ctypes.Structure):
    _fields_ = [("datatype", ctypes.c_int32), ("state", ctypes.c_int32), ("ndims", ctypes.c_int32),
                ("has_badvalue", ctypes.c_int32), ("dims", ctypes.c_int64 * MAXDIMS),
                ("incs", ctypes.c_int64 * MAXDIMS), ("offs", ctypes.c_int64),
                ("data", ctypes.c_void_p), ("badvalue", Anyval),
                ("nbroadcastids", ctypes.c_int32), ("pad2_", ctypes.c_int32),
                ("broadcastids", ctypes.c_int64 * 8)]


class CPlan(ctypes.Structure):
    """pdlc_plan: the broadcast plan of a run (pdlbroadcast.h:18-61 fields,
    core ind_sizes / inc_sizes), captured for readdata-level tests"""
    _fields_ = [("npdls", ctypes.c_int), ("ndims", ctypes.c_int), ("dims", ctypes.c_int64 * 20),
                ("incs", ctypes.c_int64 * 120), ("offs", ctypes.c_int64 * 6), ("ind_sizes", ctypes.c_int64 * 3),
                ("inc_sizes", ctypes.c_int64 * 12), ("datatype", ctypes.c_int), ("bvalflag", ctypes.c_int)]


class CError(ctypes.Structure):
    _fields_ = [("error", ctypes.c_int32), ("message", ctypes.c_char * 4096)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        l = ctypes.CDLL(LIBPATH)
        l.pdlc_run.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(CPdl)), ctypes.c_int,
                               ctypes.POINTER(ctypes.c_int64), ctypes.c_int, ctypes.c_int,
                               ctypes.POINTER(CError)]
        l.pdlc_run.restype = ctypes.c_int
        l.pdlc_free.argtypes = [ctypes.c_void_p]
        l.pdlc_find_max_pthread.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_int)]
        l.pdlc_random_fill.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64]
        l.pdlc_random_fill_from.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint64]
        l.pdlc_rand_state_after.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64]
        l.pdlc_rangeb.argtypes = [ctypes.POINTER(CPdl), ctypes.POINTER(CPdl), ctypes.c_void_p, ctypes.c_int64,
                                  ctypes.c_char_p, ctypes.POINTER(CPdl), ctypes.c_int, ctypes.POINTER(CError)]
        l.pdlc_ndtri.argtypes = [ctypes.c_double]
        l.pdlc_ndtri.restype = ctypes.c_double
        l.pdlc_ndtri_n.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        l.pdlc_xoshiro256plus_next.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        l.pdlc_xoshiro256plus_next.restype = ctypes.c_uint64
        l.pdlc_splitmix64_next.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        l.pdlc_splitmix64_next.restype = ctypes.c_uint64
        l.pdlc_truth_matmult_ld.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int64] * 3
        l.pdlc_convert.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64]
        l.pdlc_capture_plan.argtypes = [ctypes.POINTER(CPlan)]
        _lib = l
    return _lib


class OracleError(Exception):
    pass


def default_badval(t):
    info = {SB: -128, B: 255, S: -32768, US: 65535, L: -2**31, UL: 2**32 - 1, IND: -2**63,
            ULL: 2**64 - 1, LL: -2**63}
    if t in info:
        return info[t]
    if t == F:
        return -np.finfo(np.float32).max
    if t == D:
        return -np.finfo(np.float64).max
    raise ValueError(t)


class OPdl:
    """Host ndarray in PDL layout (dim 0 fastest)."""

    def __init__(self, root, dtype, dims, incs=None, offs=0, state=0, badvalue=None):
        self.root = root            # 1-D numpy array (the root buffer), or None for null
        self.dtype = dtype
        self.dims = list(dims)
        if incs is None:
            incs, inc = [], 1
            for d in self.dims:
                incs.append(inc)
                inc *= d
        self.incs = list(incs)
        self.offs = offs
        self.state = state
        self.badvalue = badvalue    # None -> type default
        self.root_dims = list(self.dims)
        self.bids = None            # pdl->broadcastids when set by broadcastI

    # ---- constructors ------------------------------------------------------
    @classmethod
    def from_numpy(cls, arr, dtype=None, badflag=False, badvalue=None):
        arr = np.asarray(arr)
        if dtype is None:
            dtype = {np.dtype(v): k for k, v in NP_TYPES.items() if k != IND}.get(arr.dtype, D)
        root = np.ascontiguousarray(arr, dtype=NP_TYPES[dtype]).reshape(-1).copy()
        p = cls(root, dtype, list(arr.shape[::-1]))
        if badflag:
            p.state |= BADVAL
        p.badvalue = badvalue
        return p

    @classmethod
    def null(cls):
        return cls(None, -1, [], state=NOMYDIMS)

    @property
    def nvals(self):
        n = 1
        for d in self.dims:
            n *= d
        return n

    def isnull(self):
        return bool(self.state & NOMYDIMS)

    # ---- views (Slices.pd affine transforms) ---------------------------------
    def _view(self, dims, incs, offs):
        v = OPdl(self.root, self.dtype, dims, incs, offs, (self.state & BADVAL) | VAFFINE, self.badvalue)
        v.root_dims = self.root_dims
        v.parent = self
        return v

    def root_pdl(self):
        p = self
        while getattr(p, "parent", None) is not None:
            p = p.parent
        p.state |= self.state & BADVAL
        return p

    def xchg(self, i, j):
        dims, incs = list(self.dims), list(self.incs)
        dims[i], dims[j] = dims[j], dims[i]
        incs[i], incs[j] = incs[j], incs[i]
        return self._view(dims, incs, self.offs)

    def mv(self, i, j):
        dims, incs = list(self.dims), list(self.incs)
        d, c = dims.pop(i), incs.pop(i)
        dims.insert(j, d)
        incs.insert(j, c)
        return self._view(dims, incs, self.offs)

    def dummy(self, pos, size=1):
        dims, incs = list(self.dims), list(self.incs)
        if pos < 0:
            pos = len(dims) + 1 + pos
        while len(dims) < pos:
            dims.append(1)
            incs.append(0)
        dims.insert(pos, size)
        incs.insert(pos, 0)
        return self._view(dims, incs, self.offs)

    def slice(self, *specs):
        """specs per dim: None (all), int k -> '(k)' (drop dim), (start, stop, step) inclusive stop
        as PDL's 'a:b:s'."""
        dims, incs, offs = [], [], self.offs
        for d in range(len(self.dims)):
            sp = specs[d] if d < len(specs) else None
            if sp is None:
                dims.append(self.dims[d])
                incs.append(self.incs[d])
            elif isinstance(sp, int):
                k = sp if sp >= 0 else self.dims[d] + sp
                offs += k * self.incs[d]
            else:
                a, b, s = sp
                a = a if a >= 0 else self.dims[d] + a
                b = b if b >= 0 else self.dims[d] + b
                # Slices.pd:2367-2386 bounds
                if not (0 <= a < self.dims[d] and 0 <= b < self.dims[d]):
                    raise OracleError("slice out of bounds in pos %d" % d)
                n = (b - a) // s + 1
                dims.append(max(n, 0))
                incs.append(self.incs[d] * s)
                offs += a * self.incs[d]
        return self._view(dims, incs, offs)

    # ---- explicit broadcast ids --------------------------------------------------
    def broadcastids(self):
        return list(self.bids) if self.bids else [len(self.dims)]

    def broadcastI(self, bid, whichdims):
        """Slices.pd:1819-1898 RedoDims, step for step: the child keeps the
        parent's ndims; the dims not named go in order, leaving room for
        the named ones at broadcastids[bid] (or the end), which then go
        there; broadcastids shift by the number moved."""
        wd = list(whichdims)
        for i in range(len(wd)):
            for j in range(i + 1, len(wd)):
                if wd[i] == wd[j] and wd[i] != -1:
                    raise OracleError("Error in broadcastI:duplicate arg %d %d %d" % (i, j, wd[i]))
        nrealwhichdims = sum(1 for w in wd if w != -1)
        pbids, nd = self.broadcastids(), len(self.dims)
        dims, incs = [None] * nd, [None] * nd
        nthc = 0
        for i in range(nd):
            if len(pbids) > bid and bid >= 0 and i == pbids[bid]:
                nthc += len(wd)
            if i in wd:
                continue
            dims[nthc], incs[nthc] = self.dims[i], self.incs[i]
            nthc += 1
        for i in range(len(wd)):
            cdim = i + (pbids[bid] if len(pbids) > bid and bid >= 0 else nd) - nrealwhichdims
            if wd[i] == -1:
                dims[cdim], incs[cdim] = 1, 0
            else:
                dims[cdim], incs[cdim] = self.dims[wd[i]], self.incs[wd[i]]
        if None in dims:
            raise OracleError("broadcastI leaves a dim unset")
        n = max(bid + 1, len(pbids))
        cb = []
        for i in range(n - 1):
            cb.append((pbids[i] if len(pbids) > i else nd) +
                      (-nrealwhichdims if i <= bid else len(wd) - nrealwhichdims))
        cb.append(nd)
        v = self._view(dims, incs, self.offs)
        v.bids = cb
        return v

    def unbroadcast(self, atind):
        """Slices.pd:1900-1925"""
        nd, b0 = len(self.dims), self.broadcastids()[0]
        dims, incs = [None] * nd, [None] * nd
        for i in range(nd):
            corc = i if i < atind else (i + nd - b0 if i < b0 else i - b0 + atind)
            dims[corc], incs[corc] = self.dims[i], self.incs[i]
        return self._view(dims, incs, self.offs)

    def dimstr(self):
        """Core.pm:2105-2117"""
        ids = self.broadcastids()
        out = "D [%s]" % ",".join(map(str, self.dims[:ids[0]]))
        for i in range(1, len(ids)):
            out += " T%d [%s]" % (i, ",".join(map(str, self.dims[ids[i - 1]:ids[i]])))
        return out

    # ---- data ------------------------------------------------------------------
    def gather(self):
        """dense numpy copy, shape = reversed dims"""
        return self.root[self.lin_index()].copy()

    def lin_index(self):
        """root-buffer index of every element, numpy shape = reversed dims"""
        lin = np.full(tuple(self.dims[::-1]), self.offs, dtype=np.int64)
        for d in range(len(self.dims)):
            shape = [1] * len(self.dims)
            shape[len(self.dims) - 1 - d] = self.dims[d]
            lin = lin + (np.arange(self.dims[d], dtype=np.int64) * self.incs[d]).reshape(shape)
        return lin

    def numpy(self):
        return self.gather()

    def badval(self):
        return self.badvalue if self.badvalue is not None else default_badval(self.dtype)

    def isbad_mask(self):
        v = self.gather()
        bv = self.badval()
        if isinstance(bv, float) and np.isnan(bv):
            return np.isnan(v)
        return v == np.asarray(bv).astype(v.dtype)

    # ---- ctypes --------------------------------------------------------------
    def to_c(self):
        c = CPdl()
        c.datatype = self.dtype
        c.state = self.state
        c.ndims = len(self.dims)
        for i, d in enumerate(self.dims):
            c.dims[i] = d
            c.incs[i] = self.incs[i]
        c.offs = self.offs
        if self.bids:
            c.nbroadcastids = len(self.bids)
            for i, v in enumerate(self.bids):
                c.broadcastids[i] = v
        c.data = self.root.ctypes.data if self.root is not None else None
        if self.badvalue is not None:
            c.has_badvalue = 1
            c.badvalue.type = self.dtype
            setattr(c.badvalue.value, "ABSUKLNPQFD"[[SB, B, S, US, L, UL, IND, ULL, LL, F, D].index(self.dtype)],
                    self.badvalue)
        return c

    def from_c(self, c, created):
        self.state = c.state
        if self.state & BADVAL and getattr(self, "parent", None) is not None:
            self.root_pdl()
        if created:
            self.dtype = c.datatype
            self.dims = [c.dims[i] for i in range(c.ndims)]
            self.incs = [c.incs[i] for i in range(c.ndims)]
            self.offs = 0
            n = max(self.nvals, 1)
            size = np.dtype(NP_TYPES[self.dtype]).itemsize
            buf = (ctypes.c_char * (n * size)).from_address(c.data)
            self.root = np.frombuffer(bytes(buf), dtype=NP_TYPES[self.dtype]).copy()
            lib().pdlc_free(c.data)


def run(op, *pdls, swap=0, nthreads=0, size_mi=1, m_size=None, dpar=None):
    """pdl_run_<op>(pdls..., swap).  m_size: the *_n_ind OtherPar; dpar: a
    double OtherPar (setvaltobad's value, setbadtoval's newval), passed as
    its IEEE bit pattern.  Raises OracleError with PDL's message."""
    cs = [p.to_c() for p in pdls]
    # the same OPdl passed twice (inplace) must be one C struct
    uniq = {}
    ptrs = (ctypes.POINTER(CPdl) * len(pdls))()
    for i, p in enumerate(pdls):
        if id(p) not in uniq:
            uniq[id(p)] = cs[i]
        ptrs[i] = ctypes.pointer(uniq[id(p)])
    created = [p.isnull() for p in pdls]
    if dpar is not None:
        other = (ctypes.c_int64 * 1)(int(np.array([dpar], np.float64).view(np.int64)[0]))
    else:
        other = (ctypes.c_int64 * 1)(swap if m_size is None else m_size)
    err = CError()
    rc = lib().pdlc_run(op.encode(), ptrs, len(pdls), other, nthreads, size_mi, ctypes.byref(err))
    if rc:
        for i, p in enumerate(pdls):
            if created[i] and cs[i].data and not (cs[i].state & NOMYDIMS):
                pass
        raise OracleError(err.message.decode())
    done = set()
    for i, p in enumerate(pdls):
        if id(p) in done:
            continue
        done.add(id(p))
        p.from_c(uniq[id(p)], created[i])
    return pdls


def run_capturing_plan(op, *pdls, swap=0, m_size=None, dpar=None):
    """run() on one thread, returning the broadcast plan PDL's redodims built
    for it (a CPlan)"""
    plan = CPlan()
    lib().pdlc_capture_plan(ctypes.byref(plan))
    try:
        run(op, *pdls, swap=swap, m_size=m_size, dpar=dpar)
    finally:
        lib().pdlc_capture_plan(None)
    return plan


def find_max_pthread(dims, target, realdims=0):
    arr = (ctypes.c_int64 * max(len(dims), 1))(*dims)
    mp, md = ctypes.c_int(0), ctypes.c_int(-1)
    lib().pdlc_find_max_pthread(arr, len(dims), realdims, target, ctypes.byref(mp), ctypes.byref(md))
    return mp.value, md.value


def random_uniform(n, seed):
    """n draws of PDL's pdl_drand from one xoshiro256+ stream (autopthread off)."""
    out = np.empty(n, dtype=np.float64)
    lib().pdlc_random_fill(out.ctypes.data, n, seed)
    return out


def rangeb(parent, ind, size=None, boundary="", back_child=None):
    """Slices.pd rangeb (the engine of range / indexND): the child as a new
    OPdl; with back_child, its BackCode writes that child into parent.
    size: None (undefined) or a list of sizes; ind must be indx."""
    pc, ic = parent.to_c(), ind.to_c()
    cc = back_child.to_c() if back_child is not None else CPdl()
    sz = None if size is None else np.ascontiguousarray(np.atleast_1d(np.asarray(size, dtype=np.int64)))
    err = CError()
    rc = lib().pdlc_rangeb(ctypes.byref(pc), ctypes.byref(ic), sz.ctypes.data if sz is not None and sz.size else None,
                           0 if sz is None else sz.size, (boundary or "").encode(), ctypes.byref(cc),
                           1 if back_child is not None else 0, ctypes.byref(err))
    if rc:
        raise OracleError(err.message.decode())
    if back_child is not None:
        return back_child
    child = OPdl.null()
    child.from_c(cc, True)
    return child


def random_from(n, seed, skip=0):
    """draws [skip, skip + n) of the single stream seeded with `seed`"""
    out = np.empty(n, dtype=np.float64)
    lib().pdlc_random_fill_from(out.ctypes.data, n, seed, skip)
    return out


def rand_state_after(seed, skip):
    """the xoshiro256+ state after `skip` draws of the stream seeded with `seed`"""
    out = np.empty(4, dtype=np.uint64)
    lib().pdlc_rand_state_after(out.ctypes.data, seed, skip)
    return out


def ndtri(y):
    """PDL::Math::ndtri (Cephes) of a float64 array"""
    y = np.ascontiguousarray(y, dtype=np.float64)
    x = np.empty_like(y)
    lib().pdlc_ndtri_n(y.ctypes.data, x.ctypes.data, y.size)
    return x


def grandom(dtype, n, seed, skip=0):
    """PDL's grandom(type, n) after srandom(seed) and `skip` earlier draws:
    ndtri of the randsym draws rounded to the type, rounded back
    (Primitive.pd:2256-2261; Math.pd:346-356)"""
    u = random_from(n, seed, skip)
    assert not np.any(u == 0.0)          # randsym would draw again
    return ndtri(u.astype(dtype).astype(np.float64)).astype(dtype)


def truth_matmult_ld(a, b):
    """row-major A(h,t) @ B(t,w) accumulated in long double"""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    h, t = a.shape
    w = b.shape[1]
    c = np.empty((h, w), dtype=np.float64)
    lib().pdlc_truth_matmult_ld(a.ctypes.data, b.ctypes.data, c.ctypes.data, h, t, w)
    return c
