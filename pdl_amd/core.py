"""Host-side mirror of PDL's ndarray API for the broadcast-loop path.

An `ndarray` is PDL's `struct pdl` as the broadcast engine sees it: a root
buffer in HBM, dims (dim 0 fastest), the composed virtual-affine increments
and offset of a view, the bad-value state and value.  Views (xchg, mv, dummy,
slice, transpose) are zero-copy like Slices.pd's affine transforms.  Every op
goes through the C ABI (include/pdlhip.h) to the HIP kernels; data reaches
the host only at PDL's physical boundaries (`numpy()`, `at()`, `sclr()`,
printing), as Core.xs's make_physical / get_dataref do.

Perl-level wrapper logic restated here (cited where defined):
  matmult PMCode        lib/PDL/Primitive.pd:197-209
  scalar literal typing lib/PDL/Core/pdlperl.h:185-205
  whole-array reducers  lib/PDL/Ufunc.pd:618-663 (`$x->flat->op`)
  operator overloads    lib/PDL/PP.pm:1420-1484 (swap for reversed operands)
"""
import ctypes
import weakref

import numpy as np

from . import _lib as L
from ._lib import PDLError

TYPE_NAMES = ["sbyte", "byte", "short", "ushort", "long", "ulong", "indx", "ulonglong", "longlong",
              "float", "double", "ldouble", "cfloat", "cdouble", "cldouble"]
# the types' ctype names without PDL_ (Types.pm:31-200, shortctype at 265)
SHORT_CTYPES = ["SByte", "Byte", "Short", "Ushort", "Long", "ULong", "Indx", "ULongLong", "LongLong",
                "Float", "Double", "LDouble", "CFloat", "CDouble", "CLDouble"]
NP_TYPES = {L.SB: np.int8, L.B: np.uint8, L.S: np.int16, L.US: np.uint16, L.L: np.int32, L.UL: np.uint32,
            L.IND: np.int64, L.ULL: np.uint64, L.LL: np.int64, L.F: np.float32, L.D: np.float64,
            L.CF: np.complex64, L.CD: np.complex128}
_FROM_NP = {np.dtype(np.int8): L.SB, np.dtype(np.uint8): L.B, np.dtype(np.int16): L.S,
            np.dtype(np.uint16): L.US, np.dtype(np.int32): L.L, np.dtype(np.uint32): L.UL,
            np.dtype(np.int64): L.LL, np.dtype(np.uint64): L.ULL, np.dtype(np.float32): L.F,
            np.dtype(np.float64): L.D, np.dtype(np.bool_): L.B, np.dtype(np.complex64): L.CF,
            np.dtype(np.complex128): L.CD}
_VALFIELD = {L.SB: "A", L.B: "B", L.S: "S", L.US: "U", L.L: "L", L.UL: "K", L.IND: "N", L.ULL: "P",
             L.LL: "Q", L.F: "F", L.D: "D"}

_inited = {"device": None, "plan_only": False}


def init(device=0):
    """Bind the library to one GPU (one process per GPU)."""
    if _inited["device"] is None:
        L.check(L.lib().pdlhip_init(device))
        _inited["device"] = device


def _need():
    if _inited["device"] is None and not _inited["plan_only"]:
        init(0)


def set_plan_only(flag):
    """Dry-run mode of the C ABI: ops plan, type-coerce and create output
    dims/types but allocate and launch nothing (no device needed)."""
    _inited["plan_only"] = bool(flag)
    L.check(L.lib().pdlhip_set_plan_only(1 if flag else 0))


def shell(dtype, dims, state=0, incs=None, offs=0):
    """an ndarray header without data (for plan-only use)"""
    p = ndarray()
    p.dtype = typecode(dtype)
    p.dims = list(dims)
    p.incs = list(incs) if incs is not None else _dense_incs(p.dims)
    p.offs = offs
    p.state = state
    return p


def typecode(t):
    if isinstance(t, int):
        return t
    if isinstance(t, str):
        return TYPE_NAMES.index(t)
    return _FROM_NP[np.dtype(t)]


class _Buf:
    """An HBM allocation owned by this process (released to the cache on GC),
    with the state PDL keeps for an ndarray's data (pdlapi.c:871-945):

    version       bumped whenever an op writes the buffer (pdl_changed)
    host          a host mirror of the whole buffer, filled by one download
                  when an element is read (the Core.xs sync points: at, set,
                  sclr), current while host_version == version
    dev_stale     the mirror holds writes (set) not yet in HBM; they go up
                  before any device op uses the buffer
    flow          the dataflow of a child made by index / index1d / index2d
                  (PDL keeps that trans): recomputed from its parents when
                  they changed, written back into them when it is written
    """
    __slots__ = ("ptr", "nbytes", "version", "host", "host_version", "host_epoch", "dev_stale", "flow",
                 "__weakref__")

    def __init__(self, ptr, nbytes):
        self.ptr = ptr
        self.nbytes = nbytes
        self.version = 0
        self.host = None
        self.host_version = -1
        self.host_epoch = -1
        self.dev_stale = False
        self.flow = None

    def __del__(self):
        if self.ptr and L._lib is not None and _inited["device"] is not None:
            try:
                L._lib.pdlhip_free(self.ptr)
            except Exception:
                pass
            self.ptr = None


def _alloc(nbytes):
    _need()
    p = ctypes.c_void_p()
    L.check(L.lib().pdlhip_alloc(max(int(nbytes), 1), ctypes.byref(p)))
    return _Buf(p.value, nbytes)


# ---------------------------------------------------------------- mirrors and dataflow
_MIRROR_MAX = 64 << 20          # buffers up to this size get a whole host mirror
_EPOCH = [0]                    # bumped by launches that do not say which buffers they write
_STALE = weakref.WeakSet()      # buffers whose mirror holds writes not yet uploaded


def _launched():
    """a list / chain / graph launch wrote buffers this module does not track:
    every host mirror is out of date"""
    _EPOCH[0] += 1


def _flush(buf):
    if buf is not None and buf.dev_stale:
        if buf.nbytes:
            L.check(L.lib().pdlhip_upload(buf.ptr, buf.host.ctypes.data, buf.nbytes))
        buf.dev_stale = False
        _STALE.discard(buf)


def _flush_all():
    for b in list(_STALE):
        _flush(b)


def _mirror(buf):
    """the buffer's host mirror, current (one download when it is not)"""
    if buf.dev_stale:
        return buf.host
    if buf.host is None or buf.host_version != buf.version or buf.host_epoch != _EPOCH[0]:
        if buf.host is None:
            buf.host = np.empty(buf.nbytes, np.uint8)
        if buf.nbytes:
            L.check(L.lib().pdlhip_download(buf.host.ctypes.data, buf.ptr, buf.nbytes))
        buf.host_version, buf.host_epoch = buf.version, _EPOCH[0]
    return buf.host


class _Flow:
    """a child's dataflow: its parents (the source first, then the index
    ndarrays), the child over its whole buffer, the parents' buffer versions
    its data reflects, and the two calls PDL's trans would make: fwd
    (readdata: compute the child again) and back (writebackdata: the child
    into the source)"""
    __slots__ = ("parents", "child", "seen", "fwd", "back")

    def __init__(self, parents, child, fwd, back):
        self.parents, self.child, self.fwd, self.back = list(parents), child, fwd, back
        self.seen = self.versions()

    def versions(self):
        return tuple((p._buf.version if p._buf is not None else 0) for p in self.parents)


_OPINFO = {}


def _op_info(op):
    if op not in _OPINFO:
        n, k = ctypes.c_int(0), ctypes.c_int(0)
        L.check(L.lib().pdlhip_op_info(op.encode(), ctypes.byref(n), ctypes.byref(k)))
        _OPINFO[op] = (n.value, k.value)
    return _OPINFO[op]


def _refresh(buf):
    """before a read: a child whose parents changed is computed again
    (PDL's make_physical of a dataflow child)"""
    f = buf.flow if buf is not None else None
    if f is None:
        return
    for p in f.parents:
        _refresh(p._buf)
        _flush(p._buf)
    if f.versions() != f.seen:
        f.fwd(f)
        buf.version += 1
        f.seen = f.versions()


def _written(buf):
    """after a write: the buffer's version moves on and a dataflow child
    writes itself back into its parent (writebackdata)"""
    if buf is None:
        return
    buf.version += 1
    _writeback(buf)


def _writeback(buf):
    f = buf.flow
    if f is None:
        return
    _flush(buf)
    for p in f.parents:
        _flush(p._buf)
    f.back(f)
    _written(f.parents[0]._buf)
    f.seen = f.versions()


def _index_flow(op):
    """the forward and back calls of an index-family child"""
    def fwd(f):
        _call(op, f.parents + [f.child])

    def back(f):
        cs = [p._to_c() for p in f.parents + [f.child]]
        arr = (L.PP * len(cs))(*[ctypes.pointer(c) for c in cs])
        L.check(L.lib().pdlhip_writebackdata(op.encode(), arr, len(cs)))
    return fwd, back


def _range_call(entry, parent, ind, size, boundary, child):
    """pdlhip_run_rangeb / pdlhip_writeback_rangeb; size None (undefined) or
    a list of ints; returns the child's C struct"""
    cp, ci, cc = parent._to_c(), ind._to_c(), child._to_c()
    if size is None:
        sz, n = None, -1
    else:
        arr = np.ascontiguousarray(np.asarray(size, dtype=np.int64).reshape(-1))
        sz, n = (ctypes.c_int64 * max(arr.size, 1))(*arr.tolist()), arr.size
    L.check(getattr(L.lib(), entry)(ctypes.byref(cp), ctypes.byref(ci), sz, n,
                                    (boundary or "").encode(), ctypes.byref(cc)))
    return cc


def _dense_incs(dims):
    incs, inc = [], 1
    for d in dims:
        incs.append(inc)
        inc *= d
    return incs


class ndarray:
    """PDL ndarray with data in HBM."""
    __array_priority__ = 1000

    def __init__(self):
        self._buf = None
        self.dtype = -1
        self.dims = []
        self.incs = []
        self.offs = 0
        self.state = L.NOMYDIMS
        self.badvalue_ = None
        self._parent = None
        self.bids = None          # explicit broadcast ids (broadcastI); None: a plain ndarray

    __hash__ = object.__hash__

    # ------------------------------------------------------------ construction
    @classmethod
    def from_numpy(cls, arr, dtype=None, badflag=False, badvalue=None):
        """Upload a numpy array (C order: numpy shape = reversed PDL dims)."""
        arr = np.asarray(arr)
        t = typecode(dtype) if dtype is not None else _FROM_NP.get(arr.dtype, L.D)
        host = np.array(arr, dtype=NP_TYPES[t], order="C", copy=True)
        p = cls()
        p.dtype = t
        p.dims = list(host.shape[::-1])
        p.incs = _dense_incs(p.dims)
        p.state = L.BADVAL if badflag else 0
        p.badvalue_ = badvalue
        p._buf = _alloc(host.nbytes)
        if host.nbytes:
            L.check(L.lib().pdlhip_upload(p._buf.ptr, host.ctypes.data, host.nbytes))
        return p

    @classmethod
    def null(cls):
        return cls()

    @classmethod
    def new_dense(cls, dtype, dims):
        p = cls()
        p.dtype = typecode(dtype)
        p.dims = list(dims)
        p.incs = _dense_incs(p.dims)
        p.state = 0
        p._buf = _alloc(p.nelem * np.dtype(NP_TYPES[p.dtype]).itemsize)
        return p

    # ------------------------------------------------------------ info
    @property
    def ndims(self):
        return len(self.dims)

    @property
    def nelem(self):
        n = 1
        for d in self.dims:
            n *= d
        return n

    def isnull(self):
        return bool(self.state & L.NOMYDIMS)

    def isempty(self):
        return self.nelem == 0

    def type(self):
        return TYPE_NAMES[self.dtype]

    def dim(self, i):
        return self.dims[i] if i < len(self.dims) else 1

    def getndims(self):
        return len(self.dims)

    def badflag(self, flag=None):
        if flag is not None:
            if flag:
                self.state |= L.BADVAL
            else:
                self.state &= ~L.BADVAL
        return bool(self.state & L.BADVAL)

    def badvalue(self, v=None):
        if v is not None:
            self.badvalue_ = v
        if self.badvalue_ is not None:
            return self.badvalue_
        return default_badvalue(self.dtype)

    # ------------------------------------------------------------ views
    def _view(self, dims, incs, offs):
        v = ndarray()
        v._buf = self._buf
        v.dtype = self.dtype
        v.dims, v.incs, v.offs = list(dims), list(incs), offs
        v.state = (self.state & L.BADVAL) | L.VAFFINE
        v.badvalue_ = self.badvalue_
        v._parent = self
        return v

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

    def transpose(self):
        return self.dummy(0) if self.ndims < 2 else self.xchg(0, 1)

    # ------------------------------------------------------------ explicit broadcasting
    def broadcastids(self):
        """pdl->broadcastids (Core.xs:1278-1290): [ndims] for a plain ndarray"""
        return list(self.bids) if self.bids else [self.ndims]

    def broadcastI(self, bid, whichdims):
        """Slices.pd:1819-1898: a view whose dims `whichdims` (in that order)
        form explicit broadcast id `bid`, placed after the dims of the ids
        below it; the other dims keep their order.  Every view op other than
        broadcastI and unbroadcast gives a plain ndarray again (its dims are
        set with one broadcast id, pdlapi.c:613)."""
        wd = [int(x) for x in whichdims]
        for i in range(len(wd)):
            for j in range(i + 1, len(wd)):
                if wd[i] == wd[j] and wd[i] != -1:
                    raise PDLError(L.EUSERERROR, "Error in broadcastI:duplicate arg %d %d %d" % (i, j, wd[i]))
        nd = self.ndims
        if any(w == -1 for w in wd):
            raise PDLError(L.EUSERERROR, "broadcastI: a whichdims entry of -1 is not supported")
        for w in wd:
            if not 0 <= w < nd:
                raise PDLError(L.EUSERERROR, "broadcastI: dim %d out of range (ndarray has %d dims)" % (w, nd))
        pb = self.broadcastids()
        pn = len(pb)
        nreal, count = len(wd), len(wd)
        dims, incs = [0] * nd, [0] * nd
        nthc = 0
        for i in range(nd):
            if pn > bid >= 0 and i == pb[bid]:
                nthc += count
            if i in wd:
                continue
            dims[nthc], incs[nthc] = self.dims[i], self.incs[i]
            nthc += 1
        base = pb[bid] if pn > bid >= 0 else nd
        for i, w in enumerate(wd):
            dims[i + base - nreal], incs[i + base - nreal] = self.dims[w], self.incs[w]
        # the moved dims land where the others left room only when none of
        # them lies past broadcastids[bid]; PDL leaves a hole otherwise
        if pn > bid >= 0 and base < nd and any(w >= base for w in wd):
            raise PDLError(L.EUSERERROR, "broadcastI: dims %s lie past broadcast id %d's start %d" % (wd, bid, base))
        cn = max(bid + 1, pn)
        if cn > L.MAXBIDS:
            raise PDLError(L.EUSERERROR, "broadcastI: at most %d broadcast ids" % L.MAXBIDS)
        bids = [(pb[i] if pn > i else nd) + (-nreal if i <= bid else count - nreal) for i in range(cn - 1)]
        v = self._view(dims, incs, self.offs)
        v.bids = bids + [nd]
        return v

    def broadcast(self, *dims):
        """explicit broadcasting over dims with broadcast id 1 (Core.pm:1995-1998)"""
        return self.broadcastI(1, dims)

    broadcast1 = broadcast

    def broadcast2(self, *dims):
        """Core.pm:2035-2038"""
        return self.broadcastI(2, dims)

    def broadcast3(self, *dims):
        """Core.pm:2055-2058"""
        return self.broadcastI(3, dims)

    def unbroadcast(self, atind):
        """Slices.pd:1900-1925: the broadcast-id dims (from broadcastids[0]
        on) move back to position atind; a plain ndarray comes out"""
        nd, b0 = self.ndims, self.broadcastids()[0]
        dims, incs = [0] * nd, [0] * nd
        for i in range(nd):
            if i < atind:
                corc = i
            elif i < b0:
                corc = i + nd - b0
            else:
                corc = i - b0 + atind
            dims[corc], incs[corc] = self.dims[i], self.incs[i]
        return self._view(dims, incs, self.offs)

    def dimstr(self):
        """Core.pm:2105-2117: 'D [..]' then ' Tk [..]' per explicit id"""
        ids = self.broadcastids()
        out = "D [" + ",".join(str(d) for d in self.dims[:ids[0]]) + "]"
        for i in range(1, len(ids)):
            out += " T%d [" % i + ",".join(str(d) for d in self.dims[ids[i - 1]:ids[i]]) + "]"
        return out

    def info(self):
        """PDL::info with the default format '%C: %T %D' (Core.pm:56, 2226-2241)"""
        if self.isnull():
            return "PDL->null"
        return "PDL: %s %s" % (SHORT_CTYPES[self.dtype], self.dimstr())

    def slice(self, *specs):
        """per dim: None (all), int k ('(k)', drops the dim) or (start, stop, step) with an
        inclusive stop as in PDL's 'a:b:s'"""
        dims, incs, offs = [], [], self.offs
        for d in range(len(self.dims)):
            sp = specs[d] if d < len(specs) else None
            if sp is None:
                dims.append(self.dims[d])
                incs.append(self.incs[d])
            elif isinstance(sp, (int, np.integer)):
                k = int(sp) if sp >= 0 else self.dims[d] + int(sp)
                if not 0 <= k < self.dims[d]:
                    raise PDLError(L.EUSERERROR, "slice starts out of bounds in pos %d (start is %d; end is %d; "
                                   "inc is 0; source dim %d runs 0 to %d)" % (d, k, k, d, self.dims[d] - 1))
                offs += k * self.incs[d]
            else:
                # bounds and regularisation of Slices.pd:2367-2393
                a, b, s = sp
                size = self.dims[d]
                a = a if a >= 0 else size + a
                if not 0 <= a < size:
                    raise PDLError(L.EUSERERROR, "slice starts out of bounds in pos %d (start is %d; end is %d; "
                                   "inc is %d; source dim %d runs 0 to %d)" % (d, a, b, s, d, size - 1))
                b = b if b >= 0 else size + b
                if not 0 <= b < size:
                    raise PDLError(L.EUSERERROR, "slice ends out of bounds in pos %d (end is %d; source dim %d "
                                   "runs 0 to %d)" % (d, b, d, size - 1))
                s = s if s else (1 if a <= b else -1)
                n = (b - a + s) // s
                dims.append(max(n, 0))
                incs.append(self.incs[d] * s)
                offs += a * self.incs[d]
        return self._view(dims, incs, offs)

    def copy(self):
        out = ndarray()
        _run("assgn", [self, out])
        return out

    def assign(self, other):
        """PDL's `.=`: other's values into this ndarray (through a view into
        its parent; a dataflow child flows back into its own)"""
        _run("assgn", [topdl(other), self])
        return self

    def _alias(self):
        v = ndarray()
        v._buf, v.dtype, v.dims, v.incs, v.offs = self._buf, self.dtype, list(self.dims), list(self.incs), self.offs
        v.state, v.badvalue_ = self.state, self.badvalue_
        return v

    def _flowing(self, op, parents):
        c = ndarray()
        _run(op, parents + [c])
        c._buf.flow = _Flow(parents, c._alias(), *_index_flow(op))
        return c

    def rangeb(self, ind, size=None, boundary=""):
        """Slices.pd rangeb (Slices.pd:561-1015): the chunks of this ndarray at
        the coordinates in ind (dim 0: a coordinate vector), each of `size`
        (None: single elements), with per-dim boundary letters (f forbid, t
        truncate, e extend, p periodic, m mirror); a child with dataflow"""
        ind = topdl(ind)
        for p in (self, ind):
            _refresh(p._buf)
            _flush(p._buf)
        c = ndarray()
        cc = _range_call("pdlhip_run_rangeb", self, ind, size, boundary, c)
        c._from_c(cc, True)
        _written(c._buf)

        def fwd(f):
            _range_call("pdlhip_run_rangeb", f.parents[0], f.parents[1], size, boundary, f.child)

        def back(f):
            _range_call("pdlhip_writeback_rangeb", f.parents[0], f.parents[1], size, boundary, f.child)
        if c._buf is not None:
            c._buf.flow = _Flow([self, ind], c._alias(), fwd, back)
        return c

    def range(self, ind, size=None, boundary=None):
        """PDL::range (Slices.pd:272-304): the index as indx (a 0-dim index
        gets a dummy dim 0), an empty index gives an empty result of the
        right dims, boundary lists pack to their first letters"""
        index = ind if isinstance(ind, ndarray) and ind.dtype == L.IND else topdl(ind).convert(L.IND) \
            if isinstance(ind, ndarray) else ndarray.from_numpy(np.asarray(ind, dtype=np.int64), L.IND)
        sz = None if size is None else [int(v) for v in np.asarray(size.numpy() if isinstance(size, ndarray)
                                                                    else size).reshape(-1)]
        if index.nelem == 0:
            sdims = list(self.dims)
            n0 = index.dims[0] if index.ndims else 0
            del sdims[:n0 + (1 if n0 == 0 else 0)]
            if sz is not None:
                sdims = sz + sdims
            idims = [0] * (index.ndims - 1) or [0]
            return ndarray.new_dense(self.dtype, idims + sdims)
        if index.ndims == 0:
            index = index.dummy(0, 1)
        bound = boundary
        if isinstance(boundary, (list, tuple)):
            bound = ""
            for el in boundary:
                el = str(el)
                if not el or el[0] not in "0123fFtTeEpPmM":
                    raise PDLError(L.EUSERERROR, "Illegal boundary value '%s' in range" % el)
                bound += el[0]
        elif boundary is not None:
            import re
            if not re.match(r"^[0123ftepx]+$", boundary) and re.match(r"^([0123ftepx])", boundary, re.I):
                bound = boundary[0]
        return self.rangeb(index, sz, bound or "")

    def indexND(self, ind, boundary=None):
        """Slices.pd:265-268: range with no size"""
        return self.range(ind, None, boundary)

    # Slices.pd:142-220 -- non-affine slices: a new ndarray with dataflow
    # both ways (recomputed when a parent changes, written back when written)
    def index(self, ind):
        return self._flowing("index", [self, topdl(ind)])

    def index1d(self, ind):
        return self._flowing("index1d", [self, topdl(ind)])

    def index2d(self, inda, indb):
        return self._flowing("index2d", [self, topdl(inda), topdl(indb)])

    def dice_axis(self, axis, idx):
        """Slices.pd:2077-2082: mv(axis,0)->index1d(idx)->mv(0,axis)"""
        ix = topdl(idx)
        if ix.ndims > 1:
            raise PDLError(L.EUSERERROR, "dice_axis: index must be <=1D")
        axis = axis + self.ndims if axis < 0 else axis
        return self.mv(axis, 0).index1d(ix).mv(0, axis)

    def dice(self, *dim_indexes):
        """Slices.pd:2001-2026: dice_axis along each dim in turn ('X' keeps a dim)"""
        if len(dim_indexes) > self.ndims:
            raise PDLError(L.EUSERERROR, "PDL::dice: Number of index arrays (%d) not equal to the dimensions of the "
                           "PDL (%d" % (len(dim_indexes), self.ndims))
        out = self
        for k, ix in enumerate(dim_indexes):
            if isinstance(ix, str) and ix == "X":
                continue
            out = out.dice_axis(k, ix)
        return out

    def sever(self):
        """detach a dataflow child from its parents (pdlapi.c:1043-1050)"""
        if self._buf is not None and self._buf.flow is not None:
            _refresh(self._buf)
            self._buf.flow = None
        return self

    def make_physical(self):
        return self

    def clump(self, n):
        """Slices.pd:1373-1400 (_clump_int): the first n dims as one (n = -1:
        up to broadcastids[0]; n < -1 counts from the end; n past ndims
        clumps them all), a child with dataflow.  Dims that sit contiguously
        in memory clump into a view of the same data; others into a dense
        copy that is refreshed from the parent and written back into it."""
        nd = self.ndims
        if n > nd:
            n = nd
        if n < -1:
            n = nd + n + 1
        nrem = self.broadcastids()[0] if n == -1 else n
        d1 = 1
        for d in self.dims[:nrem]:
            d1 *= d
        dims = [d1] + list(self.dims[nrem:])
        # broadcast ids move with the dims ($SETDELTABROADCASTIDS(1-nrem))
        bids = [b + 1 - nrem for b in self.bids] if self.bids else None
        # the clumped dims merge into one stride when each non-unit dim
        # starts where the one before it ends
        nz = [(self.dims[i], self.incs[i]) for i in range(nrem) if self.dims[i] != 1]
        if d1 == 0 or all(nz[k + 1][1] == nz[k][1] * nz[k][0] for k in range(len(nz) - 1)):
            v = self._view(dims, [nz[0][1] if nz else 1] + list(self.incs[nrem:]), self.offs)
            v.bids = bids
            return v
        # element i of the child is element i of the parent in PDL's order
        # (EquivCPOffsCode): the copies run over plain views, without the
        # parent's broadcast ids
        c = ndarray.new_dense(self.dtype, self.dims)
        c.state |= self.state & L.BADVAL
        c.badvalue_ = self.badvalue_
        _run("assgn", [self._view(self.dims, self.incs, self.offs), c])
        c.dims, c.incs, c.bids = dims, _dense_incs(dims), bids
        pdims = list(self.dims)

        def as_parent(child):
            return child._view(pdims, _dense_incs(pdims), 0)

        def plain(p):
            return p._view(p.dims, p.incs, p.offs)

        def fwd(f):
            _call("assgn", [plain(f.parents[0]), as_parent(f.child)])

        def back(f):
            _call("assgn", [as_parent(f.child), plain(f.parents[0])])
        c._buf.flow = _Flow([self], c._alias(), fwd, back)
        return c

    def flat(self):
        """clump(-1) (Core.pm flat): all dims as one, with dataflow"""
        return self.clump(-1)

    def clump_all(self):
        return self.flat()

    def convert(self, t):
        t = typecode(t)
        if t == self.dtype:
            return self
        out = ndarray.new_dense(t, self.dims)
        out.state |= self.state & L.BADVAL
        _run("assgn_convert", [self, out])
        return out

    # ------------------------------------------------------------ host mirror
    def _is_dense(self):
        return self.offs == 0 and self.incs == _dense_incs(self.dims) and not (self.state & L.VAFFINE)

    def numpy(self):
        """make_physical + copy to host; numpy shape = reversed dims"""
        if self.isnull():
            raise PDLError(L.EUSERERROR, "numpy(): ndarray is null")
        if self._buf is not None:
            _refresh(self._buf)
            _flush(self._buf)
        src = self if self._is_dense() else self.copy()
        out = np.empty(src.dims[::-1], dtype=NP_TYPES[src.dtype])
        if out.nbytes:
            L.check(L.lib().pdlhip_download(out.ctypes.data, src._buf.ptr, out.nbytes))
        return out

    def at(self, *idx):
        """one element (a Core.xs sync point): read from the buffer's host
        mirror, which one download fills while the device has not written
        the buffer since (buffers above 64 MiB: the element alone)"""
        if len(idx) != self.ndims or any(not 0 <= i < d for i, d in zip(idx, self.dims)):
            raise PDLError(L.EUSERERROR, "Position %s out of range" % (list(idx),))
        off = self.offs + sum(i * c for i, c in zip(idx, self.incs))
        dt = np.dtype(NP_TYPES[self.dtype])
        _refresh(self._buf)
        if self._buf.nbytes <= _MIRROR_MAX:
            return np.frombuffer(_mirror(self._buf), dtype=dt, count=1, offset=off * dt.itemsize)[0]
        _flush(self._buf)
        out = np.empty(1, dtype=dt)
        L.check(L.lib().pdlhip_download(out.ctypes.data, self._buf.ptr + off * dt.itemsize, dt.itemsize))
        return out[0]

    def set(self, *args):
        """PDL's set(i, j, ..., value): one element written through the host
        mirror (uploaded before the next device op reads the buffer); a
        dataflow child writes itself back into its parent at once"""
        *idx, v = args
        if len(idx) != self.ndims or any(not 0 <= i < d for i, d in zip(idx, self.dims)):
            raise PDLError(L.EUSERERROR, "Position %s out of range" % (list(idx),))
        off = self.offs + sum(i * c for i, c in zip(idx, self.incs))
        dt = np.dtype(NP_TYPES[self.dtype])
        val = np.array([v]).astype(dt)
        buf = self._buf
        _refresh(buf)
        if buf.nbytes <= _MIRROR_MAX:
            h = _mirror(buf)
            h[off * dt.itemsize:(off + 1) * dt.itemsize] = val.view(np.uint8)
            buf.version += 1
            buf.host_version, buf.host_epoch = buf.version, _EPOCH[0]
            buf.dev_stale = True
            _STALE.add(buf)
        else:
            L.check(L.lib().pdlhip_upload(buf.ptr + off * dt.itemsize, val.ctypes.data, dt.itemsize))
            buf.version += 1
        _writeback(buf)
        return self

    def sclr(self):
        if self.nelem < 1:
            raise PDLError(L.EUSERERROR, "sclr: empty ndarray")
        return self.numpy().reshape(-1)[0]

    def __repr__(self):
        if self.isnull():
            return "ndarray(null)"
        return "ndarray(%s %s: %s)" % (self.type(), self.dims, self.numpy())

    # ------------------------------------------------------------ ctypes
    def _to_c(self):
        c = L.Pdl()
        c.datatype = self.dtype
        c.state = self.state
        c.ndims = len(self.dims)
        if c.ndims > L.MAXDIMS:
            raise PDLError(L.EUSERERROR, "too many dims")
        for i, d in enumerate(self.dims):
            c.dims[i] = d
            c.incs[i] = self.incs[i]
        c.offs = self.offs
        if self.bids:
            c.nbroadcastids = len(self.bids)
            for i, v in enumerate(self.bids):
                c.broadcastids[i] = v
        c.data = self._buf.ptr if self._buf is not None else None
        if self.badvalue_ is not None:
            c.has_badvalue = 1
            c.badvalue.type = self.dtype
            setattr(c.badvalue.value, _VALFIELD[self.dtype], self.badvalue_)
        return c

    def _from_c(self, c, created):
        self.state = c.state
        if created:
            self.dtype = c.datatype
            self.dims = [c.dims[i] for i in range(c.ndims)]
            self.incs = [c.incs[i] for i in range(c.ndims)]
            self.offs = 0
            self._buf = _Buf(c.data, self.nelem * np.dtype(NP_TYPES[self.dtype]).itemsize) if c.data else None
        p = self._parent
        while p is not None and (self.state & L.BADVAL):
            p.state |= L.BADVAL
            p = p._parent

    # ------------------------------------------------------------ operators
    def _bin(self, op, other, swap):
        out = ndarray()
        _run(op, [self, topdl(other), out], swap=swap)
        return out

    def _mut(self, op, other):
        _run(op, [self, topdl(other), self], swap=0)
        return self

    def __add__(self, o): return self._bin("plus", o, 0)
    def __radd__(self, o): return self._bin("plus", o, 1)
    def __sub__(self, o): return self._bin("minus", o, 0)
    def __rsub__(self, o): return self._bin("minus", o, 1)
    def __mul__(self, o): return self._bin("mult", o, 0)
    def __rmul__(self, o): return self._bin("mult", o, 1)
    def __truediv__(self, o): return self._bin("divide", o, 0)
    def __rtruediv__(self, o): return self._bin("divide", o, 1)
    def __gt__(self, o): return self._bin("gt", o, 0)
    def __lt__(self, o): return self._bin("lt", o, 0)
    def __ge__(self, o): return self._bin("ge", o, 0)
    def __le__(self, o): return self._bin("le", o, 0)
    def __eq__(self, o): return self._bin("eq", o, 0)
    def __ne__(self, o): return self._bin("ne", o, 0)
    def __lshift__(self, o): return self._bin("shiftleft", o, 0)
    def __rlshift__(self, o): return self._bin("shiftleft", o, 1)
    def __rshift__(self, o): return self._bin("shiftright", o, 0)
    def __rrshift__(self, o): return self._bin("shiftright", o, 1)
    def __or__(self, o): return self._bin("or2", o, 0)
    def __ror__(self, o): return self._bin("or2", o, 1)
    def __and__(self, o): return self._bin("and2", o, 0)
    def __rand__(self, o): return self._bin("and2", o, 1)
    def __xor__(self, o): return self._bin("xor", o, 0)
    def __rxor__(self, o): return self._bin("xor", o, 1)
    def __pow__(self, o): return self._bin("power", o, 0)
    def __rpow__(self, o): return self._bin("power", o, 1)
    def __mod__(self, o): return self._bin("modulo", o, 0)
    def __rmod__(self, o): return self._bin("modulo", o, 1)
    def __matmul__(self, o): return matmult(self, o)
    def __rmatmul__(self, o): return matmult(topdl(o), self)
    def __iadd__(self, o): return self._mut("plus", o)
    def __isub__(self, o): return self._mut("minus", o)
    def __imul__(self, o): return self._mut("mult", o)
    def __itruediv__(self, o): return self._mut("divide", o)
    def __ipow__(self, o): return self._mut("power", o)
    def __imod__(self, o): return self._mut("modulo", o)
    def __ior__(self, o): return self._mut("or2", o)
    def __iand__(self, o): return self._mut("and2", o)
    def __ixor__(self, o): return self._mut("xor", o)
    def __ilshift__(self, o): return self._mut("shiftleft", o)
    def __irshift__(self, o): return self._mut("shiftright", o)

    def spaceship(self, o, swap=0): return self._bin("spaceship", o, swap)
    def atan2(self, o, swap=0): return self._bin("atan2", o, swap)

    # ------------------------------------------------------------ Ufunc
    def _red(self, op, out=None):
        out = ndarray() if out is None else out
        _run(op, [self, out])
        return out

    # random / randsym / grandom of an ndarray: a new one of its dims and type
    def random(self): return random(*self.dims, dtype=self.dtype)
    def randsym(self): return randsym(*self.dims, dtype=self.dtype)
    def grandom(self): return grandom(*self.dims, dtype=self.dtype)

    def sumover(self, out=None): return self._red("sumover", out)
    def prodover(self, out=None): return self._red("prodover", out)
    def dsumover(self, out=None): return self._red("dsumover", out)
    def dprodover(self, out=None): return self._red("dprodover", out)
    def cumusumover(self, out=None): return self._red("cumusumover", out)
    def cumuprodover(self, out=None): return self._red("cumuprodover", out)
    def dcumusumover(self, out=None): return self._red("dcumusumover", out)
    def dcumuprodover(self, out=None): return self._red("dcumuprodover", out)
    def average(self, out=None): return self._red("average", out)
    avgover = average
    def daverage(self, out=None): return self._red("daverage", out)
    def minimum(self, out=None): return self._red("minimum", out)
    minover = minimum
    def maximum(self, out=None): return self._red("maximum", out)
    maxover = maximum
    def minimum_ind(self, out=None): return self._red("minimum_ind", out)
    def maximum_ind(self, out=None): return self._red("maximum_ind", out)

    def minimum_n_ind(self, m, out=None):
        out = ndarray() if out is None else out
        _run("minimum_n_ind", [self, out], m_size=m)
        return out

    def maximum_n_ind(self, m, out=None):
        out = ndarray() if out is None else out
        _run("maximum_n_ind", [self, out], m_size=m)
        return out

    def minmaximum(self):
        outs = [ndarray() for _ in range(4)]
        _run("minmaximum", [self] + outs)
        return tuple(outs)

    # ------------------------------------------------------------ Ops.pd ufuncs
    def bitnot(self, out=None): return self._red("bitnot", out)
    def __invert__(self): return self._red("bitnot")
    def sqrt(self, out=None): return self._red("sqrt", out)
    def sin(self, out=None): return self._red("sin", out)
    def cos(self, out=None): return self._red("cos", out)
    def not_(self, out=None): return self._red("not", out)
    def exp(self, out=None): return self._red("exp", out)
    def log(self, out=None): return self._red("log", out)
    def log10(self, out=None): return self._red("log10", out)
    def abs(self, out=None): return self._red("abs", out)
    def __abs__(self): return self._red("abs")

    # ------------------------------------------------------------ Bad.pd
    def isbad(self, out=None): return self._red("isbad", out)
    def isgood(self, out=None): return self._red("isgood", out)
    def setnantobad(self, out=None): return self._red("setnantobad", out)
    def setinftobad(self, out=None): return self._red("setinftobad", out)
    def setnonfinitetobad(self, out=None): return self._red("setnonfinitetobad", out)
    def setbadtonan(self, out=None): return self._red("setbadtonan", out)

    def setbadif(self, mask, out=None):
        out = ndarray() if out is None else out
        _run("setbadif", [self, topdl(mask) if not isinstance(mask, ndarray) else mask, out])
        return out

    def setvaltobad(self, value, out=None):
        out = ndarray() if out is None else out
        _run("setvaltobad", [self, out], dpar=float(value))
        return out

    def setbadtoval(self, newval, out=None):
        out = ndarray() if out is None else out
        _run("setbadtoval", [self, out], dpar=float(newval))
        return out

    # whole-array reducers: $x->flat->op (Ufunc.pd:618-663)
    def sum(self): return self.flat().sumover()
    def prod(self): return self.flat().prodover()
    def dsum(self): return self.flat().dsumover()
    def avg(self): return self.flat().average()
    def min(self): return self.flat().minimum()
    def max(self): return self.flat().maximum()

    def minmax(self):
        mn, mx, _, _ = self.flat().minmaximum()
        return mn, mx

    def matmult(self, b, out=None):
        return matmult(self, b, out)


def default_badvalue(t):
    info = {L.SB: -128, L.B: 255, L.S: -32768, L.US: 65535, L.L: -2**31, L.UL: 2**32 - 1, L.IND: -2**63,
            L.ULL: 2**64 - 1, L.LL: -2**63, L.F: float(-np.finfo(np.float32).max),
            L.D: float(-np.finfo(np.float64).max)}
    return info[t]


def _scalar_type(v):
    """PDL's typing of a Perl scalar (pdlperl.h:185-205): integers take the
    smallest signed type that holds them, floats are double."""
    if isinstance(v, (bool, np.bool_)):
        v = int(v)
    if isinstance(v, np.generic):
        return _FROM_NP[v.dtype]
    if isinstance(v, int):
        for t, lo, hi in ((L.SB, -2**7, 2**7 - 1), (L.S, -2**15, 2**15 - 1), (L.L, -2**31, 2**31 - 1),
                          (L.IND, -2**63, 2**63 - 1)):
            if lo <= v <= hi:
                return t
        if 0 <= v < 2**64:
            return L.ULL
        raise PDLError(L.EUSERERROR, "integer %d out of range" % v)
    return L.D


def topdl(x, dtype=None):
    if isinstance(x, ndarray):
        return x
    if isinstance(x, (int, float, bool, np.generic)):
        t = typecode(dtype) if dtype is not None else _scalar_type(x)
        return ndarray.from_numpy(np.array(x, dtype=NP_TYPES[t]), t)
    return ndarray.from_numpy(np.asarray(x), dtype)


def pdl(data, dtype=None, badflag=False):
    """PDL->new: nested lists / numpy in C order (numpy shape = reversed dims)"""
    a = np.asarray(data)
    if dtype is None and a.dtype.kind in "iub":
        dtype = L.D   # pdl(...) of Perl numbers is double
    return ndarray.from_numpy(a, dtype if dtype is not None else L.D, badflag=badflag)


def null():
    return ndarray()


def _fill_const(p, v):
    c = p._to_c()
    L.check(L.lib().pdlhip_fill_uniform(ctypes.byref(c), 0, float(v), float(v)))
    _written(p._buf)
    return p


def zeroes(*dims, dtype=L.D):
    return _fill_const(ndarray.new_dense(dtype, dims), 0)


def ones(*dims, dtype=L.D):
    return _fill_const(ndarray.new_dense(dtype, dims), 1)


def empty(dtype=L.SB):
    return ndarray.new_dense(dtype, [0])


def sequence(*dims, dtype=L.D):
    n = 1
    for d in dims:
        n *= d
    return ndarray.from_numpy(np.arange(n).reshape(dims[::-1]), dtype)


# ---------------------------------------------------------------- PDL's random numbers
def srandom(seed):
    """PDL::srandom (Primitive.pd:2113-2140): seed the stream"""
    _need()
    L.check(L.lib().pdlhip_srandom(int(seed) & (2**64 - 1)))


def rand_skip(n):
    """move the stream on by n draws, as n elements of random would"""
    _need()
    L.check(L.lib().pdlhip_rand_skip(int(n)))


def rand_state():
    """the xoshiro256+ state the next draw starts from (4 words)"""
    _need()
    st = (ctypes.c_uint64 * 4)()
    L.check(L.lib().pdlhip_rand_state(st))
    return [int(v) for v in st]


def _rand_into(x, entry):
    _refresh(x._buf)
    _flush(x._buf)
    c = x._to_c()
    L.check(getattr(L.lib(), entry)(ctypes.byref(c)))
    _written(x._buf)
    return x


def _rand_new(entry, dims, dtype):
    _need()
    if len(dims) == 1 and isinstance(dims[0], (list, tuple)):
        dims = tuple(dims[0])
    return _rand_into(ndarray.new_dense(L.D if dtype is None else dtype, dims), entry)


def random(*dims, dtype=None):
    """PDL::random([type], dims...) (Primitive.pd:2142-2179): uniform [0,1)
    draws of PDL's stream, one per element in PDL order; double by default"""
    return _rand_new("pdlhip_run_random", dims, dtype)


def randsym(*dims, dtype=None):
    """PDL::randsym (Primitive.pd:2181-2219): the same draws, 0 excluded"""
    return _rand_new("pdlhip_run_randsym", dims, dtype)


def grandom(*dims, dtype=None):
    """PDL::grandom (Primitive.pd:2221-2262): ndtri(randsym) in the type
    (float or double)"""
    return _rand_new("pdlhip_grandom", dims, dtype)


def random_into(x):
    """pdl_run_random on an existing ndarray (a view fills through its
    dims/incs, the parent's other elements untouched)"""
    _need()
    return _rand_into(x, "pdlhip_run_random")


def randsym_into(x):
    _need()
    return _rand_into(x, "pdlhip_run_randsym")


def grandom_into(x):
    _need()
    return _rand_into(x, "pdlhip_grandom")


def _run(op, pdls, swap=0, m_size=None, dpar=None):
    """One pdl_run_<op> call through the C ABI (dpar: a double OtherPar,
    passed as its IEEE bit pattern), with the data kept in step: inputs that
    are dataflow children are recomputed if their parents changed, host
    writes go up first, written buffers move on and write dataflow children
    back into their parents."""
    _need()
    if op == "assgn_convert":
        op = "assgn"
    _npdls, npar = _op_info(op)
    for p in pdls:
        if p._buf is not None:
            _refresh(p._buf)
            _flush(p._buf)
    _call(op, pdls, swap, m_size, dpar)
    seen = set()
    for p in pdls[npar:]:
        if p._buf is not None and id(p._buf) not in seen:
            seen.add(id(p._buf))
            _written(p._buf)


def _call(op, pdls, swap=0, m_size=None, dpar=None):
    """the bare C ABI call (no dataflow bookkeeping)"""
    uniq, cs = {}, []
    for p in pdls:
        if id(p) not in uniq:
            uniq[id(p)] = p._to_c()
        cs.append(uniq[id(p)])
    created = {id(p): p.isnull() for p in pdls}
    arr = (L.PP * len(pdls))(*[ctypes.pointer(c) for c in cs])
    if dpar is not None:
        other = (ctypes.c_int64 * 1)(int(np.array([dpar], np.float64).view(np.int64)[0]))
    else:
        other = (ctypes.c_int64 * 1)(swap if m_size is None else m_size)
    L.check(L.lib().pdlhip_run(op.encode(), arr, len(pdls), other))
    done = set()
    for p in pdls:
        if id(p) in done:
            continue
        done.add(id(p))
        p._from_c(uniq[id(p)], created[id(p)])
    return pdls


def run(op, *pdls, swap=0, m_size=None, dpar=None):
    """PDL::<op>(pdls...) with explicit outputs (null() to create)"""
    return _run(op, [topdl(p) if not isinstance(p, ndarray) else p for p in pdls], swap=swap, m_size=m_size,
                dpar=dpar)


def matmult(x, y, c=None):
    """PDL::matmult PMCode (Primitive.pd:197-209): promote <2-D operands with
    dummy dims, 1x1 operands multiply elementwise, mismatched inner dims barf."""
    x = topdl(x)
    y = topdl(y)
    c = ndarray() if c is None else c
    while x.ndims < 2:
        x = x.dummy(-1)
    while y.ndims < 2:
        y = y.dummy(-1)
    if (x.dim(0) == 1 and x.dim(1) == 1) or (y.dim(0) == 1 and y.dim(1) == 1):
        t = x * y
        _run("assgn", [t, c])
        return c
    if y.dim(1) != x.dim(0):
        raise PDLError(L.EUSERERROR, "Dim mismatch in matmult of [%dx%d] x [%dx%d]: %d != %d"
                       % (x.dim(0), x.dim(1), y.dim(0), y.dim(1), x.dim(0), y.dim(1)))
    _run("matmult", [x, y, c])
    return c


def matmult_list(As, Bs, Cs=None, mode="queue"):
    """[A_i x B_i] for a list of uneven products in one persistent
    work-queue launch (mode "queue") or one launch per product ("each")."""
    n = len(As)
    if len(Bs) != n:
        raise PDLError(L.EUSERERROR, "matmult_list: %d left and %d right operands" % (n, len(Bs)))
    Cs = [ndarray() for _ in range(n)] if Cs is None else list(Cs)
    _need()
    ca = [topdl(a)._to_c() for a in As]
    cb = [topdl(b)._to_c() for b in Bs]
    cc = [c._to_c() for c in Cs]
    created = [c.isnull() for c in Cs]
    pa = (L.PP * n)(*[ctypes.pointer(x) for x in ca])
    pb = (L.PP * n)(*[ctypes.pointer(x) for x in cb])
    pc = (L.PP * n)(*[ctypes.pointer(x) for x in cc])
    _flush_all()
    L.check(L.lib().pdlhip_run_matmult_list(n, pa, pb, pc, {"queue": 0, "each": 1}[mode]))
    _launched()
    for c, s, cr in zip(Cs, cc, created):
        c._from_c(s, cr)
    return Cs


def _nodes(chain):
    def operand(v):
        return -1 - v[1] if isinstance(v, tuple) else v
    return (L.Node * max(1, len(chain)))(*[L.Node(L.BIOPS.index(op), operand(x), operand(y)) for op, x, y in chain])


class _Prepared:
    """a prepared ragged list / chain handle (pdlhip_ragged)"""

    _h = None

    def run(self):
        _flush_all()
        L.check(L.lib().pdlhip_ragged_run(self._h))
        _launched()
        return self

    def _flags(self):
        n = ctypes.c_int64()
        L.check(L.lib().pdlhip_ragged_count(self._h, ctypes.byref(n)))
        f = (ctypes.c_int32 * max(1, n.value))()
        L.check(L.lib().pdlhip_ragged_badflags(self._h, f))
        return [bool(f[i]) for i in range(n.value)]

    def free(self):
        if self._h:
            L.check(L.lib().pdlhip_ragged_free(self._h))
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Ragged(_Prepared):
    """A prepared list of 1-D ndarrays of uneven lengths, evaluated in a fixed
    set of launches: per element a chain of binops over the input slots,
    optionally kept (values=True), then `final` (sumover, dsumover, average,
    daverage, minimum, maximum, cumusumover, cumuprodover, or "none") per
    item.

    items:  list of tuples, one ndarray (1-D, or a 0-dim scalar slot) per slot
    chain:  list of (op, x, y); op is an Ops.pd binop name, an operand is a
            slot index k >= 0 or ("node", j) for the result of node j
    final:  the Ufunc op applied to the chain's value
    Equivalent to the per-ndarray loop
        for a, b, ... in items: out_i = (chain evaluated eagerly).final()
    with the same rounding at every step."""

    def __init__(self, items, chain, final, outputs=None, values=None):
        _need()
        n = len(items)
        nin = len(items[0]) if n else 1
        if any(len(t) != nin for t in items):
            raise PDLError(L.EUSERERROR, "Ragged: items have different numbers of inputs")
        objs = [topdl(x) for t in items for x in t]
        cin = [o._to_c() for o in objs]
        pin = (L.PP * max(1, len(cin)))(*[ctypes.pointer(x) for x in cin])
        self.outputs = None
        if final != "none":
            self.outputs = [ndarray() for _ in range(n)] if outputs is None else list(outputs)
        self.values = None
        if values is not None and values is not False:
            self.values = [ndarray() for _ in range(n)] if values is True else list(values)
        cout = [o._to_c() for o in self.outputs] if self.outputs is not None else []
        cval = [o._to_c() for o in self.values] if self.values is not None else []
        pout = (L.PP * max(1, n))(*[ctypes.pointer(x) for x in cout]) if self.outputs is not None else None
        pval = (L.PP * max(1, n))(*[ctypes.pointer(x) for x in cval]) if self.values is not None else None
        created = [o.isnull() for o in (self.outputs or [])] + [o.isnull() for o in (self.values or [])]
        h = ctypes.c_void_p()
        L.check(L.lib().pdlhip_ragged_prepare(n, nin, pin, len(chain), _nodes(chain), final.encode(), pout, pval,
                                              ctypes.byref(h)))
        self._h = h
        self._keep = objs
        for o, s, cr in zip((self.outputs or []) + (self.values or []), cout + cval, created):
            o._from_c(s, cr)

    def badflags(self):
        """after a run: per item, whether its output is BAD (minimum/maximum
        of an item with no good value); sets those outputs' bad-value flag"""
        f = self._flags()
        for o, b in zip(self.outputs or [], f):
            if b:
                o.state |= L.BADVAL
        return f


class Chain(_Prepared):
    """A chain of binops over ndarrays that broadcast together, fused into one
    pass: `value` (value=True or an ndarray) receives the chain's result over
    every broadcast dim, `result` the final Ufunc op along dim 0.  For
    example c = a*b + a and s = sumover(c) in one pass:

        ch = Chain([a, b], [("mult", 0, 1), ("plus", ("node", 0), 0)], "sumover", value=True)
        ch.run(); c, s = ch.value, ch.result
    """

    def __init__(self, inputs, chain, final="none", value=None, result=None):
        _need()
        objs = [topdl(x) for x in inputs]
        cin = [o._to_c() for o in objs]
        pin = (L.PP * len(cin))(*[ctypes.pointer(x) for x in cin])
        self.value = None
        if value is not None and value is not False:
            self.value = ndarray() if value is True else value
        self.result = None
        if final != "none":
            self.result = ndarray() if result is None else result
        cv = self.value._to_c() if self.value is not None else None
        cr = self.result._to_c() if self.result is not None else None
        created = (self.value.isnull() if self.value is not None else False,
                   self.result.isnull() if self.result is not None else False)
        h = ctypes.c_void_p()
        L.check(L.lib().pdlhip_chain_prepare(len(cin), pin, len(chain), _nodes(chain), final.encode(),
                                             ctypes.pointer(cv) if cv is not None else None,
                                             ctypes.pointer(cr) if cr is not None else None, ctypes.byref(h)))
        self._h = h
        self._keep = objs
        if self.value is not None:
            self.value._from_c(cv, created[0])
        if self.result is not None:
            self.result._from_c(cr, created[1])

    def badflags(self):
        """after a run: whether any row's minimum/maximum found no good value;
        sets the result's bad-value flag then"""
        bad = any(self._flags())
        if bad and self.result is not None:
            self.result.state |= L.BADVAL
        return bad


def tune_stats():
    """{'measured': n, 'reused': m}: autotuner choices this process timed and
    found in the cache file"""
    _need()
    m, r = ctypes.c_int64(), ctypes.c_int64()
    L.check(L.lib().pdlhip_tune_stats(ctypes.byref(m), ctypes.byref(r)))
    return {"measured": m.value, "reused": r.value}


class Graph:
    """A HIP graph of the work `fn` enqueues on the library's stream.  Ops
    may take scratch memory only from the library's cache while capturing
    (run an op once before capturing it); blocks they release stay with the
    graph until destroy().  Ops that wait for the stream (downloads, bad-value
    *_n_ind, ragged bad flags) are refused."""

    def __init__(self, fn):
        _need()
        self._g = None
        L.check(L.lib().pdlhip_graph_begin())
        g = ctypes.c_void_p()
        try:
            fn()
        except BaseException:
            L.lib().pdlhip_graph_end(ctypes.byref(g))
            if g.value:
                L.lib().pdlhip_graph_destroy(g)
            raise
        L.check(L.lib().pdlhip_graph_end(ctypes.byref(g)))
        self._g = g

    def launch(self):
        _flush_all()
        L.check(L.lib().pdlhip_graph_launch(self._g))
        _launched()

    def destroy(self):
        if self._g:
            L.check(L.lib().pdlhip_graph_destroy(self._g))
            self._g = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class MatmultList:
    """A prepared list of products A_i x B_i (uneven sizes): planned once
    (outputs created as matmult creates them), then each launch() runs the
    whole list in one work-queue launch ("queue") or one launch per product
    ("each") without allocating, so launches can be captured in a Graph."""

    def __init__(self, As, Bs, Cs=None):
        _need()
        n = len(As)
        if len(Bs) != n:
            raise PDLError(L.EUSERERROR, "MatmultList: %d left and %d right operands" % (n, len(Bs)))
        self.outputs = [ndarray() for _ in range(n)] if Cs is None else list(Cs)
        objs = [topdl(a) for a in As] + [topdl(b) for b in Bs]
        ca = [o._to_c() for o in objs[:n]]
        cb = [o._to_c() for o in objs[n:]]
        cc = [c._to_c() for c in self.outputs]
        created = [c.isnull() for c in self.outputs]
        pa = (L.PP * max(1, n))(*[ctypes.pointer(x) for x in ca])
        pb = (L.PP * max(1, n))(*[ctypes.pointer(x) for x in cb])
        pc = (L.PP * max(1, n))(*[ctypes.pointer(x) for x in cc])
        h = ctypes.c_void_p()
        L.check(L.lib().pdlhip_matmult_list_prepare(n, pa, pb, pc, ctypes.byref(h)))
        self._h = h
        self._keep = objs
        for o, s, cr in zip(self.outputs, cc, created):
            o._from_c(s, cr)

    def launch(self, mode="queue"):
        _flush_all()
        L.check(L.lib().pdlhip_matmult_list_launch(self._h, {"queue": 0, "each": 1}[mode]))
        _launched()
        return self

    def free(self):
        if self._h:
            L.check(L.lib().pdlhip_matmult_list_free(self._h))
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def view_into(parent, dims, offs):
    """a dense-strided view of `dims` starting at element `offs` of parent's
    root buffer (for packing many small ndarrays into one allocation)"""
    return parent._view(list(dims), _dense_incs(dims), offs)


def set_matmult_precision(mode):
    """'split16' (default), 'fp32', 'pdlorder' or 'fp8' for float matmult"""
    m = {"split16": L.MM_SPLIT16, "fp32": L.MM_FP32, "pdlorder": L.MM_PDLORDER, "fp8": L.MM_FP8}[mode]
    L.check(L.lib().pdlhip_set_matmult_precision(m))


def synchronize():
    _need()
    L.check(L.lib().pdlhip_synchronize())


class Timer:
    """elapsed ms of work enqueued on the library's stream (HIP events)"""

    def __enter__(self):
        _need()
        L.check(L.lib().pdlhip_timer_start())
        return self

    def __exit__(self, *a):
        ms = ctypes.c_float()
        L.check(L.lib().pdlhip_timer_stop(ctypes.byref(ms)))
        self.ms = ms.value


class Event:
    """a HIP event recorded on the library's stream"""

    def __init__(self):
        _need()
        p = ctypes.c_void_p()
        L.check(L.lib().pdlhip_event_create(ctypes.byref(p)))
        self.ev = p.value

    def record(self):
        L.check(L.lib().pdlhip_event_record(self.ev))
        return self

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        L.check(L.lib().pdlhip_event_elapsed(self.ev, end.ev, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        if getattr(self, "ev", None) and L._lib is not None:
            try:
                L._lib.pdlhip_event_destroy(self.ev)
            except Exception:
                pass


# functional forms (PDL::plus($a, $b, $c, $swap) ...)
def _make_biop(name):
    def f(a, b, c=None, swap=0):
        c = ndarray() if c is None else c
        _run(name, [topdl(a), topdl(b), c], swap=swap)
        return c
    f.__name__ = name
    return f


def _make_unary(name):
    def f(a, b=None):
        b = ndarray() if b is None else b
        _run(name, [topdl(a), b])
        return b
    f.__name__ = name
    return f


for _n in L.BIOPS:
    globals()[_n] = _make_biop(_n)
for _n in L.UNARY:
    globals()[_n] = _make_unary(_n)
