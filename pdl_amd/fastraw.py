"""PDL::IO::FastRaw for ndarrays in HBM (lib/PDL/IO/FastRaw.pm).

The format: the data file holds the ndarray's elements in PDL's own order
(dim 0 fastest), native byte order, nothing else; a text header (default
"<name>.hdr", or the Header option) holds three lines: the datatype code
(PDL::Types numbering, the same codes as pdl_amd), the number of dims, and
the dims separated by spaces (FastRaw.pm:202-231).

    writefraw(x, name, header=None)   FastRaw.pm:255-262 (data + header)
    readfraw(name, header=None)       FastRaw.pm:293-311 (a new ndarray, uploaded to HBM)
    gluefraw(x, name, header=None)    FastRaw.pm:323-345 (append one more slice along the last dim)
    mapfraw(name, header=None)        a read-only numpy memmap of the data (FastRaw.pm's mapfraw maps it)

The host-level halves (read_header / write_header / read_host / write_host)
work on numpy arrays and need no GPU.  Error texts follow FastRaw.pm's
barf messages.
"""
import os

import numpy as np

from . import _lib as L
from ._lib import PDLError

TYPE_NAMES = ["sbyte", "byte", "short", "ushort", "long", "ulong", "indx", "ulonglong", "longlong",
              "float", "double", "ldouble", "cfloat", "cdouble", "cldouble"]
NP_OF_CODE = {0: np.int8, 1: np.uint8, 2: np.int16, 3: np.uint16, 4: np.int32, 5: np.uint32, 6: np.int64,
              7: np.uint64, 8: np.int64, 9: np.float32, 10: np.float64, 12: np.complex64, 13: np.complex128}


def _hname(name, header):
    return header if header else name + ".hdr"


def read_header(name, header=None):
    """FastRaw.pm:202-219: {'Type': code, 'Dims': [...], 'NDims': n}"""
    hname = _hname(name, header)
    try:
        with open(hname) as h:
            lines = h.read().split("\n")
    except OSError as e:
        raise PDLError(L.EUSERERROR, "Couldn't open '%s' for reading: %s" % (hname, e.strerror))
    if len(lines) < 3 or lines[2] == "" and len(lines) < 4:
        raise PDLError(L.EUSERERROR, "Format error in '%s'" % hname)
    tid, ndims, dims = lines[0].strip(), lines[1].strip(), lines[2].split()
    if len(dims) != int(ndims):
        raise PDLError(L.EUSERERROR, "Format error reading fraw header file '%s'" % hname)
    return {"Type": int(tid), "Dims": [int(d) for d in dims], "NDims": int(ndims)}


def write_header(name, code, dims, header=None):
    """FastRaw.pm:221-229: datatype, ndims, dims -- one per line"""
    hname = _hname(name, header)
    try:
        with open(hname, "w") as h:
            h.write("%d\n%d\n%s\n" % (code, len(dims), " ".join(str(int(d)) for d in dims)))
    except OSError as e:
        raise PDLError(L.EUSERERROR, "Couldn't open '%s' for writing: %s" % (hname, e.strerror))


def _np_type(code):
    if code not in NP_OF_CODE:
        raise PDLError(L.EUSERERROR, "fastraw: type %s has no numpy counterpart" %
                       (TYPE_NAMES[code] if 0 <= code < len(TYPE_NAMES) else code))
    return np.dtype(NP_OF_CODE[code])


def write_host(name, code, dims, data, header=None):
    """data: the elements in PDL order (a numpy array of the type's dtype)"""
    write_header(name, code, dims, header)
    flat = np.ascontiguousarray(data, dtype=_np_type(code)).reshape(-1)
    with open(name, "wb") as d:
        d.write(flat.tobytes())


def read_host(name, header=None):
    """(code, dims, numpy array shaped reversed(dims), i.e. C order of PDL's layout)"""
    hdr = read_header(name, header)
    dt = _np_type(hdr["Type"])
    n = 1
    for d in hdr["Dims"]:
        n *= d
    try:
        with open(name, "rb") as f:
            raw = f.read(n * dt.itemsize)
    except OSError as e:
        raise PDLError(L.EUSERERROR, "Couldn't open '%s' for reading: %s" % (name, e.strerror))
    # a short file leaves the rest zero, as readfraw's zeroes() + sysread loop does
    buf = np.zeros(n, dtype=dt)
    got = np.frombuffer(raw, dtype=dt)
    buf[:got.size] = got
    return hdr["Type"], hdr["Dims"], buf.reshape(hdr["Dims"][::-1])


def glue_host(name, code, dims, data, header=None):
    """FastRaw.pm:323-345: append one item whose dims equal the file's
    lower dims; the header's last dim grows by one"""
    if not os.path.isfile(name):
        raise PDLError(L.EUSERERROR, "'%s' must be real filename: No such file or directory" % name)
    hdr = read_header(name, header)
    if code != hdr["Type"]:
        raise PDLError(L.EUSERERROR, "gluefraw: ndarray has type '%s' but file has type '%s'"
                       % (TYPE_NAMES[code], hdr["Type"]))
    fdims = hdr["Dims"]
    if len(fdims) < 2:
        raise PDLError(L.EUSERERROR, "gluefraw: header dims needs at least 2 dims, got (%s)"
                       % " ".join(map(str, fdims)))
    low = fdims[:-1]
    if list(dims) != list(low):
        raise PDLError(L.EUSERERROR, "gluefraw: incompatible lower dims, ndarray (%s) vs header (%s)"
                       % (" ".join(map(str, dims)), " ".join(map(str, low))))
    with open(name, "ab") as d:
        d.write(np.ascontiguousarray(data, dtype=_np_type(code)).reshape(-1).tobytes())
    write_header(name, code, low + [fdims[-1] + 1], header)


# ---------------------------------------------------------------- ndarrays in HBM
def writefraw(x, name, header=None):
    """x: a pdl_amd ndarray (its data comes back from HBM once, at this
    physical boundary)"""
    write_host(name, x.dtype, x.dims, x.numpy(), header)


def readfraw(name, header=None):
    from . import core as C
    code, dims, arr = read_host(name, header)
    return C.ndarray.from_numpy(arr, code) if arr.ndim else C.ndarray.from_numpy(arr.reshape(()), code)


def gluefraw(x, name, header=None):
    glue_host(name, x.dtype, x.dims, x.numpy(), header)


def mapfraw(name, header=None):
    """the file's data as a read-only numpy memmap in C order of PDL's layout"""
    hdr = read_header(name, header)
    dt = _np_type(hdr["Type"])
    return np.memmap(name, dtype=dt, mode="r", shape=tuple(hdr["Dims"][::-1]))
