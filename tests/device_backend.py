"""Adapter exposing pdl_amd (HIP through the C ABI) as a harness backend."""
import pdl_amd
from pdl_amd import core as C


class DeviceBackend:
    name = "device"
    Error = pdl_amd.PDLError

    def make(self, tname, arr, badflag=False):
        return C.ndarray.from_numpy(arr, dtype=C.TYPE_NAMES.index(tname), badflag=badflag)

    def null(self):
        return C.ndarray()

    def run(self, op, pdls, swap=0, m_size=None, dpar=None):
        C._run(op, pdls, swap=swap, m_size=m_size, dpar=dpar)

    def dims(self, p):
        return list(p.dims)

    def nvals(self, p):
        return p.nelem

    def read(self, p):
        return C.TYPE_NAMES[p.dtype], list(p.dims), p.numpy().reshape(-1), p.badflag()

    def root_of(self, p):
        while p._parent is not None:
            p = p._parent
        return p
