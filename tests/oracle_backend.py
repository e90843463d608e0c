"""Adapter exposing the oracle (oracle/pdlcpu.py) as a harness backend."""
import numpy as np

from oracle import pdlcpu as O


class OracleBackend:
    name = "oracle"
    Error = O.OracleError

    def make(self, tname, arr, badflag=False):
        return O.OPdl.from_numpy(arr, dtype=O.TYPE_NAMES.index(tname), badflag=badflag)

    def null(self):
        return O.OPdl.null()

    def run(self, op, pdls, swap=0, m_size=None, dpar=None):
        O.run(op, *pdls, swap=swap, m_size=m_size, dpar=dpar)

    def dims(self, p):
        return list(p.dims)

    def nvals(self, p):
        return p.nvals

    def read(self, p):
        return O.TYPE_NAMES[p.dtype], list(p.dims), p.gather().reshape(-1), bool(p.state & O.BADVAL)

    def root_of(self, p):
        return p.root_pdl()
