"""FASTA input and similarity-matrix output of the aligner path
(pa_fasta_*, pa_triples_dev, pa_write_mm in include/pastis_hip.h)."""
import ctypes as C

import numpy as np

from ._lib import TRIPLE_DTYPE, check, lib


def fasta_parse(data, k=6):
    """Sequences of a FASTA buffer as the reference's reader and FastaData
    normalise them: (ASCII residues uint8, offsets uint64, record index of
    each sequence, count of lengths wrapped past 65535)."""
    L = lib()
    data = bytes(data)
    buf = C.create_string_buffer(data, len(data)) if data else None
    h = C.c_void_p()
    check(None, L.pa_fasta_parse(buf, len(data), k, C.byref(h)),
          "pa_fasta_parse")
    try:
        ns, nr, nw = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(None, L.pa_fasta_info(h, C.byref(ns), C.byref(nr),
                                    C.byref(nw)), "pa_fasta_info")
        n, m = ns.value, nr.value
        res = np.zeros(m, np.uint8)
        if m:
            C.memmove(res.ctypes.data, L.pa_fasta_residues(h), m)
        offs = np.zeros(n + 1, np.uint64)
        C.memmove(offs.ctypes.data, L.pa_fasta_offsets(h), 8 * (n + 1))
        recs = np.zeros(n, np.uint64)
        if n:
            C.memmove(recs.ctypes.data, L.pa_fasta_records(h), 8 * n)
    finally:
        L.pa_fasta_free(h)
    return res, offs, recs, nw.value


def write_mm(path, nrows, ncols, triples):
    """Matrix Market file of the similarity triples (TRIPLE_DTYPE)."""
    t = np.ascontiguousarray(triples, TRIPLE_DTYPE)
    check(None, lib().pa_write_mm(path.encode(), nrows, ncols,
                                  C.c_void_p(t.ctypes.data) if t.size
                                  else None, len(t)), "pa_write_mm")


def triples_dev(al, d_pairs, d_res, n, d_out, row_base=0, col_base=0):
    """Compacts the passing pairs of a device batch into triples; returns
    their number."""
    m = C.c_uint64()
    check(al.ctx, lib().pa_triples_dev(al.ctx, d_pairs, d_res, n, row_base,
                                       col_base, d_out, C.byref(m), None),
          "pa_triples_dev")
    return m.value
