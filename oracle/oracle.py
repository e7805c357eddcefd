"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle restates, on the CPU, the alignment semantics that the reference
reaches through SeqAn (src/align/SeqanFullAligner.cpp:173-266,
src/align/SeqanXdropAligner.cpp:190-361).  Only tests/, the smoke check and
the cpu_baseline leg of bench.py import this module; the product library
never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

# Residue order of src/Alphabet.cpp:26.
ALPHABET = "ARNDCQEGHILKMFPSTWYVBZX*J"
X_CODE = ALPHABET.index("X")


class OrAln(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "score", "v_beg", "v_end", "h_beg", "h_end", "matches", "cols",
        "gap_opens")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class OrXdrop(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "beg_h", "beg_v", "end_h", "end_v", "score_left", "score_right",
        "upd_left", "upd_right")] + [("cells", C.c_uint64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class OrParams(C.Structure):
    _fields_ = [("gap_open", C.c_int32), ("gap_ext", C.c_int32),
                ("xdrop_gap", C.c_int32), ("xdrop", C.c_int32),
                ("seed_len", C.c_int32), ("seed_cnt", C.c_int32),
                ("cov_thr", C.c_double), ("ani_thr", C.c_int32)]


class OrPairOut(C.Structure):
    _fields_ = [("aln", OrAln), ("identity", C.c_float),
                ("score_aln", C.c_float), ("pass_", C.c_int32),
                ("seed", C.c_int32), ("ext", OrXdrop * 2)]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u8p = C.POINTER(C.c_uint8)
        i8p = C.POINTER(C.c_int8)
        L.or_blosum62_32.argtypes = [i8p]
        L.or_sw_local.argtypes = [u8p, C.c_int, u8p, C.c_int, i8p, C.c_int,
                                  C.c_int, C.c_int, C.POINTER(OrAln)]
        L.or_nw_global.argtypes = L.or_sw_local.argtypes
        L.or_xdrop_extend.argtypes = [u8p, C.c_int, u8p, C.c_int, C.c_int,
                                      C.c_int, C.c_int, i8p, C.c_int, C.c_int,
                                      C.POINTER(OrXdrop)]
        L.or_sfa_pair.argtypes = [u8p, C.c_int, u8p, C.c_int, i8p,
                                  C.POINTER(OrParams), C.POINTER(OrPairOut)]
        L.or_sxa_pair.argtypes = [u8p, C.c_int, u8p, C.c_int,
                                  C.POINTER(C.c_uint16), i8p,
                                  C.POINTER(OrParams), C.POINTER(OrPairOut)]
        vp = C.c_void_p
        L.or_sw_batch.argtypes = [vp, vp, vp, vp, C.c_uint64, vp, C.c_int,
                                  C.c_int, vp, vp, vp, C.c_int]
        L.or_sfa_batch.argtypes = [vp, vp, vp, vp, C.c_uint64, vp,
                                   C.POINTER(OrParams), vp, C.c_int]
        L.or_sxa_batch.argtypes = [vp, vp, vp, vp, vp, C.c_uint64, vp,
                                   C.POINTER(OrParams), vp, C.c_int]
        _lib = L
    return _lib


def encode(s):
    """ASCII protein string -> codes in ALPHABET order (unknown -> X)."""
    tbl = np.full(256, X_CODE, dtype=np.uint8)
    for i, ch in enumerate(ALPHABET):
        tbl[ord(ch)] = i
        tbl[ord(ch.lower())] = i
    return tbl[np.frombuffer(s.encode(), dtype=np.uint8)]


def blosum62():
    m = np.zeros(32 * 32, dtype=np.int8)
    lib().or_blosum62_32(m.ctypes.data_as(C.POINTER(C.c_int8)))
    return m


def _u8(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint8))


def _mat(mat):
    m = blosum62() if mat is None else np.ascontiguousarray(mat, np.int8)
    return m, m.ctypes.data_as(C.POINTER(C.c_int8))


def sw_local(h, v, gap_open, gap_ext, stats=True, mat=None):
    h, hp = _u8(h)
    v, vp = _u8(v)
    m, mp = _mat(mat)
    out = OrAln()
    rc = lib().or_sw_local(hp, len(h), vp, len(v), mp, gap_open, gap_ext,
                           int(stats), C.byref(out))
    assert rc == 0
    return out.as_dict()


def nw_global(h, v, gap_open, gap_ext, stats=True, mat=None):
    h, hp = _u8(h)
    v, vp = _u8(v)
    m, mp = _mat(mat)
    out = OrAln()
    rc = lib().or_nw_global(hp, len(h), vp, len(v), mp, gap_open, gap_ext,
                            int(stats), C.byref(out))
    assert rc == 0
    return out.as_dict()


def xdrop_extend(h, v, beg_h, beg_v, seed_len, gap, xdrop, mat=None):
    h, hp = _u8(h)
    v, vp = _u8(v)
    m, mp = _mat(mat)
    out = OrXdrop()
    rc = lib().or_xdrop_extend(hp, len(h), vp, len(v), beg_h, beg_v,
                               seed_len, mp, gap, xdrop, C.byref(out))
    if rc:
        raise ValueError("seed out of range")
    return out.as_dict()


def params(gap_open=-2, gap_ext=-11, xdrop_gap=-11, xdrop=49, seed_len=6,
           seed_cnt=2, cov_thr=0.7, ani_thr=30):
    return OrParams(gap_open, gap_ext, xdrop_gap, xdrop, seed_len, seed_cnt,
                    cov_thr, ani_thr)


def _pair_dict(o):
    d = o.aln.as_dict()
    d.update(identity=o.identity, score_aln=o.score_aln, passed=o.pass_,
             seed=o.seed, ext=[o.ext[0].as_dict(), o.ext[1].as_dict()])
    return d


def sfa_pair(h, v, p, mat=None):
    h, hp = _u8(h)
    v, vp = _u8(v)
    m, mp = _mat(mat)
    out = OrPairOut()
    assert lib().or_sfa_pair(hp, len(h), vp, len(v), mp, C.byref(p),
                             C.byref(out)) == 0
    return _pair_dict(out)


def sxa_pair(h, v, seeds, p, mat=None):
    h, hp = _u8(h)
    v, vp = _u8(v)
    m, mp = _mat(mat)
    s = (C.c_uint16 * 4)(*seeds)
    out = OrPairOut()
    rc = lib().or_sxa_pair(hp, len(h), vp, len(v), s, mp, C.byref(p),
                           C.byref(out))
    if rc:
        raise ValueError("oracle error %d" % rc)
    return _pair_dict(out)


PAIR_OUT_DTYPE = np.dtype([
    ("score", "<i4"), ("v_beg", "<i4"), ("v_end", "<i4"), ("h_beg", "<i4"),
    ("h_end", "<i4"), ("matches", "<i4"), ("cols", "<i4"),
    ("gap_opens", "<i4"), ("identity", "<f4"), ("score_aln", "<f4"),
    ("passed", "<i4"), ("seed", "<i4"),
    ("ext", [("beg_h", "<i4"), ("beg_v", "<i4"), ("end_h", "<i4"),
             ("end_v", "<i4"), ("score_left", "<i4"), ("score_right", "<i4"),
             ("upd_left", "<i4"), ("upd_right", "<i4"), ("cells", "<u8")],
     (2,))])
assert PAIR_OUT_DTYPE.itemsize == C.sizeof(OrPairOut)


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


def sw_batch(seqs, offs, pr, pc, gap_open, gap_ext, nthreads=0, mat=None):
    seqs = np.ascontiguousarray(seqs, np.uint8)
    offs = np.ascontiguousarray(offs, np.uint64)
    pr = np.ascontiguousarray(pr, np.uint32)
    pc = np.ascontiguousarray(pc, np.uint32)
    m, _ = _mat(mat)
    n = len(pr)
    score = np.zeros(n, np.int32)
    ve = np.zeros(n, np.int32)
    he = np.zeros(n, np.int32)
    rc = lib().or_sw_batch(_ptr(seqs), _ptr(offs), _ptr(pr), _ptr(pc), n,
                           _ptr(m), gap_open, gap_ext, _ptr(score), _ptr(ve),
                           _ptr(he), nthreads)
    assert rc == 0
    return score, ve, he


def sfa_batch(seqs, offs, pr, pc, p, nthreads=0, mat=None):
    seqs = np.ascontiguousarray(seqs, np.uint8)
    offs = np.ascontiguousarray(offs, np.uint64)
    pr = np.ascontiguousarray(pr, np.uint32)
    pc = np.ascontiguousarray(pc, np.uint32)
    m, _ = _mat(mat)
    out = np.zeros(len(pr), PAIR_OUT_DTYPE)
    rc = lib().or_sfa_batch(_ptr(seqs), _ptr(offs), _ptr(pr), _ptr(pc),
                            len(pr), _ptr(m), C.byref(p), _ptr(out), nthreads)
    assert rc == 0
    return out


def sxa_batch(seqs, offs, pr, pc, seeds, p, nthreads=0, mat=None):
    seqs = np.ascontiguousarray(seqs, np.uint8)
    offs = np.ascontiguousarray(offs, np.uint64)
    pr = np.ascontiguousarray(pr, np.uint32)
    pc = np.ascontiguousarray(pc, np.uint32)
    seeds = np.ascontiguousarray(seeds, np.uint16).reshape(-1, 4)
    m, _ = _mat(mat)
    out = np.zeros(len(pr), PAIR_OUT_DTYPE)
    rc = lib().or_sxa_batch(_ptr(seqs), _ptr(offs), _ptr(pr), _ptr(pc),
                            _ptr(seeds), len(pr), _ptr(m), C.byref(p),
                            _ptr(out), nthreads)
    if rc:
        raise ValueError("oracle batch error")
    return out


# ---- k-mer candidates and FASTA normalisation ----------------------------

CAND_DTYPE = np.dtype([("r", "<u4"), ("c", "<u4"), ("count", "<u2"),
                       ("kscore", "<i2"), ("seed", "<u2", (4,))])
ALPHABETS = {"pdefault": 0, "murphy10": 1, "dssp10": 2, "gbmr10": 3,
             "td10": 4, "diamond": 5}
SYMMETRIC = {"none": 0, "upper": 1, "even_odd": 2}


def _kmer_lib():
    L = lib()
    if not hasattr(L, "_kmer_ready"):
        vp = C.c_void_p
        L.or_kmer_candidates.argtypes = [vp, vp, C.c_uint64, vp, vp,
                                         C.c_uint64, C.c_int, C.c_int,
                                         C.c_int, C.c_int, C.c_uint64,
                                         C.c_uint64, C.POINTER(vp),
                                         C.POINTER(C.c_uint64)]
        L.or_free.argtypes = [vp]
        L.or_fasta_normalize.argtypes = [vp, C.c_uint64, C.c_int,
                                         C.POINTER(vp), C.POINTER(vp),
                                         C.POINTER(vp), C.POINTER(C.c_uint64)]
        L.or_alph_tables.argtypes = [C.c_int, vp, vp, C.POINTER(C.c_int)]
        L._kmer_ready = True
    return L


def alph_tables(alph="pdefault"):
    """(code_of[128] with 255 = not a k-mer letter, self_of[128], size)."""
    L = _kmer_lib()
    code = np.zeros(128, np.uint8)
    self_ = np.zeros(128, np.int16)
    size = C.c_int()
    assert L.or_alph_tables(ALPHABETS[alph], _ptr(code), _ptr(self_),
                            C.byref(size)) == 0
    return code, self_, size.value


def _take(vp, dtype, n):
    if n == 0:
        lib().or_free(vp)
        return np.zeros(0, dtype)
    buf = (C.c_char * (n * np.dtype(dtype).itemsize)).from_address(vp.value)
    out = np.frombuffer(bytes(buf), dtype).copy()
    lib().or_free(vp)
    return out


def kmer_candidates(rres, roffs, cres=None, coffs=None, k=6, stride=1,
                    alph="pdefault", sym="none", roff=0, coff=0):
    """Overlap-matrix nonzeros (CAND_DTYPE, column-major) of ASCII row and
    column sets; cres None means the column set is the row set."""
    L = _kmer_lib()
    rres = np.ascontiguousarray(rres, np.uint8)
    roffs = np.ascontiguousarray(roffs, np.uint64)
    if cres is not None:
        cres = np.ascontiguousarray(cres, np.uint8)
        coffs = np.ascontiguousarray(coffs, np.uint64)
    out = C.c_void_p()
    n = C.c_uint64()
    rc = L.or_kmer_candidates(
        _ptr(rres) if rres.size else None, _ptr(roffs), len(roffs) - 1,
        None if cres is None else (_ptr(cres) if cres.size else None),
        None if cres is None else _ptr(coffs),
        0 if cres is None else len(coffs) - 1, k, stride, ALPHABETS[alph],
        SYMMETRIC[sym], roff, coff, C.byref(out), C.byref(n))
    assert rc == 0
    return _take(out, CAND_DTYPE, n.value)


def fasta_normalize(data, k=6):
    """(residues uint8, offsets uint64, record index) of a FASTA file's
    bytes as one process of the reference sees them."""
    L = _kmer_lib()
    data = np.frombuffer(bytes(data), np.uint8)
    res, offs, orig = C.c_void_p(), C.c_void_p(), C.c_void_p()
    ns = C.c_uint64()
    assert L.or_fasta_normalize(_ptr(data) if data.size else None, data.size,
                                k, C.byref(res), C.byref(offs),
                                C.byref(orig), C.byref(ns)) == 0
    n = ns.value
    if not offs.value:
        return (np.zeros(0, np.uint8), np.zeros(1, np.uint64),
                np.zeros(0, np.uint64))
    o = _take(offs, np.uint64, n + 1)
    r = _take(res, np.uint8, int(o[-1]))
    g = _take(orig, np.uint64, n)
    return r, o, g
