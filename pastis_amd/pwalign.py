"""Python face of the C++ operator layer (libpastis_pw.so, include/pastis_pw.h):
one diagonal block through HipAligner + pw_aln_batch, the path the
reference's similarity search takes (src/sim_search.cpp:119-135,
src/DistPWRunner.cpp:30-160)."""
import ctypes as C
import os

import numpy as np

from . import _lib
from .aligner import ALPHABET

PW_PATH = os.path.join(_lib.HERE, "libpastis_pw.so")
_pw = None

STATS = ("nnz_in", "nnz_kmer", "nnz_aln", "batches", "aligned",
         "filtered_out", "passed", "aln_pairs")


def lib():
    """Load libpastis_pw.so; raises if it is missing (no fallback)."""
    global _pw
    if _pw is None:
        if not os.path.exists(PW_PATH):
            raise ImportError("libpastis_pw.so is not built: run "
                              "__graft_entry__.build()")
        _lib.lib()   # libpastis_hip.so first (same directory, rpath)
        L = C.CDLL(PW_PATH)
        vp, u64, i32 = C.c_void_p, C.c_uint64, C.c_int
        L.pw_run_block.argtypes = ([i32] * 8 + [vp, vp, u64, vp, vp, u64,
                                                u64, u64, u64] + [vp] * 5 +
                                   [i32, C.c_float, u64, C.c_double, i32,
                                    vp, vp, vp, C.c_float, vp, vp])
        L.pw_bench_aln_batch.argtypes = ([i32] * 8 + [vp, vp, u64, vp, u64,
                                                      C.c_double, i32, i32,
                                                      i32, i32, vp, vp])
        _pw = L
    return _pw


def ascii_of(codes):
    tbl = np.frombuffer(ALPHABET.encode(), np.uint8)
    return tbl[np.asarray(codes, np.uint8)]


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else None


def run_block(mode, codes, offsets, rows, cols, count, score, seeds=None,
              gap_open=-11, gap_ext=-2, klength=6, xdrop=49, seed_count=2,
              gap_model="as_called", ckthr=0, mosthr=-1.0, batch_sz=10 ** 10,
              cov_thr=0.7, ani_thr=30, filt=None, device=0, ccodes=None,
              coffsets=None, bl_roffset=0, bl_coffset=0):
    """Runs one block; returns (score_aln per nonzero, stats dict).
    filt = (params, mean, istd, tau) enables the learned prefilter.
    ccodes / coffsets give a separate column set (default: the row set);
    rows / cols are block-local, offset by bl_roffset / bl_coffset."""
    res = np.ascontiguousarray(ascii_of(codes))
    offs = np.ascontiguousarray(offsets, np.uint64)
    cres = cof = None
    if ccodes is not None:
        cres = np.ascontiguousarray(ascii_of(ccodes))
        cof = np.ascontiguousarray(coffsets, np.uint64)
    rows = np.ascontiguousarray(rows, np.uint64)
    cols = np.ascontiguousarray(cols, np.uint64)
    count = np.ascontiguousarray(count, np.uint16)
    score = np.ascontiguousarray(score, np.int16)
    seeds = (np.ascontiguousarray(seeds, np.uint16).reshape(-1)
             if seeds is not None else None)
    n = len(rows)
    out = np.zeros(n, np.float32)
    st = np.zeros(8, np.uint64)
    gm = {"as_called": _lib.PA_GAP_AS_CALLED,
          "as_documented": _lib.PA_GAP_AS_DOCUMENTED}[gap_model]
    fp = fm = fs = None
    tau = 0.0
    if filt is not None:
        fp = np.ascontiguousarray(filt[0], np.float32)
        fm = np.ascontiguousarray(filt[1], np.float32)
        fs = np.ascontiguousarray(filt[2], np.float32)
        tau = float(filt[3])
    rc = lib().pw_run_block(mode, gap_open, gap_ext, klength, xdrop,
                            seed_count, gm, device, _p(res), _p(offs),
                            len(offs) - 1, _p(cres), _p(cof),
                            0 if cof is None else len(cof) - 1, bl_roffset,
                            bl_coffset, n, _p(rows), _p(cols), _p(count),
                            _p(score), _p(seeds), ckthr, mosthr, batch_sz,
                            cov_thr, ani_thr, _p(fp), _p(fm), _p(fs), tau,
                            _p(out), _p(st))
    _lib.check(None, rc, "pw_run_block")
    return out, dict(zip(STATS, (int(v) for v in st)))


def bench_aln_batch(mode, codes, offsets, pairs, gap_open=-11, gap_ext=-2,
                    klength=6, xdrop=49, seed_count=2, gap_model="as_called",
                    cov_thr=0.7, ani_thr=30, autotune=True, warmup=1,
                    steps=3, device=0):
    """Wall time of HipAligner::aln_batch on host tuples made from `pairs`
    (PAIR_DTYPE, block-local indices of a diagonal block): returns
    (ms per call, pairs passed)."""
    res = np.ascontiguousarray(ascii_of(codes))
    offs = np.ascontiguousarray(offsets, np.uint64)
    pairs = np.ascontiguousarray(pairs, _lib.PAIR_DTYPE)
    gm = {"as_called": _lib.PA_GAP_AS_CALLED,
          "as_documented": _lib.PA_GAP_AS_DOCUMENTED}[gap_model]
    ms = C.c_double(0)
    passed = C.c_uint64(0)
    rc = lib().pw_bench_aln_batch(mode, gap_open, gap_ext, klength, xdrop,
                                  seed_count, gm, device, _p(res), _p(offs),
                                  len(offs) - 1, _p(pairs), len(pairs),
                                  cov_thr, ani_thr, 1 if autotune else 0,
                                  warmup, steps, C.byref(ms),
                                  C.byref(passed))
    _lib.check(None, rc, "pw_bench_aln_batch")
    return ms.value, passed.value
