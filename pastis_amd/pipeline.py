"""The candidate stream of one GPU: the similarity search's inner loop on a
resident sequence set.

The reference runs it as blocked SpGEMM + symmetric pruning + pw_aln_batch
+ triple collection (src/sim_search.cpp:58-278: getBlockId, PruneI,
pw_aln_batch, the DCSC walk into rids/cids/vals).  Here every step stays on
the device:
  1. pa_kmer_candidates: the overlap matrix for a block of columns, with the
     symmetric rule and the k-mer threshold prune (ckthr / mosthr) fused;
  2. batches of at most `batch` candidates (--bsz) go through the learned
     filter when one is installed (pa_filter_select_dev), then the X-drop
     aligner (pa_align_dev);
  3. pa_triples_dev keeps the passing pairs as (row, column, score_aln).
Only counts cross to the host until the triples are copied out."""
import time

import numpy as np

from ._lib import PAIR_DTYPE, RESULT_DTYPE, TRIPLE_DTYPE
from . import io


class CandidateStream:
    def __init__(self, al, k=6, ckthr=1, mosthr=-1.0, symmetric="upper",
                 batch=10_000_000, col_block=None, tau=None):
        self.al = al
        self.k = k
        self.ckthr = ckthr
        self.mosthr = mosthr
        self.symmetric = symmetric
        self.batch = int(batch)
        self.col_block = col_block
        self.tau = tau            # learned filter threshold (None: off)
        self._bufs = {}

    def _buf(self, name, nbytes):
        cap, p = self._bufs.get(name, (0, None))
        if cap < nbytes:
            if p is not None:
                self.al.device_free(p)
            p = self.al.device_alloc(max(nbytes, 1))
            self._bufs[name] = (nbytes, p)
        return self._bufs[name][1]

    def close(self):
        for cap, p in self._bufs.values():
            self.al.device_free(p)
        self._bufs = {}

    def blocks(self, ncols):
        step = self.col_block or ncols
        return [(a, min(a + step, ncols)) for a in range(0, ncols, step)]

    def candidates(self, c_begin, c_end):
        """Generates one column block's candidates on the device; returns
        (device pointer, count)."""
        n = self.al.kmer_candidates_dev(
            None, 0, k=self.k, symmetric=self.symmetric, ckthr=self.ckthr,
            mosthr=self.mosthr, c_begin=c_begin, c_end=c_end)
        d = self._buf("cand", max(n, 1) * PAIR_DTYPE.itemsize)
        self.al.kmer_candidates_dev(
            d, n, k=self.k, symmetric=self.symmetric, ckthr=self.ckthr,
            mosthr=self.mosthr, c_begin=c_begin, c_end=c_end)
        return d, n

    def align(self, d_cand, n, d_tri, tri_off):
        """Filter (when on), align and collect triples of n device
        candidates; returns (triples written, pairs aligned)."""
        import ctypes as C
        al = self.al
        b = min(self.batch, max(n, 1))
        d_res = self._buf("res", b * RESULT_DTYPE.itemsize)
        if self.tau is not None:
            d_kept = self._buf("kept", b * PAIR_DTYPE.itemsize)
            d_idx = self._buf("idx", b * 4)
        ntri = aligned = 0
        for off in range(0, n, self.batch):
            m = min(self.batch, n - off)
            d_p = C.c_void_p(d_cand.value + off * PAIR_DTYPE.itemsize)
            if self.tau is not None:
                m = al.filter_select_dev(d_p, m, float(self.tau), d_kept,
                                         d_idx)
                d_p = d_kept
            if m == 0:
                continue
            al.align_dev(d_p, m, d_res)
            aligned += m
            d_t = C.c_void_p(d_tri.value + (tri_off + ntri) *
                             TRIPLE_DTYPE.itemsize)
            ntri += io.triples_dev(al, d_p, d_res, m, d_t)
        return ntri, aligned

    def run(self, ncols=None, keep_candidates=False):
        """Runs every column block; returns (triples, stats)."""
        al = self.al
        ncols = ncols if ncols is not None else self._ncols()
        st = dict(candidates=0, aligned=0, triples=0, gen_s=0.0, align_s=0.0)
        tris, cands = [], []
        for (a, b) in self.blocks(ncols):
            t0 = time.perf_counter()
            d_cand, n = self.candidates(a, b)
            al.synchronize()
            t1 = time.perf_counter()
            d_tri = self._buf("tri", max(n, 1) * TRIPLE_DTYPE.itemsize)
            ntri, aligned = self.align(d_cand, n, d_tri, 0)
            al.synchronize()
            t2 = time.perf_counter()
            st["gen_s"] += t1 - t0
            st["align_s"] += t2 - t1
            st["candidates"] += n
            st["aligned"] += aligned
            st["triples"] += ntri
            t = np.zeros(ntri, TRIPLE_DTYPE)
            if ntri:
                al.d2h(t, d_tri)
            tris.append(t)
            if keep_candidates and n:
                c = np.zeros(n, PAIR_DTYPE)
                al.d2h(c, d_cand)
                cands.append(c)
        out = np.concatenate(tris) if tris else np.zeros(0, TRIPLE_DTYPE)
        if keep_candidates:
            st["candidate_pairs"] = (np.concatenate(cands) if cands
                                     else np.zeros(0, PAIR_DTYPE))
        return out, st

    def _ncols(self):
        raise ValueError("pass ncols (the number of loaded column sequences)")
