"""Python face of the aligner: a thin owner of one pa_ctx.

The reference's operator is the C++ class pastis::PWAlign
(inc/align/PWAlign.hpp:20-82); its MI355X replacement is HipAligner
(pastis_amd/csrc/HipAligner.hpp) over the C ABI of include/pastis_hip.h.  This
module drives the same C ABI from Python for tests and benchmarks.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (EXTENT_DTYPE, PAIR_DTYPE, RESULT_DTYPE, SCORE_DTYPE,
                   PaConfig, PaStats, check, lib)

ALPHABET = "ARNDCQEGHILKMFPSTWYVBZX*J"


def config_for_cli(gap_open=-11, gap_ext=-2, klength=6, xdrop=0,
                   seed_count=2, gap_model="as_called"):
    """pa_config from the reference's CLI values (src/main.cpp:158-176)."""
    cfg = PaConfig()
    model = {"as_called": _lib.PA_GAP_AS_CALLED,
             "as_documented": _lib.PA_GAP_AS_DOCUMENTED}[gap_model]
    check(None, lib().pa_config_for_cli(C.byref(cfg), gap_open, gap_ext,
                                        klength, xdrop, seed_count, model),
          "pa_config_for_cli")
    return cfg


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else None


class Aligner:
    """Owns a pa_ctx: sequences are loaded once, batches aligned many times."""

    def __init__(self, cfg=None, **kw):
        self.cfg = cfg if cfg is not None else config_for_cli(**kw)
        self._ctx = C.c_void_p()
        check(None, lib().pa_create(C.byref(self.cfg), C.byref(self._ctx)),
              "pa_create")

    @property
    def ctx(self):
        return self._ctx

    def close(self):
        if self._ctx:
            lib().pa_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_seqs(self, codes, offsets, side=_lib.PA_SIDE_BOTH,
                  encoding=_lib.PA_ENC_CODES):
        codes = np.ascontiguousarray(codes, np.uint8)
        offsets = np.ascontiguousarray(offsets, np.uint64)
        check(self._ctx, lib().pa_load_seqs(self._ctx, side, _ptr(codes),
                                            _ptr(offsets), len(offsets) - 1,
                                            encoding), "pa_load_seqs")

    def align(self, pairs):
        pairs = np.ascontiguousarray(pairs, PAIR_DTYPE)
        out = np.zeros(len(pairs), RESULT_DTYPE)
        check(self._ctx, lib().pa_align(self._ctx, _ptr(pairs), len(pairs),
                                        _ptr(out)), "pa_align")
        return out

    def sw_scores(self, pairs):
        pairs = np.ascontiguousarray(pairs, PAIR_DTYPE)
        out = np.zeros(len(pairs), SCORE_DTYPE)
        check(self._ctx, lib().pa_sw_scores(self._ctx, _ptr(pairs), len(pairs),
                                            _ptr(out)), "pa_sw_scores")
        return out

    def xdrop_extents(self, pairs, seed=0):
        pairs = np.ascontiguousarray(pairs, PAIR_DTYPE)
        out = np.zeros(len(pairs), EXTENT_DTYPE)
        check(self._ctx, lib().pa_xdrop_extents(self._ctx, _ptr(pairs),
                                                len(pairs), seed, _ptr(out)),
              "pa_xdrop_extents")
        return out

    # device-resident batches (benchmarks)
    def device_alloc(self, nbytes):
        p = C.c_void_p()
        check(self._ctx, lib().pa_device_alloc(self._ctx, nbytes, C.byref(p)),
              "pa_device_alloc")
        return p

    def device_free(self, p):
        check(self._ctx, lib().pa_device_free(self._ctx, p), "pa_device_free")

    def h2d(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        check(self._ctx, lib().pa_memcpy_h2d(self._ctx, dst, _ptr(arr),
                                             arr.nbytes), "pa_memcpy_h2d")

    def d2h(self, arr, src):
        check(self._ctx, lib().pa_memcpy_d2h(self._ctx, _ptr(arr), src,
                                             arr.nbytes), "pa_memcpy_d2h")
        return arr

    def align_dev(self, d_pairs, n, d_out):
        check(self._ctx, lib().pa_align_dev(self._ctx, d_pairs, n, d_out, None),
              "pa_align_dev")

    def sw_scores_dev(self, d_pairs, n, d_out):
        check(self._ctx, lib().pa_sw_scores_dev(self._ctx, d_pairs, n, d_out,
                                                None), "pa_sw_scores_dev")

    def sw_scores_noalloc(self, d_pairs, n, d_out):
        check(self._ctx, lib().pa_sw_scores_noalloc(self._ctx, d_pairs, n,
                                                    d_out, None),
              "pa_sw_scores_noalloc")

    def sw_scores_graph(self, d_pairs, n, d_out):
        check(self._ctx, lib().pa_sw_scores_graph(self._ctx, d_pairs, n, d_out,
                                                  None), "pa_sw_scores_graph")

    def align_graph(self, d_pairs, n, d_out):
        """pa_align_graph: the traced operator as a replayed HIP graph."""
        check(self._ctx, lib().pa_align_graph(self._ctx, d_pairs, n, d_out,
                                              None), "pa_align_graph")

    def synchronize(self):
        check(self._ctx, lib().pa_synchronize(self._ctx), "pa_synchronize")

    def stats(self):
        s = PaStats()
        check(self._ctx, lib().pa_get_stats(self._ctx, C.byref(s)),
              "pa_get_stats")
        return s.as_dict()

    def kernel_stats(self):
        arr = (_lib.PaKernelStat * 64)()
        n = C.c_int()
        check(self._ctx, lib().pa_get_kernel_stats(self._ctx, arr, 64,
                                                   C.byref(n)),
              "pa_get_kernel_stats")
        return [dict(name=arr[i].name.decode(), launches=arr[i].launches,
                     cells=arr[i].cells, ms=arr[i].ms)
                for i in range(min(n.value, 64))]

    def reset_stats(self):
        check(self._ctx, lib().pa_reset_stats(self._ctx), "pa_reset_stats")

    def tuning(self):
        t = _lib.PaTuning()
        check(self._ctx, lib().pa_get_tuning(self._ctx, C.byref(t)),
              "pa_get_tuning")
        return t.as_dict()

    def set_tuning(self, **kw):
        t = _lib.PaTuning()
        check(self._ctx, lib().pa_get_tuning(self._ctx, C.byref(t)),
              "pa_get_tuning")
        t.update(**kw)
        check(self._ctx, lib().pa_set_tuning(self._ctx, C.byref(t)),
              "pa_set_tuning")

    def autotune(self, d_pairs, n, target="align", sample=0, force=False):
        """pa_autotune on n device-resident pairs: installs the fastest
        tuning for this context and returns it (a tuning() dict)."""
        t = _lib.PaTuning()
        check(self._ctx, lib().pa_autotune(
            self._ctx, d_pairs, n, _lib.TUNE_TARGETS[target], sample,
            1 if force else 0, C.byref(t)), "pa_autotune")
        return t.as_dict()

    # ---- k-mer candidate generation ---------------------------------------
    def kmer_candidates(self, k=6, stride=1, alphabet="pdefault",
                        symmetric="none", ckthr=0, mosthr=-1.0, row_base=0,
                        col_base=0, c_begin=0, c_end=None):
        """Overlap-matrix nonzeros of the loaded sets (pa_kmer_candidates),
        as PAIR_DTYPE in column-major order."""
        kc = _lib.PaKmerConfig(k, stride, _lib.ALPHABETS[alphabet],
                               _lib.SYMMETRIC[symmetric], ckthr, mosthr,
                               row_base, col_base)
        ce = (1 << 63) if c_end is None else c_end
        n = C.c_uint64()
        check(self._ctx, lib().pa_kmer_candidates(
            self._ctx, C.byref(kc), c_begin, ce, None, 0, C.byref(n)),
              "pa_kmer_candidates")
        out = np.zeros(n.value, PAIR_DTYPE)
        if n.value == 0:
            return out
        d = self.device_alloc(out.nbytes)
        try:
            check(self._ctx, lib().pa_kmer_candidates(
                self._ctx, C.byref(kc), c_begin, ce, d, n.value, C.byref(n)),
                  "pa_kmer_candidates")
            self.d2h(out, d)
        finally:
            self.device_free(d)
        return out

    def kmer_candidates_dev(self, d_out, cap, k=6, stride=1,
                            alphabet="pdefault", symmetric="none", ckthr=0,
                            mosthr=-1.0, row_base=0, col_base=0, c_begin=0,
                            c_end=None):
        """Device variant: fills d_out (cap pairs); returns the count
        (raises when it exceeds cap)."""
        kc = _lib.PaKmerConfig(k, stride, _lib.ALPHABETS[alphabet],
                               _lib.SYMMETRIC[symmetric], ckthr, mosthr,
                               row_base, col_base)
        ce = (1 << 63) if c_end is None else c_end
        n = C.c_uint64()
        check(self._ctx, lib().pa_kmer_candidates(
            self._ctx, C.byref(kc), c_begin, ce, d_out, cap, C.byref(n)),
              "pa_kmer_candidates")
        return n.value

    # ---- learned prefilter ------------------------------------------------
    def filter_features(self, pairs):
        pairs = np.ascontiguousarray(pairs, PAIR_DTYPE)
        out = np.zeros((len(pairs), _lib.PA_FILTER_NF), np.float32)
        check(self._ctx, lib().pa_filter_features(self._ctx, _ptr(pairs),
                                                  len(pairs), _ptr(out)),
              "pa_filter_features")
        return out

    def filter_set_model(self, params, mean=None, istd=None):
        params = np.ascontiguousarray(params, np.float32)
        assert params.shape == (_lib.PA_FILTER_NPARAM,)
        mean = None if mean is None else np.ascontiguousarray(mean, np.float32)
        istd = None if istd is None else np.ascontiguousarray(istd, np.float32)
        check(self._ctx, lib().pa_filter_set_model(self._ctx, _ptr(params),
                                                   _ptr(mean), _ptr(istd)),
              "pa_filter_set_model")

    def filter_get_model(self):
        p = np.zeros(_lib.PA_FILTER_NPARAM, np.float32)
        m = np.zeros(_lib.PA_FILTER_NF, np.float32)
        s = np.zeros(_lib.PA_FILTER_NF, np.float32)
        check(self._ctx, lib().pa_filter_get_model(self._ctx, _ptr(p), _ptr(m),
                                                   _ptr(s)),
              "pa_filter_get_model")
        return p, m, s

    def filter_grad(self, feats, labels, pos_weight=1.0):
        feats = np.ascontiguousarray(feats, np.float32)
        labels = np.ascontiguousarray(labels, np.uint8)
        g = np.zeros(_lib.PA_FILTER_NPARAM, np.float32)
        loss = C.c_float()
        check(self._ctx, lib().pa_filter_grad(self._ctx, _ptr(feats),
                                              _ptr(labels), len(labels),
                                              C.c_float(pos_weight), _ptr(g),
                                              C.byref(loss)),
              "pa_filter_grad")
        return float(loss.value), g

    def filter_train(self, feats, labels, epochs=1, batch=4096, seed=1,
                     lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8,
                     pos_weight=1.0, standardize=True):
        feats = np.ascontiguousarray(feats, np.float32)
        labels = np.ascontiguousarray(labels, np.uint8)
        tc = _lib.PaTrainConfig(epochs, batch, seed, lr, beta1, beta2, eps,
                                pos_weight, int(bool(standardize)))
        losses = np.zeros(epochs, np.float32)
        check(self._ctx, lib().pa_filter_train(self._ctx, _ptr(feats),
                                               _ptr(labels), len(labels),
                                               C.byref(tc), _ptr(losses)),
              "pa_filter_train")
        return losses

    def filter_logits(self, feats):
        feats = np.ascontiguousarray(feats, np.float32)
        out = np.zeros(len(feats), np.float32)
        check(self._ctx, lib().pa_filter_logits(self._ctx, _ptr(feats),
                                                len(feats), _ptr(out)),
              "pa_filter_logits")
        return out

    def filter_score_dev(self, d_pairs, n, d_logits):
        check(self._ctx, lib().pa_filter_score_dev(self._ctx, d_pairs, n,
                                                   d_logits, None),
              "pa_filter_score_dev")

    def filter_select_dev(self, d_pairs, n, tau, d_kept, d_kept_index=None):
        kept = C.c_uint64()
        check(self._ctx, lib().pa_filter_select_dev(
            self._ctx, d_pairs, n, C.c_float(tau), d_kept, d_kept_index,
            C.byref(kept), None), "pa_filter_select_dev")
        return kept.value


def kmer_prune(pairs, ckthr, mosthr):
    """DistPWRunner.cpp:48-54 k-mer prefilter: boolean keep mask."""
    pairs = np.ascontiguousarray(pairs, PAIR_DTYPE)
    keep = np.zeros(len(pairs), np.uint8)
    kept = C.c_uint64()
    check(None, lib().pa_kmer_prune(_ptr(pairs), len(pairs), ckthr,
                                    C.c_float(mosthr), _ptr(keep),
                                    C.byref(kept)), "pa_kmer_prune")
    return keep.astype(bool)
