"""ctypes binding of libpastis_hip.so (the C ABI of include/pastis_hip.h).

The library is built in-tree by __graft_entry__.build().  There is no fallback:
a missing or unloadable library raises immediately.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpastis_hip.so")
GEN_PATH = os.path.join(HERE, "libpastis_gen.so")

PA_OK = 0
PA_MODE_SW, PA_MODE_XDROP = 0, 1
PA_SIDE_ROW, PA_SIDE_COL, PA_SIDE_BOTH = 0, 1, 2
PA_ENC_CODES, PA_ENC_ASCII = 0, 1
PA_GAP_AS_CALLED, PA_GAP_AS_DOCUMENTED = 0, 1
PA_FLAG_PASS, PA_FLAG_RERUN32, PA_FLAG_EMPTY, PA_FLAG_NOTRACE = 1, 2, 4, 8

# Every extern "C" function of include/pastis_hip.h.
EXPORTS = [
    "pa_create", "pa_config_for_cli", "pa_destroy", "pa_strerror",
    "pa_last_error", "pa_load_seqs", "pa_align", "pa_align_dev",
    "pa_sw_scores", "pa_sw_scores_dev", "pa_xdrop_extents", "pa_kmer_prune",
    "pa_device_alloc", "pa_device_free", "pa_host_alloc", "pa_host_free",
    "pa_memcpy_h2d", "pa_memcpy_d2h",
    "pa_synchronize", "pa_get_stats", "pa_reset_stats",
    "pa_get_kernel_stats", "pa_filter_features", "pa_filter_features_dev",
    "pa_filter_set_model", "pa_filter_get_model", "pa_filter_grad",
    "pa_filter_train", "pa_filter_logits", "pa_filter_score_dev",
    "pa_filter_select_dev", "pa_set_thresholds", "pa_get_tuning",
    "pa_set_tuning", "pa_autotune", "pa_sw_scores_graph", "pa_align_graph",
    "pa_kmer_candidates",
    "pa_fasta_parse", "pa_fasta_info", "pa_fasta_residues",
    "pa_fasta_offsets", "pa_fasta_records", "pa_fasta_free", "pa_load_fasta",
    "pa_triples_dev", "pa_write_mm", "pa_sw_scores_noalloc",
]

TRIPLE_DTYPE = np.dtype([("row", "<u8"), ("col", "<u8"),
                         ("score_aln", "<f4"), ("pad", "<u4")])

ALPHABETS = {"pdefault": 0, "murphy10": 1, "dssp10": 2, "gbmr10": 3,
             "td10": 4, "diamond": 5}
SYMMETRIC = {"none": 0, "upper": 1, "even_odd": 2}

PA_FILTER_NF = 16
PA_FILTER_NPARAM = 5313

PAIR_DTYPE = np.dtype([("r", "<u4"), ("c", "<u4"), ("count", "<u2"),
                       ("kscore", "<i2"), ("seed", "<u2", (4,))])
RESULT_DTYPE = np.dtype([("score", "<i4"), ("r_beg", "<u2"), ("r_end", "<u2"),
                         ("c_beg", "<u2"), ("c_end", "<u2"),
                         ("matches", "<u4"), ("cols", "<u4"),
                         ("gap_opens", "<u4"), ("identity", "<f4"),
                         ("score_aln", "<f4"), ("flags", "<u2"),
                         ("seed", "<u2")])
SCORE_DTYPE = np.dtype([("score", "<i4"), ("r_end", "<u2"), ("c_end", "<u2")])
EXTENT_DTYPE = np.dtype([("r_beg", "<u2"), ("r_end", "<u2"), ("c_beg", "<u2"),
                         ("c_end", "<u2")])


class PaConfig(C.Structure):
    _fields_ = [("mode", C.c_int32), ("gap_open", C.c_int32),
                ("gap_ext", C.c_int32), ("xdrop", C.c_int32),
                ("xdrop_gap", C.c_int32), ("seed_len", C.c_int32),
                ("seed_cnt", C.c_int32), ("ani_thr", C.c_int32),
                ("cov_thr", C.c_double), ("device", C.c_int32),
                ("flags", C.c_int32), ("matrix", C.c_void_p)]


class PaStats(C.Structure):
    _fields_ = [("pairs", C.c_uint64), ("dp_cells", C.c_uint64),
                ("xdrop_cells", C.c_uint64), ("rerun_pairs", C.c_uint64),
                ("ms_sw", C.c_double), ("ms_xdrop", C.c_double),
                ("ms_nw", C.c_double), ("ms_trace", C.c_double),
                ("ms_total", C.c_double), ("launches", C.c_uint64),
                ("graph_captures", C.c_uint64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class PaKernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 24), ("launches", C.c_uint64),
                ("cells", C.c_uint64), ("ms", C.c_double)]


DP_SHAPES = {"l16r4": 0, "l16r8": 1, "l64r4": 2, "l64r8": 3, "l32r8": 4,
             "l8r8": 5}
TUNE_TARGETS = {"sw_scores": 0, "align": 1}


class PaTuning(C.Structure):
    _fields_ = [("dp_wg_per_cu", C.c_int32), ("xd_wg_per_cu", C.c_int32),
                ("xd_ring", C.c_int32), ("xd_coop", C.c_int32),
                ("dp_edge", C.c_int32 * 3), ("dp_shape", C.c_int32 * 4),
                ("tb_stage", C.c_int32), ("xd_wave", C.c_int32),
                ("tb_overlap", C.c_int32)]

    def as_dict(self):
        """Scalars as ints, dp_edge / dp_shape as tuples."""
        out = {}
        for n, _ in self._fields_:
            v = getattr(self, n)
            out[n] = tuple(v) if n in ("dp_edge", "dp_shape") else v
        return out

    def update(self, **kw):
        for k, v in kw.items():
            if k in ("dp_edge", "dp_shape"):
                arr = getattr(self, k)
                if len(v) != len(arr):
                    raise ValueError("%s takes %d values" % (k, len(arr)))
                for i, x in enumerate(v):
                    arr[i] = DP_SHAPES[x] if isinstance(x, str) else int(x)
            elif k in dict(self._fields_):
                setattr(self, k, int(v))
            else:
                raise KeyError("no tuning field %r" % k)


class PaKmerConfig(C.Structure):
    _fields_ = [("k", C.c_int32), ("stride", C.c_int32),
                ("alphabet", C.c_int32), ("symmetric", C.c_int32),
                ("ckthr", C.c_int32), ("mosthr", C.c_float),
                ("row_base", C.c_uint64), ("col_base", C.c_uint64)]


class PaTrainConfig(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("batch", C.c_uint32),
                ("seed", C.c_uint64), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float),
                ("pos_weight", C.c_float), ("standardize", C.c_int32)]


assert PAIR_DTYPE.itemsize == 20 and RESULT_DTYPE.itemsize == 36
assert TRIPLE_DTYPE.itemsize == 24
assert SCORE_DTYPE.itemsize == 8 and EXTENT_DTYPE.itemsize == 8

_lib = None


class PaError(RuntimeError):
    pass


def lib():
    """Load libpastis_hip.so; raises if it is missing (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libpastis_hip.so is not built: run "
                          "__graft_entry__.build() (expected at %s)" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, u64, i32 = C.c_void_p, C.c_uint64, C.c_int
    L.pa_create.argtypes = [C.POINTER(PaConfig), C.POINTER(vp)]
    L.pa_config_for_cli.argtypes = [C.POINTER(PaConfig), i32, i32, i32, i32,
                                    i32, i32]
    L.pa_destroy.argtypes = [vp]
    L.pa_destroy.restype = None
    L.pa_strerror.argtypes = [i32]
    L.pa_strerror.restype = C.c_char_p
    L.pa_last_error.argtypes = [vp]
    L.pa_last_error.restype = C.c_char_p
    L.pa_load_seqs.argtypes = [vp, i32, vp, vp, u64, i32]
    for f in ("pa_align", "pa_sw_scores"):
        getattr(L, f).argtypes = [vp, vp, u64, vp]
    for f in ("pa_align_dev", "pa_sw_scores_dev", "pa_sw_scores_graph",
              "pa_align_graph",
              "pa_sw_scores_noalloc"):
        getattr(L, f).argtypes = [vp, vp, u64, vp, vp]
    L.pa_xdrop_extents.argtypes = [vp, vp, u64, i32, vp]
    L.pa_kmer_prune.argtypes = [vp, u64, i32, C.c_float, vp,
                                C.POINTER(C.c_uint64)]
    L.pa_device_alloc.argtypes = [vp, C.c_size_t, C.POINTER(vp)]
    L.pa_device_free.argtypes = [vp, vp]
    L.pa_memcpy_h2d.argtypes = [vp, vp, vp, C.c_size_t]
    L.pa_memcpy_d2h.argtypes = [vp, vp, vp, C.c_size_t]
    L.pa_synchronize.argtypes = [vp]
    L.pa_get_stats.argtypes = [vp, C.POINTER(PaStats)]
    L.pa_reset_stats.argtypes = [vp]
    L.pa_get_kernel_stats.argtypes = [vp, C.POINTER(PaKernelStat), i32,
                                      C.POINTER(C.c_int)]
    L.pa_set_thresholds.argtypes = [vp, C.c_double, i32]
    L.pa_get_tuning.argtypes = [vp, C.POINTER(PaTuning)]
    L.pa_set_tuning.argtypes = [vp, C.POINTER(PaTuning)]
    L.pa_autotune.argtypes = [vp, vp, C.c_uint64, C.c_int, C.c_uint64,
                              C.c_uint32, C.POINTER(PaTuning)]
    fp = C.POINTER(C.c_float)
    L.pa_filter_features.argtypes = [vp, vp, u64, vp]
    L.pa_filter_features_dev.argtypes = [vp, vp, u64, vp, vp]
    L.pa_filter_set_model.argtypes = [vp, vp, vp, vp]
    L.pa_filter_get_model.argtypes = [vp, vp, vp, vp]
    L.pa_filter_grad.argtypes = [vp, vp, vp, u64, C.c_float, vp, fp]
    L.pa_filter_train.argtypes = [vp, vp, vp, u64, C.POINTER(PaTrainConfig),
                                  vp]
    L.pa_filter_logits.argtypes = [vp, vp, u64, vp]
    L.pa_filter_score_dev.argtypes = [vp, vp, u64, vp, vp]
    L.pa_filter_select_dev.argtypes = [vp, vp, u64, C.c_float, vp, vp,
                                       C.POINTER(C.c_uint64), vp]
    L.pa_kmer_candidates.argtypes = [vp, C.POINTER(PaKmerConfig), u64, u64,
                                     vp, u64, C.POINTER(C.c_uint64)]
    L.pa_fasta_parse.argtypes = [vp, u64, i32, C.POINTER(vp)]
    L.pa_fasta_info.argtypes = [vp, C.POINTER(C.c_uint64),
                                C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    for f in ("pa_fasta_residues", "pa_fasta_offsets", "pa_fasta_records"):
        getattr(L, f).argtypes = [vp]
        getattr(L, f).restype = vp
    L.pa_fasta_free.argtypes = [vp]
    L.pa_fasta_free.restype = None
    L.pa_load_fasta.argtypes = [vp, i32, vp, u64, i32]
    L.pa_triples_dev.argtypes = [vp, vp, vp, u64, u64, u64, vp,
                                 C.POINTER(C.c_uint64), vp]
    L.pa_write_mm.argtypes = [C.c_char_p, u64, u64, vp, u64]
    _lib = L
    return L


def check(ctx, rc, what):
    if rc != PA_OK:
        L = lib()
        msg = L.pa_strerror(rc).decode()
        if ctx:
            detail = L.pa_last_error(ctx).decode()
            if detail:
                msg += ": " + detail
        raise PaError("%s failed (%d): %s" % (what, rc, msg))
