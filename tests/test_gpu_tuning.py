"""Launch shapes change speed only.

Every layout of the DP row classes (class edges, a kernel shape per class,
traceback staging) gives the oracle's results on the score-only, traced
local and X-drop paths, including multi-pass tasks and int32 re-runs;
pa_autotune installs a valid choice, caches it per device and changes no
result; and the operator's double-buffered host staging writes the same
tuple values whatever its chunk size.
"""
import numpy as np
import pytest

from helpers import (oracle_params, oracle_sw, rand_seq, seqset,
                     small_config1, small_config2)
from oracle import oracle as orc
from pastis_amd import _lib, pwalign
from pastis_amd.aligner import Aligner, config_for_cli

pytestmark = pytest.mark.gpu

# (dp_shape per class, dp_edge, tb_stage, tb_overlap); shapes 0..5 = l16r4,
# l16r8, l64r4, l64r8, l32r8, l8r8
LAYOUTS = [
    ((5, 1, 4, 3), (128, 256, 512), 1, 0),   # the defaults
    ((0, 1, 2, 3), (128, 256, 512), 0, 1),   # 4-row shapes on short rows
    ((0, 0, 0, 0), (16, 40, 100), 1, 2),     # 128-row passes everywhere
    ((3, 3, 3, 3), (128, 256, 512), 0, 0),   # 1024-row passes everywhere
    ((3, 2, 1, 0), (16, 40, 100), 1, 1),     # wide on short, narrow on long
    ((1, 0, 3, 2), (200, 300, 1000), 1, 0),
    ((5, 1, 4, 3), (128, 256, 512), 0, 2),   # the defaults, walks unstaged
    ((4, 5, 4, 5), (16, 40, 100), 1, 1),     # 32- and 8-lane multi-pass tasks
    ((5, 4, 5, 4), (16, 40, 100), 0, 2),
    ((5, 1, 4, 3), (128, 256, 512), 1, 2),   # walks beside the last DP
]


def _with_long(w, seed):
    """The workload's sequences plus a 1,500- and a 2,600-residue sequence
    and a near-identical copy of the latter (its local score saturates the
    traced int16 kernels), with pairs among them appended."""
    rng = np.random.default_rng(seed)
    n0 = len(w.offsets) - 1
    seqs = [w.codes[int(w.offsets[i]):int(w.offsets[i + 1])]
            for i in range(n0)]
    a = rand_seq(rng, 1500)
    b = rand_seq(rng, 2600)
    c = b.copy()
    c[::97] = rand_seq(rng, len(c[::97]))
    codes, offs = seqset(seqs + [a, b, c])
    extra = np.zeros(5, w.pairs.dtype)
    for i, (r, cc) in enumerate([(n0, n0 + 1), (n0 + 1, n0 + 2),
                                 (n0 + 2, n0 + 1), (n0, 3), (5, n0 + 2)]):
        extra["r"][i], extra["c"][i], extra["count"][i] = r, cc, 2
    return codes, offs, np.concatenate([w.pairs, extra])


def _set(al, layout):
    sh, ed, st, ov = layout
    al.set_tuning(dp_shape=sh, dp_edge=ed, tb_stage=st, tb_overlap=ov)
    t = al.tuning()
    assert t["dp_shape"] == sh and t["dp_edge"] == ed and \
        t["tb_stage"] == st and t["tb_overlap"] == ov


def test_sw_layouts_match_the_oracle():
    w = small_config1(n_seq=80, n_pairs=300, seed=81)
    codes, offs, pairs = _with_long(w, 81)
    go, ge = -2, -11
    ref = oracle_sw(codes, offs, pairs, go, ge)
    for layout in LAYOUTS:
        print("sw layout", layout, flush=True)
        al = Aligner(config_for_cli(-11, -2, 6, 0, 2, "as_called"))
        _set(al, layout)
        al.load_seqs(codes, offs)
        al.reset_stats()
        got = al.sw_scores(pairs)
        assert np.array_equal(got["score"], ref[0]), layout
        assert np.array_equal(got["r_end"].astype(np.int32), ref[1]), layout
        assert np.array_equal(got["c_end"].astype(np.int32), ref[2]), layout
        al.close()


@pytest.mark.parametrize("model", ["as_called", "as_documented"])
def test_sfa_layouts_match_the_oracle(model):
    w = small_config1(n_seq=80, n_pairs=300, seed=82)
    codes, offs, pairs = _with_long(w, 82)
    ref = orc.sfa_batch(codes, offs, pairs["r"], pairs["c"],
                        oracle_params(model))
    for layout in LAYOUTS:
        print("sfa layout", model, layout, flush=True)
        al = Aligner(config_for_cli(-11, -2, 6, 0, 2, model))
        _set(al, layout)
        al.load_seqs(codes, offs)
        al.reset_stats()
        got = al.align(pairs)
        for gf, rf in [("score", "score"), ("r_end", "v_end"),
                       ("c_end", "h_end"), ("r_beg", "v_beg"),
                       ("c_beg", "h_beg"), ("matches", "matches"),
                       ("cols", "cols"), ("gap_opens", "gap_opens")]:
            assert np.array_equal(got[gf].astype(np.int64),
                                  ref[rf].astype(np.int64)), (layout, gf)
        assert np.array_equal(got["score_aln"], ref["score_aln"]), layout
        if model == "as_called":
            # the near-identical 2,600-residue pair re-runs in int32
            assert al.stats()["rerun_pairs"] >= 1, layout
        al.close()


@pytest.mark.parametrize("model", ["as_called", "as_documented"])
def test_sxa_layouts_match_the_oracle(model):
    w = small_config2(n_seq=120, n_pairs=400, seed=83)
    ref = orc.sxa_batch(w.codes, w.offsets, w.pairs["r"], w.pairs["c"],
                        w.pairs["seed"], oracle_params(model))
    for layout in LAYOUTS:
        print("sxa layout", model, layout, flush=True)
        al = Aligner(config_for_cli(-11, -2, 6, 49, 2, model))
        _set(al, layout)
        al.load_seqs(w.codes, w.offsets)
        got = al.align(w.pairs)
        for gf in ("score", "matches", "cols", "gap_opens"):
            assert np.array_equal(got[gf].astype(np.int64),
                                  ref[gf].astype(np.int64)), (layout, gf)
        assert np.array_equal(got["score_aln"], ref["score_aln"]), layout
        al.close()


# Scoring matrices other than BLOSUM62: twice BLOSUM62 (-8 .. 22) and
# BLOSUM62 + 4 (0 .. 15) do not fit the 4-bit profiles of the 8-rows-per-lane
# kernels, which the context then replaces by their 4-row counterparts;
# BLOSUM62 - 1 (-5 .. 10) fits with another minimum.
MATRICES = {"x2": lambda m: 2 * m, "plus4": lambda m: m + 4,
            "minus1": lambda m: m - 1}


def _mat_cfg(cfg, m):
    cfg.matrix = m.ctypes.data
    return cfg


@pytest.mark.parametrize("name", sorted(MATRICES))
def test_other_matrices_match_the_oracle(name):
    base = orc.blosum62().astype(np.int16)
    m = np.ascontiguousarray(MATRICES[name](base).astype(np.int8))
    w = small_config1(n_seq=80, n_pairs=300, seed=85)
    codes, offs, pairs = _with_long(w, 85)
    go, ge = -2, -11
    ref_sw = orc.sw_batch(codes, offs, pairs["r"], pairs["c"], go, ge, mat=m)
    ref_sfa = orc.sfa_batch(codes, offs, pairs["r"], pairs["c"],
                            oracle_params("as_called"), mat=m)
    w2 = small_config2(n_seq=120, n_pairs=300, seed=86)
    ref_sxa = orc.sxa_batch(w2.codes, w2.offsets, w2.pairs["r"],
                            w2.pairs["c"], w2.pairs["seed"],
                            oracle_params("as_called"), mat=m)
    for layout in (LAYOUTS[0], LAYOUTS[1], LAYOUTS[8]):
        print("matrix", name, "layout", layout, flush=True)
        al = Aligner(_mat_cfg(config_for_cli(-11, -2, 6, 0, 2, "as_called"),
                              m))
        _set(al, layout)
        al.load_seqs(codes, offs)
        got = al.sw_scores(pairs)
        assert np.array_equal(got["score"], ref_sw[0]), layout
        assert np.array_equal(got["r_end"].astype(np.int32), ref_sw[1])
        assert np.array_equal(got["c_end"].astype(np.int32), ref_sw[2])
        got = al.align(pairs)
        for gf, rf in [("score", "score"), ("r_end", "v_end"),
                       ("c_end", "h_end"), ("r_beg", "v_beg"),
                       ("c_beg", "h_beg"), ("matches", "matches"),
                       ("cols", "cols"), ("gap_opens", "gap_opens")]:
            assert np.array_equal(got[gf].astype(np.int64),
                                  ref_sfa[rf].astype(np.int64)), (layout, gf)
        al.close()
        al = Aligner(_mat_cfg(config_for_cli(-11, -2, 6, 49, 2, "as_called"),
                              m))
        _set(al, layout)
        al.load_seqs(w2.codes, w2.offsets)
        got = al.align(w2.pairs)
        for gf in ("score", "matches", "cols", "gap_opens"):
            assert np.array_equal(got[gf].astype(np.int64),
                                  ref_sxa[gf].astype(np.int64)), (layout, gf)
        assert np.array_equal(got["score_aln"], ref_sxa["score_aln"])
        al.close()


def test_bad_layouts_are_rejected():
    al = Aligner(config_for_cli(-11, -2, 6, 49, 2, "as_called"))
    for kw in [dict(dp_edge=(128, 128, 512)), dict(dp_edge=(0, 10, 20)),
               dict(dp_edge=(1, 2, 70000)), dict(dp_shape=(0, 1, 2, 6)),
               dict(tb_stage=2), dict(tb_overlap=3), dict(tb_overlap=-1)]:
        with pytest.raises(_lib.PaError):
            al.set_tuning(**kw)
    al.close()


@pytest.mark.parametrize("target", ["align", "sw_scores"])
def test_autotune_changes_no_result_and_is_cached(target):
    xdrop = 49 if target == "align" else 0
    w = (small_config2(n_seq=150, n_pairs=3000, seed=84) if xdrop else
         small_config1(n_seq=150, n_pairs=3000, seed=84))
    cfg = config_for_cli(-11, -2, 6, xdrop, 2, "as_called")
    run = (lambda a: a.align(w.pairs)) if xdrop else \
        (lambda a: a.sw_scores(w.pairs))
    al = Aligner(cfg)
    al.load_seqs(w.codes, w.offsets)
    base = run(al)
    dp = al.device_alloc(w.pairs.nbytes)
    al.h2d(dp, w.pairs)
    print("autotune", target, flush=True)
    t = al.autotune(dp, len(w.pairs), target, force=True)
    assert t == al.tuning()
    assert run(al).tobytes() == base.tobytes()
    al.device_free(dp)
    # a second context on the device installs the cached choice
    al2 = Aligner(cfg)
    al2.load_seqs(w.codes, w.offsets)
    dp2 = al2.device_alloc(w.pairs.nbytes)
    al2.h2d(dp2, w.pairs)
    assert al2.autotune(dp2, len(w.pairs), target) == t
    assert al2.tuning() == t
    assert run(al2).tobytes() == base.tobytes()
    al2.device_free(dp2)
    al2.close()
    al.close()


def _unique(pairs):
    key = pairs["r"].astype(np.uint64) << np.uint64(32) | pairs["c"]
    _, first = np.unique(key, return_index=True)
    return pairs[np.sort(first)]


@pytest.mark.parametrize("chunk", ["7", "100", "1000000"])
def test_operator_staging_chunks(monkeypatch, chunk):
    """HipAligner aligns chunk i while a host thread writes chunk i-1 back
    and packs chunk i+1: any chunk size gives the oracle's values."""
    from pastis_amd.prefilter import init_params
    monkeypatch.setenv("PA_HOST_CHUNK", chunk)
    w = small_config2(n_seq=120, n_pairs=900, seed=85)
    pairs = _unique(w.pairs)
    ref = orc.sxa_batch(w.codes, w.offsets, pairs["r"], pairs["c"],
                        pairs["seed"], oracle_params("as_called"))
    want = np.where(ref["passed"].astype(bool), ref["score_aln"],
                    0).astype(np.float32)
    print("operator chunk", chunk, len(pairs), flush=True)
    args = (1, w.codes, w.offsets, pairs["r"], pairs["c"], pairs["count"],
            pairs["kscore"], pairs["seed"])
    got, st = pwalign.run_block(*args, batch_sz=333)
    np.testing.assert_array_equal(got, want)
    assert st["passed"] == int(np.sum(want != 0))
    # with the learned filter keeping every pair (index-mapped write-back)
    filt = (init_params(0), np.zeros(16, np.float32),
            np.ones(16, np.float32), np.float32(-np.inf))
    got, st = pwalign.run_block(*args, batch_sz=333, filt=filt)
    np.testing.assert_array_equal(got, want)


def test_operator_timing_entry_point():
    w = small_config2(n_seq=120, n_pairs=900, seed=86)
    ref = orc.sxa_batch(w.codes, w.offsets, w.pairs["r"], w.pairs["c"],
                        w.pairs["seed"], oracle_params("as_called"))
    ms, passed = pwalign.bench_aln_batch(1, w.codes, w.offsets, w.pairs,
                                         autotune=False, warmup=1, steps=2)
    assert ms > 0
    assert passed == int(ref["passed"].astype(bool).sum())
