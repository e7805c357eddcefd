"""HIP Smith-Waterman path against the CPU oracle, bit-exact.

Scores and end cells (pa_sw_scores) and the full --sfa operator (pa_align:
traceback statistics, identity, predicate, score_aln) on seeded synthetic
batches, hand-made edge cases (length-1 sequences, identical and
multi-pass-long sequences, int16 saturation forcing the int32 re-run) and
both readings of the gap arguments.
"""
import numpy as np
import pytest

from helpers import (GAP_MODELS, oracle_params, pairs_of, rand_seq, seqset,
                     small_config1)
from oracle import oracle as orc
from pastis_amd import _lib
from pastis_amd.aligner import Aligner, config_for_cli

pytestmark = pytest.mark.gpu


def make_aligner(model, xdrop=0):
    cfg = config_for_cli(-11, -2, 6, xdrop, 2, model)
    return Aligner(cfg)


def check_scores(al, codes, offs, pairs, model):
    g = GAP_MODELS[model]
    got = al.sw_scores(pairs)
    sc, ve, he = orc.sw_batch(codes, offs, pairs["r"], pairs["c"],
                              g["gap_open"], g["gap_ext"])
    bad = np.nonzero((got["score"] != sc) | (got["r_end"] != ve) |
                     (got["c_end"] != he))[0]
    assert bad.size == 0, ("first mismatch", int(bad[0]), got[bad[0]],
                           sc[bad[0]], ve[bad[0]], he[bad[0]])


def check_align(al, codes, offs, pairs, model):
    p = oracle_params(model)
    got = al.align(pairs)
    ref = orc.sfa_batch(codes, offs, pairs["r"], pairs["c"], p)
    fields = [("score", "score"), ("r_end", "v_end"), ("c_end", "h_end"),
              ("r_beg", "v_beg"), ("c_beg", "h_beg"), ("matches", "matches"),
              ("cols", "cols"), ("gap_opens", "gap_opens")]
    for gf, rf in fields:
        bad = np.nonzero(got[gf].astype(np.int64) != ref[rf].astype(np.int64))[0]
        assert bad.size == 0, (gf, int(bad[0]), got[bad[0]], ref[bad[0]])
    both_nan = np.isnan(got["identity"]) & np.isnan(ref["identity"])
    assert ((got["identity"] == ref["identity"]) | both_nan).all()
    assert (got["score_aln"] == ref["score_aln"]).all()
    assert ((got["flags"] & _lib.PA_FLAG_PASS).astype(bool)
            == ref["passed"].astype(bool)).all()
    return got


@pytest.mark.parametrize("model", ["as_called", "as_documented"])
def test_sw_scores_synthetic(model):
    w = small_config1()
    al = make_aligner(model)
    al.load_seqs(w.codes, w.offsets)
    check_scores(al, w.codes, w.offsets, w.pairs, model)


@pytest.mark.parametrize("model", ["as_called", "as_documented"])
def test_sfa_operator_synthetic(model):
    w = small_config1(n_seq=100, n_pairs=800, seed=21)
    al = make_aligner(model)
    al.load_seqs(w.codes, w.offsets)
    got = check_align(al, w.codes, w.offsets, w.pairs, model)
    assert (got["flags"] & _lib.PA_FLAG_PASS).any()


def test_sw_edge_cases():
    rng = np.random.default_rng(5)
    s_long = rand_seq(rng, 2600)           # three passes at 1024 rows
    s_mid = rand_seq(rng, 700)
    seqs = [orc.encode("W"), orc.encode("A"), orc.encode("WW"),
            rand_seq(rng, 1), rand_seq(rng, 257), rand_seq(rng, 513),
            s_long, s_long.copy(), s_mid,
            np.concatenate([s_mid[:300], rand_seq(rng, 40), s_mid[300:]]),
            orc.encode("PPPPPPPPPP"), orc.encode("WWWWWWWWWW")]
    codes, offs = seqset(seqs)
    n = len(seqs)
    rc = [(i, j) for i in range(n) for j in range(n)]
    pairs = pairs_of(rc)
    for model in GAP_MODELS:
        al = make_aligner(model)
        al.load_seqs(codes, offs)
        check_scores(al, codes, offs, pairs, model)
        check_align(al, codes, offs, pairs, model)


def test_int16_saturation_reruns_in_int32():
    # identical 3,300-residue tryptophan runs score 36,300 > 32,767
    rng = np.random.default_rng(8)
    s = orc.encode("W" * 3300)
    t = s.copy()
    t[1000] = orc.encode("A")[0]
    codes, offs = seqset([s, t, rand_seq(rng, 900)])
    pairs = pairs_of([(0, 1), (1, 0), (0, 0), (2, 0)])
    al = make_aligner("as_documented")
    al.load_seqs(codes, offs)
    check_scores(al, codes, offs, pairs, "as_documented")
    got = check_align(al, codes, offs, pairs, "as_documented")
    assert got["score"][2] > 32767
    assert got["flags"][2] & _lib.PA_FLAG_RERUN32


def test_empty_batch_and_separate_row_col_sets():
    rng = np.random.default_rng(2)
    rows = [rand_seq(rng, int(x)) for x in rng.integers(20, 400, 30)]
    cols = [rand_seq(rng, int(x)) for x in rng.integers(20, 400, 25)]
    rcodes, roffs = seqset(rows)
    ccodes, coffs = seqset(cols)
    al = make_aligner("as_called")
    al.load_seqs(rcodes, roffs, side=_lib.PA_SIDE_ROW)
    al.load_seqs(ccodes, coffs, side=_lib.PA_SIDE_COL)
    assert len(al.align(pairs_of([]))) == 0
    rc = [(int(rng.integers(30)), int(rng.integers(25))) for _ in range(300)]
    pairs = pairs_of(rc)
    got = al.sw_scores(pairs)
    for k, (r, c) in enumerate(rc):
        o = orc.sw_local(cols[c], rows[r], -2, -11, stats=False)
        assert (got["score"][k], got["r_end"][k], got["c_end"][k]) == \
            (o["score"], o["v_end"], o["h_end"])


def test_bad_pair_index_is_an_error():
    rng = np.random.default_rng(1)
    codes, offs = seqset([rand_seq(rng, 50), rand_seq(rng, 60)])
    al = make_aligner("as_called")
    al.load_seqs(codes, offs)
    with pytest.raises(_lib.PaError):
        al.align(pairs_of([(0, 2)]))


def test_repeat_runs_are_deterministic():
    w = small_config1(n_seq=60, n_pairs=400, seed=33)
    al = make_aligner("as_called")
    al.load_seqs(w.codes, w.offsets)
    a = al.align(w.pairs)
    b = al.align(w.pairs)
    assert a.tobytes() == b.tobytes()
