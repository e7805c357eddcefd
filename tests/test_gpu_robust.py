"""Workspace, chunking and boundary cases of the HIP path.

- A captured score graph is recaptured after the context's workspace moved
  (a larger eager batch) or its sequence sets changed, and then still gives
  the eager results.
- A trace store far smaller than the batch's trace (PA_TRACE_BYTES) splits
  traced SW and X-drop batches into many chunks; results stay bit-exact, and
  an alignment whose own trace exceeds the store is PA_ERR_LIMIT.
- The C++ operator path with separate row and column sets and nonzero block
  offsets (bl_roffset, bl_coffset of DistPWRunner.cpp:127-130).
- The device entry points align pairs with a bad index or seed as empty
  alignments instead of reading outside the sequence store.
"""
import os

import numpy as np
import pytest

from helpers import (oracle_params, pairs_of, rand_seq, seqset, small_config1,
                     small_config2)
from oracle import oracle as orc
from pastis_amd import _lib, pwalign
from pastis_amd._lib import PAIR_DTYPE, RESULT_DTYPE, SCORE_DTYPE
from pastis_amd.aligner import Aligner, config_for_cli

pytestmark = pytest.mark.gpu


def _scores_eager(al, d_pairs, n, d_out):
    al.sw_scores_dev(d_pairs, n, d_out)
    out = np.zeros(n, SCORE_DTYPE)
    al.synchronize()
    return al.d2h(out, d_out)


def _scores_graph(al, d_pairs, n, d_out):
    al.sw_scores_graph(d_pairs, n, d_out)
    out = np.zeros(n, SCORE_DTYPE)
    al.synchronize()
    return al.d2h(out, d_out)


def test_graph_recaptures_after_workspace_and_sequences_change():
    w = small_config1(n_seq=120, n_pairs=1200, seed=61)
    al = Aligner(config_for_cli(-11, -2, 6, 0, 2, "as_called"))
    al.load_seqs(w.codes, w.offsets)
    n = len(w.pairs)
    d_pairs = al.device_alloc(3 * w.pairs.nbytes)
    d_g = al.device_alloc(3 * n * SCORE_DTYPE.itemsize)
    d_e = al.device_alloc(3 * n * SCORE_DTYPE.itemsize)
    al.h2d(d_pairs, w.pairs)
    print("graph capture at n", flush=True)
    first = _scores_graph(al, d_pairs, n, d_g)
    assert np.array_equal(first, _scores_eager(al, d_pairs, n, d_e))
    caps = al.stats()["graph_captures"]
    _scores_graph(al, d_pairs, n, d_g)
    assert al.stats()["graph_captures"] == caps          # replay only
    # an eager batch three times larger grows (moves) the workspace
    big = np.concatenate([w.pairs] * 3)
    al.h2d(d_pairs, big)
    print("eager batch at 3n", flush=True)
    _scores_eager(al, d_pairs, 3 * n, d_e)
    al.h2d(d_pairs, w.pairs)
    got = _scores_graph(al, d_pairs, n, d_g)
    assert al.stats()["graph_captures"] == caps + 1
    assert np.array_equal(got, first)
    # longer sequences: a different set, larger store and boundary rows
    rng = np.random.default_rng(3)
    seqs = [rand_seq(rng, int(x)) for x in rng.integers(400, 2500, 150)]
    codes, offs = seqset(seqs)
    al.load_seqs(codes, offs)
    p = pairs_of([(int(rng.integers(150)), int(rng.integers(150)))
                  for _ in range(n)])
    al.h2d(d_pairs, p)
    print("graph after reload", flush=True)
    got = _scores_graph(al, d_pairs, n, d_g)
    assert al.stats()["graph_captures"] == caps + 2
    assert np.array_equal(got, _scores_eager(al, d_pairs, n, d_e))
    sc, ve, he = orc.sw_batch(codes, offs, p["r"], p["c"], -2, -11)
    assert np.array_equal(got["score"], sc)
    assert np.array_equal(got["r_end"].astype(np.int32), ve)
    assert np.array_equal(got["c_end"].astype(np.int32), he)
    for b in (d_pairs, d_g, d_e):
        al.device_free(b)


@pytest.fixture
def small_trace_store():
    old = os.environ.get("PA_TRACE_BYTES")
    os.environ["PA_TRACE_BYTES"] = str(1 << 20)
    yield
    if old is None:
        del os.environ["PA_TRACE_BYTES"]
    else:
        os.environ["PA_TRACE_BYTES"] = old


def _cmp(got, ref, fields):
    for gf, rf in fields:
        bad = np.nonzero(got[gf].astype(np.int64) != ref[rf].astype(np.int64))[0]
        assert bad.size == 0, (gf, int(bad[0]), got[bad[0]], ref[bad[0]])
    assert (got["score_aln"] == ref["score_aln"]).all()


def test_chunked_trace_store_sw_and_xdrop(small_trace_store):
    # --sfa: ~0.5-60 KB of trace per pair, ~40 MB per batch -> many chunks
    w = small_config1(n_seq=100, n_pairs=1500, seed=62)
    al = Aligner(config_for_cli(-11, -2, 6, 0, 2, "as_called"))
    al.load_seqs(w.codes, w.offsets)
    print("chunked sfa", flush=True)
    got = al.align(w.pairs)
    ref = orc.sfa_batch(w.codes, w.offsets, w.pairs["r"], w.pairs["c"],
                        oracle_params("as_called"))
    _cmp(got, ref, [("score", "score"), ("r_end", "v_end"),
                    ("c_end", "h_end"), ("r_beg", "v_beg"),
                    ("c_beg", "h_beg"), ("matches", "matches"),
                    ("cols", "cols"), ("gap_opens", "gap_opens")])
    al.close()
    # --sxa: global alignment of the infixes, traced in chunks
    w = small_config2(n_seq=150, n_pairs=1500, seed=63)
    al = Aligner(config_for_cli(-11, -2, 6, 49, 2, "as_called"))
    al.load_seqs(w.codes, w.offsets)
    print("chunked sxa", flush=True)
    got = al.align(w.pairs)
    ref = orc.sxa_batch(w.codes, w.offsets, w.pairs["r"], w.pairs["c"],
                        w.pairs["seed"], oracle_params("as_called"))
    _cmp(got, ref, [("score", "score"), ("matches", "matches"),
                    ("cols", "cols"), ("gap_opens", "gap_opens")])
    al.close()


def test_alignment_larger_than_trace_store_is_limit_error(small_trace_store):
    rng = np.random.default_rng(4)
    codes, offs = seqset([rand_seq(rng, 3000), rand_seq(rng, 3000)])
    al = Aligner(config_for_cli(-11, -2, 6, 0, 2, "as_called"))
    al.load_seqs(codes, offs)
    with pytest.raises(_lib.PaError, match="trace"):
        al.align(pairs_of([(0, 1)]))   # 4.5 MB of trace > 1 MiB
    al.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_block_offsets_separate_row_and_column_sets(mode):
    w = small_config2(n_seq=160, n_pairs=2000, seed=64)
    # rows: sequences [0, 100); columns: sequences [60, 160)
    L = w.lengths
    seqs = [w.codes[int(w.offsets[i]):int(w.offsets[i + 1])]
            for i in range(len(L))]
    rcodes, roffs = seqset(seqs[:100])
    ccodes, coffs = seqset(seqs[60:])
    roff, coff = 17, 9            # block offsets into the two sets
    p = w.pairs
    keep = (p["r"] >= roff) & (p["r"] < 100) & (p["c"] >= 60 + coff)
    p = p[keep]
    key = p["r"].astype(np.uint64) << np.uint64(32) | p["c"]
    _, first = np.unique(key, return_index=True)
    p = p[np.sort(first)]
    assert len(p) > 50
    rows_local = p["r"].astype(np.uint64) - roff
    cols_local = p["c"].astype(np.uint64) - 60 - coff
    print("block offsets mode", mode, len(p), flush=True)
    got, st = pwalign.run_block(
        mode, rcodes, roffs, rows_local, cols_local, p["count"], p["kscore"],
        p["seed"] if mode == 1 else None, ccodes=ccodes, coffsets=coffs,
        bl_roffset=roff, bl_coffset=coff, batch_sz=37)
    prm = oracle_params("as_called")
    if mode == 0:
        ref = orc.sfa_batch(w.codes, w.offsets, p["r"], p["c"], prm)
    else:
        ref = orc.sxa_batch(w.codes, w.offsets, p["r"], p["c"], p["seed"],
                            prm)
    want = np.where(ref["passed"].astype(bool), ref["score_aln"], 0)
    np.testing.assert_array_equal(got, want.astype(np.float32))
    assert st["batches"] == len(p) // 37 + 1
    assert st["nnz_aln"] == int(np.sum(want != 0))


def test_device_entry_points_align_bad_pairs_as_empty():
    w = small_config2(n_seq=80, n_pairs=300, seed=65)
    al = Aligner(config_for_cli(-11, -2, 6, 49, 2, "as_called"))
    al.load_seqs(w.codes, w.offsets)
    p = w.pairs.copy()
    L = w.lengths
    bad_idx = [3, 10]
    p["r"][3] = 10 ** 6                       # not loaded
    p["seed"][10, 0] = int(L[p["r"][10]])     # seed 0 runs past the end
    p["seed"][10, 2] = int(L[p["r"][10]])     # and seed 1 too
    n = len(p)
    d_pairs = al.device_alloc(p.nbytes)
    d_out = al.device_alloc(n * RESULT_DTYPE.itemsize)
    al.h2d(d_pairs, p)
    al.align_dev(d_pairs, n, d_out)
    got = np.zeros(n, RESULT_DTYPE)
    al.synchronize()
    al.d2h(got, d_out)
    for i in bad_idx:
        assert got["score_aln"][i] == 0 and got["cols"][i] == 0, got[i]
        assert not (got["flags"][i] & _lib.PA_FLAG_PASS)
        assert got["r_beg"][i] == got["r_end"][i] == 0
    good = np.setdiff1d(np.arange(n), bad_idx)
    ref = orc.sxa_batch(w.codes, w.offsets, p["r"][good], p["c"][good],
                        p["seed"][good], oracle_params("as_called"))
    assert np.array_equal(got["score"][good], ref["score"])
    assert np.array_equal(got["score_aln"][good], ref["score_aln"])
    # the host entry point still reports them
    with pytest.raises(_lib.PaError):
        al.align(p[bad_idx[:1]])
    with pytest.raises(_lib.PaError):
        al.align(p[bad_idx[1:]])
    # SW mode: an index outside the sets is an empty local alignment
    al2 = Aligner(config_for_cli(-11, -2, 6, 0, 2, "as_called"))
    al2.load_seqs(w.codes, w.offsets)
    d_s = al2.device_alloc(n * SCORE_DTYPE.itemsize)
    d_p2 = al2.device_alloc(p.nbytes)
    al2.h2d(d_p2, p)
    al2.sw_scores_dev(d_p2, n, d_s)
    s = np.zeros(n, SCORE_DTYPE)
    al2.synchronize()
    al2.d2h(s, d_s)
    assert s["score"][3] == 0 and s["r_end"][3] == 0
    for b in (d_pairs, d_out):
        al.device_free(b)
    for b in (d_s, d_p2):
        al2.device_free(b)


def test_graph_per_batch_of_a_resident_pair_array():
    """One graph per batch offset (the --bsz loop over resident pairs):
    each is captured once and replayed; all replays equal eager runs."""
    import ctypes as C
    w = small_config1(n_seq=120, n_pairs=3000, seed=66)
    al = Aligner(config_for_cli(-11, -2, 6, 0, 2, "as_called"))
    al.load_seqs(w.codes, w.offsets)
    n, b = len(w.pairs), 1000
    d_pairs = al.device_alloc(w.pairs.nbytes)
    d_g = al.device_alloc(n * SCORE_DTYPE.itemsize)
    d_e = al.device_alloc(n * SCORE_DTYPE.itemsize)
    al.h2d(d_pairs, w.pairs)
    want = _scores_eager(al, d_pairs, n, d_e)
    caps0 = al.stats()["graph_captures"]
    for rep in range(3):
        print("batched graph replay", rep, flush=True)
        for off in range(0, n, b):
            dp = C.c_void_p(d_pairs.value + off * w.pairs.itemsize)
            do = C.c_void_p(d_g.value + off * SCORE_DTYPE.itemsize)
            al.sw_scores_graph(dp, b, do)
        got = np.zeros(n, SCORE_DTYPE)
        al.synchronize()
        al.d2h(got, d_g)
        assert np.array_equal(got, want), rep
    assert al.stats()["graph_captures"] == caps0 + 3
    for x in (d_pairs, d_g, d_e):
        al.device_free(x)
