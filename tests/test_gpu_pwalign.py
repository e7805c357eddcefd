"""The C++ operator path (HipAligner + pw_aln_batch) against the oracle.

Each nonzero's final score_aln must equal what the reference path writes:
0 when the k-mer prune (count <= ckthr || score <= mosthr) removes it,
identity/100 of the oracle's alignment when the coverage/identity predicate
holds, else 0 (and then the post-prune removes it).  Batch counts follow
batch_cnt = l_nnz / bsz + 1 (DistPWRunner.cpp:116-131).
"""
import numpy as np
import pytest

from helpers import oracle_params, small_config1, small_config2
from oracle import oracle as orc
from pastis_amd import _lib, pwalign

pytestmark = pytest.mark.gpu


def _unique(pairs):
    key = pairs["r"].astype(np.uint64) << np.uint64(32) | pairs["c"]
    _, first = np.unique(key, return_index=True)
    return pairs[np.sort(first)]


def _expected(mode, w, pairs, model, ckthr, mosthr, cov_thr=0.7, ani_thr=30):
    keep = ~((pairs["count"] <= ckthr) |
             (pairs["kscore"].astype(np.float32) <= np.float32(mosthr)))
    p = oracle_params(model)
    p.cov_thr, p.ani_thr = cov_thr, ani_thr
    if mode == 0:
        ref = orc.sfa_batch(w.codes, w.offsets, pairs["r"], pairs["c"], p)
    else:
        ref = orc.sxa_batch(w.codes, w.offsets, pairs["r"], pairs["c"],
                            pairs["seed"], p)
    want = np.where(keep & ref["passed"].astype(bool), ref["score_aln"], 0)
    return want.astype(np.float32), int(keep.sum())


def _run(mode, w, pairs, model, **kw):
    return pwalign.run_block(mode, w.codes, w.offsets, pairs["r"], pairs["c"],
                             pairs["count"], pairs["kscore"],
                             pairs["seed"] if mode == 1 else None,
                             gap_model=model, **kw)


@pytest.mark.parametrize("model", ["as_called", "as_documented"])
def test_sfa_block(model):
    w = small_config1(n_seq=100, n_pairs=900, seed=41)
    pairs = _unique(w.pairs)
    rng = np.random.default_rng(0)
    pairs["count"] = rng.integers(1, 5, len(pairs))
    pairs["kscore"] = rng.integers(-5, 40, len(pairs))
    print("sfa block", model, len(pairs), flush=True)
    got, st = _run(0, w, pairs, model, ckthr=1, mosthr=3.0, batch_sz=97)
    want, nk = _expected(0, w, pairs, model, 1, 3.0)
    np.testing.assert_array_equal(got, want)
    assert st["nnz_in"] == len(pairs) and st["nnz_kmer"] == nk
    assert st["batches"] == nk // 97 + 1
    assert st["nnz_aln"] == int(np.sum(want != 0)) == st["passed"]
    assert st["aligned"] == nk == st["aln_pairs"]


@pytest.mark.parametrize("model", ["as_called", "as_documented"])
def test_sxa_block(model):
    w = small_config2(n_seq=120, n_pairs=900, seed=42)
    pairs = _unique(w.pairs)
    print("sxa block", model, len(pairs), flush=True)
    got, st = _run(1, w, pairs, model, batch_sz=128)
    want, nk = _expected(1, w, pairs, model, 0, -1.0)
    np.testing.assert_array_equal(got, want)
    assert st["batches"] == nk // 128 + 1
    assert st["nnz_aln"] == int(np.sum(want != 0))


def test_batch_size_dividing_nnz_gives_an_empty_last_batch():
    w = small_config2(n_seq=60, n_pairs=200, seed=43)
    pairs = _unique(w.pairs)[:100]
    got, st = _run(1, w, pairs, "as_called", batch_sz=25)
    want, nk = _expected(1, w, pairs, "as_called", 0, -1.0)
    assert nk == 100 and st["batches"] == 5
    np.testing.assert_array_equal(got, want)


def test_thresholds_come_from_params():
    w = small_config2(n_seq=80, n_pairs=400, seed=44)
    pairs = _unique(w.pairs)
    got, _ = _run(1, w, pairs, "as_called", cov_thr=0.3, ani_thr=20)
    want, _ = _expected(1, w, pairs, "as_called", 0, -1.0, 0.3, 20)
    np.testing.assert_array_equal(got, want)


def test_learned_filter_extremes():
    from pastis_amd.prefilter import init_params
    w = small_config2(n_seq=80, n_pairs=400, seed=45)
    pairs = _unique(w.pairs)
    p = init_params(0)
    mean = np.zeros(16, np.float32)
    istd = np.ones(16, np.float32)
    # tau = +inf drops every pair: nothing is aligned, everything is pruned
    got, st = _run(1, w, pairs, "as_called",
                   filt=(p, mean, istd, np.float32(np.inf)))
    assert not got.any() and st["aligned"] == 0
    assert st["filtered_out"] == len(pairs) and st["nnz_aln"] == 0
    # tau = -inf keeps every pair: identical to the plain path
    got, st = _run(1, w, pairs, "as_called",
                   filt=(p, mean, istd, np.float32(-np.inf)))
    want, _ = _expected(1, w, pairs, "as_called", 0, -1.0)
    np.testing.assert_array_equal(got, want)
    assert st["filtered_out"] == 0


def test_duplicate_coordinates_and_bad_index_are_errors():
    w = small_config2(n_seq=40, n_pairs=50, seed=46)
    pairs = _unique(w.pairs)
    dup = np.concatenate([pairs[:3], pairs[:1]])
    with pytest.raises(_lib.PaError):
        _run(1, w, dup, "as_called")
    bad = pairs[:5].copy()
    bad["r"][2] = 10 ** 5
    with pytest.raises(_lib.PaError):
        _run(1, w, bad, "as_called")
