"""Replays of a captured score-only SW batch (pa_sw_scores_graph) and of a
captured traced batch (pa_align_graph, both modes) give the eager results
every time, also after the pairs' contents change between replays and with
eager work and synchronizations between them; a different batch size
recaptures; a traced replay whose trace outgrows the captured store is
flagged, not wrong."""
import numpy as np
import pytest

from helpers import small_config1, small_config2
from pastis_amd._lib import (PA_FLAG_NOTRACE, PAIR_DTYPE, RESULT_DTYPE,
                             SCORE_DTYPE)
from pastis_amd.aligner import Aligner, config_for_cli

pytestmark = pytest.mark.gpu


def _eager(al, d_pairs, n, d_out):
    al.sw_scores_dev(d_pairs, n, d_out)
    out = np.zeros(n, SCORE_DTYPE)
    al.synchronize()
    return al.d2h(out, d_out)


def test_graph_replays_match_eager():
    w = small_config1(n_seq=150, n_pairs=3000, seed=51)
    al = Aligner(config_for_cli(-11, -2, 6, 0, 2, "as_called"))
    al.load_seqs(w.codes, w.offsets)
    n = len(w.pairs)
    d_pairs = al.device_alloc(w.pairs.nbytes)
    d_g = al.device_alloc(n * SCORE_DTYPE.itemsize)
    d_e = al.device_alloc(n * SCORE_DTYPE.itemsize)
    rng = np.random.default_rng(0)
    batches = [w.pairs]
    for _ in range(3):
        p = w.pairs.copy()
        perm = rng.permutation(n)
        p["r"], p["c"] = w.pairs["c"][perm], w.pairs["r"][perm]
        batches.append(p)
    for i, p in enumerate(batches):
        print("graph replay", i, flush=True)
        al.h2d(d_pairs, np.ascontiguousarray(p, PAIR_DTYPE))
        al.sw_scores_graph(d_pairs, n, d_g)          # capture once, replay
        got = np.zeros(n, SCORE_DTYPE)
        al.synchronize()
        al.d2h(got, d_g)
        want = _eager(al, d_pairs, n, d_e)           # eager work in between
        assert np.array_equal(got, want), i
    # a smaller batch on the same buffers recaptures
    m = n // 3
    al.sw_scores_graph(d_pairs, m, d_g)
    got = np.zeros(m, SCORE_DTYPE)
    al.synchronize()
    al.d2h(got, d_g)
    assert np.array_equal(got, _eager(al, d_pairs, m, d_e))
    for b in (d_pairs, d_g, d_e):
        al.device_free(b)


def _align_eager(al, d_pairs, n, d_out):
    al.align_dev(d_pairs, n, d_out)
    out = np.zeros(n, RESULT_DTYPE)
    al.synchronize()
    return al.d2h(out, d_out)


def _align_replays(model, xdrop, w, trace_bytes=None, tuning=None):
    """pa_align_graph over a batch whose contents change between calls,
    with eager pa_align_dev work in between; every call against eager."""
    import os
    old = os.environ.get("PA_TRACE_BYTES")
    if trace_bytes:
        os.environ["PA_TRACE_BYTES"] = str(trace_bytes)
    try:
        al = Aligner(config_for_cli(-11, -2, 6, xdrop, 2, model))
    finally:
        if trace_bytes:
            if old is None:
                del os.environ["PA_TRACE_BYTES"]
            else:
                os.environ["PA_TRACE_BYTES"] = old
    if tuning:
        al.set_tuning(**tuning)
    al.load_seqs(w.codes, w.offsets)
    n = len(w.pairs)
    d_pairs = al.device_alloc(w.pairs.nbytes)
    d_g = al.device_alloc(n * RESULT_DTYPE.itemsize)
    d_e = al.device_alloc(n * RESULT_DTYPE.itemsize)
    rng = np.random.default_rng(1)
    batches = [w.pairs]
    for _ in range(3):
        batches.append(w.pairs[rng.permutation(n)])
    al.reset_stats()
    for i, p in enumerate(batches):
        print("align graph", model, xdrop, trace_bytes, i, flush=True)
        al.h2d(d_pairs, np.ascontiguousarray(p, PAIR_DTYPE))
        al.align_graph(d_pairs, n, d_g)
        got = np.zeros(n, RESULT_DTYPE)
        al.synchronize()
        al.d2h(got, d_g)
        want = _align_eager(al, d_pairs, n, d_e)
        assert np.array_equal(got, want), i
        assert not (got["flags"] & PA_FLAG_NOTRACE).any()
    assert al.stats()["graph_captures"] == 1
    for b in (d_pairs, d_g, d_e):
        al.device_free(b)
    al.close()


@pytest.mark.parametrize("model", ["as_called", "as_documented"])
def test_traced_xdrop_graph_replays_match_eager(model):
    _align_replays(model, 49, small_config2(n_seq=150, n_pairs=2000,
                                            seed=52))


def test_traced_sw_graph_replays_match_eager():
    _align_replays("as_called", 0, small_config1(n_seq=120, n_pairs=1500,
                                                 seed=53))


def test_traced_graph_with_device_sized_chunks():
    """A 16 MiB trace store: the captured batch runs in several chunks whose
    task ranges and trace bases come from the device."""
    _align_replays("as_called", 49, small_config2(n_seq=150, n_pairs=2000,
                                                  seed=54),
                   trace_bytes=16 << 20)


@pytest.mark.parametrize("overlap", [1, 2])
@pytest.mark.parametrize("trace_bytes", [None, 16 << 20])
def test_traced_graph_with_walks_beside_the_dp(overlap, trace_bytes):
    """Traceback walks on the second stream beside the DP (each class's
    beside the next classes', or the shorter classes' beside the last
    class's): the capture forks and joins that stream through events, in
    one chunk and in many."""
    _align_replays("as_called", 49, small_config2(n_seq=150, n_pairs=2000,
                                                  seed=56),
                   trace_bytes=trace_bytes, tuning=dict(tb_overlap=overlap))


def test_traced_graph_overflow_is_flagged():
    """A replay whose trace outgrows the captured store returns
    PA_FLAG_NOTRACE results instead of wrong ones."""
    w = small_config2(n_seq=150, n_pairs=2000, seed=55)
    al = Aligner(config_for_cli(-11, -2, 6, 49, 2, "as_called"))
    al.load_seqs(w.codes, w.offsets)
    n = len(w.pairs)
    short = w.pairs.copy()
    # capture on a batch of short alignments: a pair of a sequence with
    # itself, seeded at the start, traces its whole length, so replace
    # every pair by the one with the shortest sequences
    lens = np.diff(w.offsets.astype(np.int64))
    i0 = int(np.argmin(lens[w.pairs["r"]] + lens[w.pairs["c"]]))
    short[:] = w.pairs[i0]
    d_pairs = al.device_alloc(w.pairs.nbytes)
    d_g = al.device_alloc(n * RESULT_DTYPE.itemsize)
    al.h2d(d_pairs, np.ascontiguousarray(short, PAIR_DTYPE))
    al.align_graph(d_pairs, n, d_g)          # captures on the short batch
    al.synchronize()
    # long identical pairs: far more trace than the captured store holds
    j0 = int(np.argmax(lens))
    long_ = w.pairs.copy()
    long_["r"] = long_["c"] = j0
    long_["seed"] = 0
    al.h2d(d_pairs, np.ascontiguousarray(long_, PAIR_DTYPE))
    al.align_graph(d_pairs, n, d_g)
    got = np.zeros(n, RESULT_DTYPE)
    al.synchronize()
    al.d2h(got, d_g)
    assert (got["flags"] == PA_FLAG_NOTRACE).all()
    for b in (d_pairs, d_g):
        al.device_free(b)
    al.close()
