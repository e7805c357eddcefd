"""Replays of a captured score-only SW batch (pa_sw_scores_graph) give the
eager results (pa_sw_scores_dev) every time, also after the pairs' contents
change between replays and with eager work and synchronizations between
them; a different batch size recaptures."""
import numpy as np
import pytest

from helpers import small_config1
from pastis_amd._lib import PAIR_DTYPE, SCORE_DTYPE
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
