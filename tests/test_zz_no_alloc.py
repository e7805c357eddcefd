"""pa_sw_scores_noalloc (the score-only local alignment with its workspace
in the library's fixed device array) against pa_sw_scores_dev, bit-exact,
on two calls in a row -- the second with a smaller batch, so stale contents
of the shared array from the first call must not leak into it -- and its
size limits."""
import numpy as np
import pytest

from helpers import pairs_of, rand_seq, seqset, small_config1
from pastis_amd import _lib
from pastis_amd._lib import SCORE_DTYPE
from pastis_amd.aligner import Aligner, config_for_cli

pytestmark = pytest.mark.gpu


def _run(al, fn, d_pairs, n, d_out):
    fn(d_pairs, n, d_out)
    out = np.zeros(n, SCORE_DTYPE)
    al.synchronize()
    return al.d2h(out, d_out)


def test_noalloc_matches_scores_dev_twice():
    # lengths up to 2,600: multi-pass tasks use the boundary rows
    rng = np.random.default_rng(17)
    w = small_config1(n_seq=150, n_pairs=6000, seed=81)
    L = w.lengths
    seqs = [w.codes[int(w.offsets[i]):int(w.offsets[i + 1])]
            for i in range(len(L))]
    seqs += [rand_seq(rng, 2600), rand_seq(rng, 1500)]
    codes, offs = seqset(seqs)
    n0 = len(seqs) - 2
    extra = pairs_of([(n0, n0 + 1), (n0 + 1, n0), (n0, 3), (7, n0)])
    pairs = np.concatenate([w.pairs, extra])
    al = Aligner(config_for_cli(-11, -2, 6, 0, 2, "as_called"))
    al.load_seqs(codes, offs)
    n = len(pairs)
    d_pairs = al.device_alloc(pairs.nbytes)
    d_a = al.device_alloc(n * SCORE_DTYPE.itemsize)
    d_b = al.device_alloc(n * SCORE_DTYPE.itemsize)
    al.h2d(d_pairs, pairs)
    print("noalloc call 1", n, flush=True)
    want = _run(al, al.sw_scores_dev, d_pairs, n, d_a)
    got = _run(al, al.sw_scores_noalloc, d_pairs, n, d_b)
    assert np.array_equal(got, want)
    assert (want["score"][-4:] > 0).all()
    # second call, smaller, other pairs
    m = 1700
    small = pairs[::-1][:m].copy()
    small["r"], small["c"] = small["c"].copy(), small["r"].copy()
    al.h2d(d_pairs, small)
    print("noalloc call 2", m, flush=True)
    want = _run(al, al.sw_scores_dev, d_pairs, m, d_a)
    got = _run(al, al.sw_scores_noalloc, d_pairs, m, d_b)
    assert np.array_equal(got, want)
    for b in (d_pairs, d_a, d_b):
        al.device_free(b)
    al.close()


def test_noalloc_limits():
    rng = np.random.default_rng(3)
    al = Aligner(config_for_cli(-11, -2, 6, 0, 2, "as_called"))
    codes, offs = seqset([rand_seq(rng, 100), rand_seq(rng, 120)])
    al.load_seqs(codes, offs)
    p = pairs_of([(0, 1)])
    d_p = al.device_alloc(p.nbytes)
    d_o = al.device_alloc(SCORE_DTYPE.itemsize)
    al.h2d(d_p, p)
    # the size check comes before any device work
    with pytest.raises(_lib.PaError, match="262144"):
        al.sw_scores_noalloc(d_p, 262145, d_o)
    codes, offs = seqset([rand_seq(rng, 5000), rand_seq(rng, 120)])
    al.load_seqs(codes, offs)
    with pytest.raises(_lib.PaError, match="4096"):
        al.sw_scores_noalloc(d_p, 1, d_o)
    al.device_free(d_p)
    al.device_free(d_o)
    al.close()
