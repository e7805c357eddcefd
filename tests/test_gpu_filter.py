"""GPU parity of the learned prefilter against the CPU oracle
(oracle/mlp_oracle.py), and what the trained filter gains.

Features, logits (precomputed and fused) and selection are bit-exact; the
gradient matches to 1e-6 relative (the sigmoid goes through a double exp on
both sides, which may differ in the last double bit); Adam training follows
the oracle's trajectory.
"""
import numpy as np
import pytest

from helpers import small_config2
from oracle import mlp_oracle as M
from pastis_amd import _lib, prefilter, synth
from pastis_amd.aligner import Aligner, config_for_cli

pytestmark = pytest.mark.gpu


def _seqs(w):
    o = w.offsets.astype(np.int64)
    return [w.codes[o[i]:o[i + 1]] for i in range(len(o) - 1)]


@pytest.fixture(scope="module")
def small():
    w = small_config2(n_seq=150, n_pairs=1500, seed=21)
    al = Aligner(config_for_cli(xdrop=49))
    al.load_seqs(w.codes, w.offsets)
    print("filter fixture: small_config2 seed 21", flush=True)
    return w, al


def _edge_workload():
    """Sequences of length 0..7 and ordinary ones; pairs with count 0, 1, 2."""
    rng = np.random.default_rng(4)
    seqs = [np.zeros(0, np.uint8)] + [rng.integers(0, 25, L).astype(np.uint8)
                                      for L in (1, 5, 6, 7, 30, 250)]
    offs = np.zeros(len(seqs) + 1, np.uint64)
    offs[1:] = np.cumsum([len(s) for s in seqs])
    codes = np.concatenate(seqs)
    pairs = np.zeros(64, _lib.PAIR_DTYPE)
    for i in range(64):
        r, c = rng.integers(0, len(seqs), 2)
        pairs[i]["r"], pairs[i]["c"] = r, c
        pairs[i]["count"] = i % 3
        pairs[i]["kscore"] = rng.integers(-20, 60)
        lr, lc = len(seqs[r]), len(seqs[c])
        if lr >= 6 and lc >= 6:
            pairs[i]["seed"] = [rng.integers(0, lr - 5), rng.integers(0, lc - 5),
                                rng.integers(0, lr - 5), rng.integers(0, lc - 5)]
    return seqs, codes, offs, pairs


def test_features_bit_exact(small):
    w, al = small
    f = al.filter_features(w.pairs)
    ref = M.features(w.pairs, _seqs(w), _seqs(w))
    np.testing.assert_array_equal(f, ref)


def test_features_edge_cases():
    seqs, codes, offs, pairs = _edge_workload()
    al = Aligner(config_for_cli(xdrop=49))
    al.load_seqs(codes, offs)
    f = al.filter_features(pairs)
    np.testing.assert_array_equal(f, M.features(pairs, seqs, seqs))
    assert np.all(np.isfinite(f))
    # separate row / column sets
    al2 = Aligner(config_for_cli(xdrop=49))
    al2.load_seqs(codes, offs, side=_lib.PA_SIDE_ROW)
    rev = seqs[::-1]
    o2 = np.zeros(len(rev) + 1, np.uint64)
    o2[1:] = np.cumsum([len(s) for s in rev])
    al2.load_seqs(np.concatenate(rev), o2, side=_lib.PA_SIDE_COL)
    p2 = pairs.copy()
    p2["seed"] = 0
    np.testing.assert_array_equal(al2.filter_features(p2),
                                  M.features(p2, seqs, rev))


def _random_model(seed):
    rng = np.random.default_rng(seed)
    p = M.init_params(seed)
    p[M.OFF_B1:M.OFF_W2] = rng.normal(0, 0.2, 64)
    p[M.OFF_B2:M.OFF_W3] = rng.normal(0, 0.2, 64)
    p[M.OFF_B3] = -0.5
    mean = rng.normal(0, 0.5, 16).astype(np.float32)
    istd = rng.uniform(0.5, 3, 16).astype(np.float32)
    return p, mean, istd


def test_logits_bit_exact(small):
    w, al = small
    p, mean, istd = _random_model(1)
    al.filter_set_model(p, mean, istd)
    f = al.filter_features(w.pairs)
    z_ref, _ = M.forward(p, M.standardize(f, mean, istd))
    np.testing.assert_array_equal(al.filter_logits(f), z_ref)
    flt = prefilter.LearnedFilter.__new__(prefilter.LearnedFilter)
    flt.al = al
    np.testing.assert_array_equal(flt.score_pairs(w.pairs), z_ref)
    q, m2, s2 = al.filter_get_model()
    np.testing.assert_array_equal(q, p)
    np.testing.assert_array_equal(m2, mean)
    np.testing.assert_array_equal(s2, istd)


def test_grad_matches_oracle(small):
    w, al = small
    p, mean, istd = _random_model(2)
    al.filter_set_model(p, mean, istd)
    f = al.filter_features(w.pairs)
    y = (np.random.default_rng(0).random(len(f)) < 0.3).astype(np.uint8)
    loss, g = al.filter_grad(f, y, pos_weight=2.0)
    lref, gref = M.loss_and_grad(p, M.standardize(f, mean, istd), y, 2.0)
    assert abs(loss - lref) <= 1e-4 * abs(lref)
    scale = np.abs(gref).max()
    np.testing.assert_allclose(g, gref, rtol=1e-6, atol=1e-7 * scale)
    print("grad entries bit-equal: %.4f" % np.mean(g == gref))


def test_training_follows_oracle(small):
    w, al = small
    p0 = M.init_params(5)
    al.filter_set_model(p0)
    f = al.filter_features(w.pairs)[:600]
    y = (np.random.default_rng(1).random(600) < 0.4).astype(np.uint8)
    losses = al.filter_train(f, y, epochs=1, batch=200, seed=0, lr=1e-3,
                             pos_weight=1.5, standardize=True)
    q, mean, istd = al.filter_get_model()
    mref, sref = M.standardization(f)
    np.testing.assert_allclose(mean, mref, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(istd, sref, rtol=1e-6)
    x = M.standardize(f, mean, istd)
    p = p0.copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    total = 0.0
    for t in range(1, 4):
        sl = slice((t - 1) * 200, t * 200)
        lt, g = M.loss_and_grad(p, x[sl], y[sl], 1.5)
        total += lt * 200
        M.adam_step(p, m, v, g, t, lr=1e-3)
    assert abs(losses[0] - total / 600) < 1e-4
    close = np.abs(q - p) <= 1e-6 + 1e-5 * np.abs(p)
    assert close.mean() > 0.999, close.mean()
    assert np.abs(q - p).max() <= 3 * 2e-3


def test_select_matches_logits(small):
    w, al = small
    p, mean, istd = _random_model(3)
    al.filter_set_model(p, mean, istd)
    f = al.filter_features(w.pairs)
    z, _ = M.forward(p, M.standardize(f, mean, istd))
    tau = float(np.median(z))
    n = len(w.pairs)
    dp = al.device_alloc(w.pairs.nbytes)
    dk = al.device_alloc(w.pairs.nbytes)
    di = al.device_alloc(4 * n)
    al.h2d(dp, w.pairs)
    k = al.filter_select_dev(dp, n, tau, dk, di)
    kept = np.zeros(k, _lib.PAIR_DTYPE)
    idx = np.zeros(k, np.uint32)
    al.synchronize()
    al.d2h(kept, dk)
    al.d2h(idx, di)
    for b in (dp, dk, di):
        al.device_free(b)
    want = np.nonzero(z >= tau)[0]
    assert k == len(want)
    assert np.array_equal(np.sort(idx), want)
    assert np.array_equal(kept, w.pairs[idx])


def test_select_bad_index(small):
    w, al = small
    al.filter_set_model(M.init_params(0))
    bad = w.pairs[:10].copy()
    bad["r"][3] = 10 ** 6
    dp = al.device_alloc(bad.nbytes)
    dk = al.device_alloc(bad.nbytes)
    al.h2d(dp, bad)
    with pytest.raises(_lib.PaError, match="outside"):
        al.filter_select_dev(dp, 10, 0.0, dk)
    al.device_free(dp)
    al.device_free(dk)
    with pytest.raises(_lib.PaError, match="outside"):
        al.filter_features(bad)


def test_trained_filter_recall_and_gain():
    """Train on GPU-aligner labels, calibrate for recall, evaluate held out;
    the filter must keep far fewer pairs than the k-mer count rule at the
    same recall."""
    w = synth.make(1500, 60000, model="lognormal", a=250, b=0.6,
                   clip=(30, 3000), fam_size=15, ident=(0.3, 0.95),
                   frac_hom=0.3, seed=31, name="filter-train")
    al = Aligner(config_for_cli(xdrop=49))
    al.load_seqs(w.codes, w.offsets)
    res = al.align(w.pairs)
    y = prefilter.labels_of(res)
    flt = prefilter.LearnedFilter(al, seed=0)
    f = flt.features(w.pairs)
    tr, ca, te = slice(0, 30000), slice(30000, 45000), slice(45000, 60000)
    losses = flt.train(f[tr], y[tr], epochs=6, batch=2048, lr=3e-3)
    assert losses[-1] < losses[0]
    flt.calibrate(f[ca], y[ca], recall=0.998)
    z = flt.logits(f[te])
    rep = prefilter.report(z, y[te], flt.tau)
    print("held-out:", rep)
    assert rep["positives"] > 500
    assert rep["recall"] >= 0.995
    # baseline without the learned filter: keep pairs by shared k-mer count
    # (the ckthr rule), threshold chosen for the same held-out recall
    cnt = w.pairs["count"][te].astype(np.float32)
    t_cnt = prefilter.threshold_for_recall(cnt, y[te], rep["recall"])
    base = prefilter.report(cnt, y[te], t_cnt)
    print("k-mer count baseline:", base)
    assert rep["kept_frac"] < base["kept_frac"]
    assert rep["kept_frac"] < 0.8
    # the filtered pipeline gives the exact results on every kept pair
    out, info = prefilter.filtered_align(flt, w.pairs[te], audit_rate=0.05)
    kept = info["kept_mask"]
    assert kept.sum() == info["kept"]
    ref = res[te]
    assert np.array_equal(out[kept], ref[kept])
    assert 0.98 <= info["recall_est"] <= 1.0
