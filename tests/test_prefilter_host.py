"""Host logic of the learned prefilter: threshold calibration for a recall
target, with and without the margin for the calibration set's sampling
error (no GPU)."""
import numpy as np

from pastis_amd import prefilter


def test_threshold_exact_quantile():
    # 1000 positives with logits 0..999, negatives elsewhere
    pos = np.arange(1000, dtype=np.float32)
    neg = np.full(500, -5.0, np.float32)
    z = np.concatenate([pos, neg])
    y = np.concatenate([np.ones(1000), np.zeros(500)])
    # 0.5 % of 1000 positives may be dropped: the 5 smallest
    assert prefilter.threshold_for_recall(z, y, 0.995) == 5.0
    rep = prefilter.report(z, y, 5.0)
    assert rep["recall"] == 0.995
    assert rep["positives"] == 1000


def test_threshold_margin_closed_form():
    pos = np.arange(1000, dtype=np.float32)
    y = np.ones(1000)
    # m = 5 misses allowed; z = 1.645 leaves floor(5 - 1.645 sqrt(5)) = 1
    assert prefilter.threshold_for_recall(pos, y, 0.995, z=1.645) == 1.0
    # a margin larger than m drops none
    assert prefilter.threshold_for_recall(pos, y, 0.995, z=10.0) == 0.0
    assert prefilter.threshold_for_recall(np.zeros(0), np.zeros(0)) == -np.inf


def test_margin_holds_recall_on_fresh_samples():
    """Calibrate on one sample, measure recall on another from the same
    distribution: with z = 1.645 the target holds in about 95 % of trials,
    without the margin in about half."""
    rng = np.random.default_rng(0)
    ok_margin = ok_plain = 0
    trials = 400
    for _ in range(trials):
        cal = rng.normal(size=4000).astype(np.float32)
        new = rng.normal(size=4000).astype(np.float32)
        y = np.ones(4000)
        t1 = prefilter.threshold_for_recall(cal, y, 0.995, z=1.645)
        t0 = prefilter.threshold_for_recall(cal, y, 0.995)
        ok_margin += np.mean(new >= t1) >= 0.995
        ok_plain += np.mean(new >= t0) >= 0.995
    assert ok_margin / trials > 0.9
    assert ok_plain / trials < 0.75
