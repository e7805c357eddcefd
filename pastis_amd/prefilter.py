"""Learned candidate prefilter: labels, training, calibration, filtered runs.

The filter stands beside the reference's k-mer prefilter
(src/DistPWRunner.cpp:48-54, prune when count <= ckthr || score <= mosthr):
an MLP 16 -> 64 -> 64 -> 1 on per-pair features predicts whether the exact
alignment will pass the coverage/identity predicate
(src/align/SeqanXdropAligner.cpp:344-351).  Pairs below the threshold are
dropped; every other pair (the fallback band) is aligned exactly, so the
filter can lose recall but never changes a reported score.  A random audit
sample of dropped pairs is aligned as well and estimates the recall online.

Everything runs through the C ABI (pa_filter_* in include/pastis_hip.h):
features, training (forward, backward and Adam kernels) and inference are
HIP kernels; this module only orchestrates and computes summary numbers.
"""
import numpy as np

from . import _lib
from ._lib import PAIR_DTYPE, RESULT_DTYPE

NF = _lib.PA_FILTER_NF
NPARAM = _lib.PA_FILTER_NPARAM
H = 64


def init_params(seed=0):
    """He-uniform weights, zero biases; flat layout W1 b1 W2 b2 w3 b3."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    o = 0
    p[o:o + NF * H] = rng.uniform(-1, 1, NF * H) * np.sqrt(6.0 / NF)
    o += NF * H + H
    p[o:o + H * H] = rng.uniform(-1, 1, H * H) * np.sqrt(6.0 / H)
    o += H * H + H
    p[o:o + H] = rng.uniform(-1, 1, H) * np.sqrt(6.0 / H)
    return p


def labels_of(results):
    """Pass labels from pa_result flags (the predicate of the exact path)."""
    return ((results["flags"] & _lib.PA_FLAG_PASS) != 0).astype(np.uint8)


def threshold_for_recall(logits, labels, recall=0.995, z=0.0):
    """Largest tau whose keep rule (logit >= tau) keeps >= recall of the
    positives of this calibration set.

    z > 0 leaves a one-sided margin for the sampling error of the
    calibration set: of npos positives it allows k = m - z * sqrt(m) misses
    (m = (1 - recall) * npos, the normal approximation of the binomial
    count), so that recall on new pairs from the same distribution stays
    at or above the target with probability about Phi(z)."""
    pos = np.sort(np.asarray(logits, np.float32)[np.asarray(labels) != 0])
    if len(pos) == 0:
        return np.float32(-np.inf)
    m = (1.0 - recall) * len(pos)
    k = int(np.floor(max(m - z * np.sqrt(m), 0.0)))
    return np.float32(pos[min(k, len(pos) - 1)])


def report(logits, labels, tau, band_hi=0.0):
    """Recall / precision of keep = logit >= tau, the kept fraction and the
    uncertain band [tau, band_hi) (band_hi is a logit; 0 = probability 0.5)."""
    logits = np.asarray(logits, np.float32)
    y = np.asarray(labels) != 0
    keep = logits >= tau
    tp = int(np.sum(keep & y))
    npos = int(np.sum(y))
    nkeep = int(np.sum(keep))
    band = int(np.sum(keep & (logits < band_hi)))
    return dict(recall=tp / npos if npos else 1.0,
                precision=tp / nkeep if nkeep else 1.0,
                kept_frac=nkeep / max(len(y), 1),
                band_frac=band / max(len(y), 1),
                positives=npos, kept=nkeep, n=len(y), tau=float(tau))


class LearnedFilter:
    """A trained filter bound to an Aligner's context (and its sequences)."""

    def __init__(self, aligner, seed=0):
        self.al = aligner
        self.tau = np.float32(-np.inf)
        aligner.filter_set_model(init_params(seed))

    def features(self, pairs):
        return self.al.filter_features(pairs)

    def train(self, feats, labels, epochs=4, batch=4096, lr=2e-3,
              pos_weight=None, seed=1):
        labels = np.asarray(labels, np.uint8)
        if pos_weight is None:   # balance the classes
            npos = max(int(labels.sum()), 1)
            pos_weight = max(1.0, (len(labels) - npos) / npos)
        return self.al.filter_train(feats, labels, epochs=epochs, batch=batch,
                                    seed=seed, lr=lr, pos_weight=pos_weight,
                                    standardize=True)

    def logits(self, feats):
        return self.al.filter_logits(feats)

    def calibrate(self, feats, labels, recall=0.995, z=0.0):
        self.tau = threshold_for_recall(self.logits(feats), labels, recall, z)
        return self.tau

    def score_pairs(self, pairs):
        """Fused device features + MLP for host pairs (logits)."""
        pairs = np.ascontiguousarray(pairs, PAIR_DTYPE)
        n = len(pairs)
        if n == 0:
            return np.zeros(0, np.float32)
        dp = self.al.device_alloc(pairs.nbytes)
        dl = self.al.device_alloc(4 * n)
        try:
            self.al.h2d(dp, pairs)
            self.al.filter_score_dev(dp, n, dl)
            out = np.zeros(n, np.float32)
            self.al.synchronize()
            return self.al.d2h(out, dl)
        finally:
            self.al.device_free(dp)
            self.al.device_free(dl)


def audit_mask(n, rate, seed=0):
    """Deterministic audit sample: pair i is audited when a hash of (seed, i)
    falls below rate."""
    i = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = i * np.uint64(0x9E3779B97F4A7C15) + np.uint64(seed)
        z ^= z >> np.uint64(31)
        z *= np.uint64(0xBF58476D1CE4E5B9)
        z ^= z >> np.uint64(29)
    return (z >> np.uint64(11)).astype(np.float64) / float(1 << 53) < rate


def filtered_align(flt, pairs, audit_rate=0.0, audit_seed=0):
    """Learned filter + exact aligner on host pairs.

    Returns (results, info): results has one pa_result per input pair; pairs
    the filter dropped (and did not audit) keep zeros, like pairs the
    reference never writes.  info holds the kept count, the audit counts and
    the recall estimated from the audit sample.
    """
    al = flt.al
    pairs = np.ascontiguousarray(pairs, PAIR_DTYPE)
    n = len(pairs)
    out = np.zeros(n, RESULT_DTYPE)
    if n == 0:
        return out, dict(kept=0, audited=0, audit_pass=0, recall_est=1.0)
    dp = al.device_alloc(pairs.nbytes)
    dk = al.device_alloc(pairs.nbytes)
    di = al.device_alloc(4 * n)
    dr = al.device_alloc(RESULT_DTYPE.itemsize * n)
    try:
        al.h2d(dp, pairs)
        k = al.filter_select_dev(dp, n, float(flt.tau), dk, di)
        if k:
            al.align_dev(dk, k, dr)
            res = np.zeros(k, RESULT_DTYPE)
            idx = np.zeros(k, np.uint32)
            al.synchronize()
            al.d2h(res, dr)
            al.d2h(idx, di)
            out[idx] = res
        else:
            idx = np.zeros(0, np.uint32)
    finally:
        for p in (dp, dk, di, dr):
            al.device_free(p)
    kept = np.zeros(n, bool)
    kept[idx] = True
    info = dict(kept=int(k), audited=0, audit_pass=0, kept_mask=kept)
    if audit_rate > 0:
        aud = np.nonzero(audit_mask(n, audit_rate, audit_seed) & ~kept)[0]
        info["audited"] = len(aud)
        if len(aud):
            ares = al.align(pairs[aud])
            out[aud] = ares
            info["audit_pass"] = int(np.sum(labels_of(ares)))
    npass_kept = int(np.sum(labels_of(out[kept])))
    missed = info["audit_pass"] / audit_rate if audit_rate > 0 else 0.0
    info["recall_est"] = (npass_kept / (npass_kept + missed)
                          if npass_kept + missed > 0 else 1.0)
    return out, info
