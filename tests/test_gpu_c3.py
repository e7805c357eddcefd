"""Heavy-tailed Smith-Waterman batches against the oracle, bit-exact.

The batched uneven-length workload (200,000 proteins, log-normal lengths
plus a 1 % tail of 3,000-35,000 residues, Zipf pairs) puts most of its cells
into pairs that touch a tail sequence: multi-pass tasks of up to 35 passes of
1,024 rows, boundary rows carried between passes, traces of hundreds of
megabytes per task, and near-identical long pairs whose score leaves the int16
range and takes the exact int32 re-run.  These tests run such pairs, taken
from the generator's own tail sequences plus planted near-identical copies,
through the score path (pa_sw_scores_dev), its captured graph
(pa_sw_scores_graph, whose capture then contains the int32 re-run) and the
traced --sfa operator (pa_align), and compare with the oracle."""
import numpy as np
import pytest

from helpers import oracle_params, seqset
from oracle import oracle as orc
from pastis_amd import _lib, synth
from pastis_amd._lib import PAIR_DTYPE, SCORE_DTYPE
from pastis_amd.aligner import Aligner, config_for_cli

pytestmark = pytest.mark.gpu

MODEL = ("as_called", -2, -11)   # DP open -2, extend -11 (Score(-11, -2))


def mutate(rng, s, sub=0.04, indels=4):
    """Copy of s with a fraction `sub` of substitutions and a few short
    insertions and deletions (near-identical: the local score stays far
    above the int16 range for long sequences)."""
    t = s.copy()
    pos = rng.random(len(t)) < sub
    t[pos] = rng.integers(0, 20, int(pos.sum()))
    for _ in range(indels):
        at = int(rng.integers(1, len(t) - 1))
        if rng.random() < 0.5:
            t = np.concatenate([t[:at], rng.integers(0, 20, 3).astype(np.uint8),
                                t[at:]])
        else:
            t = np.concatenate([t[:at], t[at + 3:]])
    return t.astype(np.uint8)


def tail_workload():
    w = synth.config3(scale=0.01, seed=103)
    L = w.lengths
    seqs = [w.codes[int(w.offsets[i]):int(w.offsets[i + 1])]
            for i in range(len(L))]
    tail = [i for i in range(len(L)) if L[i] >= 3000]
    short = [i for i in range(len(L)) if L[i] < 3000]
    assert len(tail) >= 5, len(tail)
    rng = np.random.default_rng(7)
    # both sides >= 30,000 residues: the longest tail sequence against a
    # near-identical copy of itself (int16 saturates in both kernels)
    big = max(tail, key=lambda i: L[i])
    base = seqs[big][:31000]
    if len(base) < 31000:
        base = np.concatenate([base, rng.integers(0, 20, 31000 - len(base))
                               .astype(np.uint8)])
    extra = [base, mutate(rng, base)]
    # planted near-identical pairs of tail sequences long enough for the
    # local score to leave the int16 range
    mid = sorted([i for i in tail if L[i] >= 7000], key=lambda i: L[i])[:4]
    assert len(mid) == 4
    for i in mid:
        extra += [seqs[i], mutate(rng, seqs[i], sub=0.08)]
    n0 = len(seqs)
    seqs = seqs + extra
    codes, offs = seqset(seqs)
    rc = [(n0, n0 + 1)]
    for k in range(4):
        rc.append((n0 + 2 + 2 * k, n0 + 3 + 2 * k))
    # tail x short, short x tail and tail x tail pairs of the generator
    for _ in range(14):
        t = int(rng.choice(tail))
        s = int(rng.choice(short))
        rc.append((t, s) if rng.random() < 0.5 else (s, t))
    for _ in range(3):
        a, b = rng.choice(tail, 2, replace=False)
        rc.append((int(a), int(b)))
    pairs = np.zeros(len(rc), PAIR_DTYPE)
    pairs["r"] = [r for r, _ in rc]
    pairs["c"] = [c for _, c in rc]
    pairs["count"] = 2
    return codes, offs, pairs


@pytest.fixture(scope="module")
def c3():
    codes, offs, pairs = tail_workload()
    L = np.diff(offs).astype(np.int64)
    print("C3 tail batch: %d pairs, lengths up to %d, %.3g cells" %
          (len(pairs), int(L.max()),
           float(np.sum(L[pairs["r"]] * L[pairs["c"]]))), flush=True)
    al = Aligner(config_for_cli(-11, -2, 6, 0, 2, MODEL[0]))
    al.load_seqs(codes, offs)
    # one oracle run gives scores, end cells and the traceback statistics
    full = orc.sfa_batch(codes, offs, pairs["r"], pairs["c"],
                         oracle_params(MODEL[0]))
    ref = (full["score"], full["v_end"], full["h_end"])
    yield al, codes, offs, pairs, ref, full
    al.close()


def _check_scores(got, ref):
    sc, ve, he = ref
    bad = np.nonzero((got["score"] != sc) | (got["r_end"] != ve) |
                     (got["c_end"] != he))[0]
    assert bad.size == 0, ("first mismatch", int(bad[0]), got[bad[0]],
                           sc[bad[0]], ve[bad[0]], he[bad[0]])


def test_c3_tail_scores(c3):
    al, codes, offs, pairs, ref, _ = c3
    got = al.sw_scores(pairs)
    _check_scores(got, ref)
    # the planted pairs leave the int16 range: the int32 re-run ran
    assert (ref[0][:5] > 32767).all(), ref[0][:5]


def test_c3_tail_scores_graph(c3):
    al, codes, offs, pairs, ref, _ = c3
    n = len(pairs)
    d_pairs = al.device_alloc(pairs.nbytes)
    d_out = al.device_alloc(n * SCORE_DTYPE.itemsize)
    al.h2d(d_pairs, pairs)
    for rep in range(3):
        print("C3 graph replay", rep, flush=True)
        al.sw_scores_graph(d_pairs, n, d_out)
        got = np.zeros(n, SCORE_DTYPE)
        al.synchronize()
        al.d2h(got, d_out)
        _check_scores(got, ref)
        al.sw_scores(pairs[:4])   # eager work between replays
    al.device_free(d_pairs)
    al.device_free(d_out)


def test_c3_tail_sfa_operator(c3):
    al, codes, offs, pairs, _, want = c3
    got = al.align(pairs)
    for gf, rf in [("score", "score"), ("r_end", "v_end"), ("c_end", "h_end"),
                   ("r_beg", "v_beg"), ("c_beg", "h_beg"),
                   ("matches", "matches"), ("cols", "cols"),
                   ("gap_opens", "gap_opens")]:
        bad = np.nonzero(got[gf].astype(np.int64) !=
                         want[rf].astype(np.int64))[0]
        assert bad.size == 0, (gf, int(bad[0]), got[bad[0]], want[bad[0]])
    assert (got["score_aln"] == want["score_aln"]).all()
    # the traced int16 kernel holds 4*score + tag: the planted pairs re-ran
    assert (got["flags"][:5] & _lib.PA_FLAG_RERUN32).all()
    assert (got["flags"][:5] & _lib.PA_FLAG_PASS).all()
