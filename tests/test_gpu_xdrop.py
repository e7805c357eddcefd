"""HIP seed-and-extend X-drop path against the CPU oracle, bit-exact:
extension extents per seed (pa_xdrop_extents) and the full --sxa operator
(global alignment of the extended infixes, best seed, predicate)."""
import numpy as np
import pytest

from helpers import (GAP_MODELS, oracle_params, pairs_of, rand_seq, seqset,
                     small_config2)
from oracle import oracle as orc
from pastis_amd import _lib
from pastis_amd.aligner import Aligner, config_for_cli

pytestmark = pytest.mark.gpu


def make_aligner(model, seed_cnt=2):
    return Aligner(config_for_cli(-11, -2, 6, 49, seed_cnt, model))


def check_extents(al, codes, offs, pairs, model, seed_cnt=2):
    g = GAP_MODELS[model]
    for s in range(seed_cnt):
        got = al.xdrop_extents(pairs, s)
        for k in range(len(pairs)):
            r, c = int(pairs["r"][k]), int(pairs["c"][k])
            v = codes[offs[r]:offs[r + 1]]
            h = codes[offs[c]:offs[c + 1]]
            ro, co = int(pairs["seed"][k][2 * s]), int(pairs["seed"][k][2 * s + 1])
            x = orc.xdrop_extend(h, v, co, ro, 6, g["xdrop_gap"], 49)
            assert (got["r_beg"][k], got["r_end"][k], got["c_beg"][k],
                    got["c_end"][k]) == (x["beg_v"], x["end_v"], x["beg_h"],
                                         x["end_h"]), (k, s)


def check_align(al, codes, offs, pairs, model, seed_cnt=2):
    p = oracle_params(model, seed_cnt=seed_cnt)
    got = al.align(pairs)
    ref = orc.sxa_batch(codes, offs, pairs["r"], pairs["c"], pairs["seed"], p)
    for gf, rf in [("score", "score"), ("matches", "matches"),
                   ("cols", "cols"), ("gap_opens", "gap_opens"),
                   ("seed", "seed")]:
        bad = np.nonzero(got[gf].astype(np.int64) != ref[rf].astype(np.int64))[0]
        assert bad.size == 0, (gf, int(bad[0]), got[bad[0]], ref[bad[0]])
    for gf, rf in [("r_beg", "v_beg"), ("r_end", "v_end"), ("c_beg", "h_beg"),
                   ("c_end", "h_end")]:
        assert (got[gf].astype(np.int64) == ref[rf].astype(np.int64)).all(), gf
    assert (got["identity"] == ref["identity"]).all()
    assert (got["score_aln"] == ref["score_aln"]).all()
    assert ((got["flags"] & _lib.PA_FLAG_PASS).astype(bool)
            == ref["passed"].astype(bool)).all()
    return got


@pytest.mark.parametrize("model", ["as_called", "as_documented"])
def test_xdrop_extents_synthetic(model):
    w = small_config2(n_pairs=600)
    al = make_aligner(model)
    al.load_seqs(w.codes, w.offsets)
    check_extents(al, w.codes, w.offsets, w.pairs, model)


# The per-lane kernel keeps cells as int8 offsets from each anti-diagonal's
# X-drop limit for X <= 160 and gaps in [-32, -1], as int16 offsets for X <=
# 30000 and gaps in [-1000, -1]; beyond either bound every item runs in the
# int32 kernels.  Extents on both sides of each bound.
@pytest.mark.parametrize("xdrop,gap", [(160, -11), (161, -11), (49, -32),
                                       (49, -33), (30000, -11), (30001, -11),
                                       (49, -1000), (49, -1001), (1, -1)])
def test_xdrop_extents_at_the_ring_encoding_limits(xdrop, gap):
    w = small_config2(n_seq=40, n_pairs=80, seed=17)
    cfg = config_for_cli(-11, -2, 6, xdrop, 2, "as_called")
    cfg.xdrop_gap = gap   # the DP gaps stay in their own range
    codes, offs, pairs = w.codes, w.offsets, w.pairs
    want = []
    for s in range(2):
        for k in range(len(pairs)):
            r, c = int(pairs["r"][k]), int(pairs["c"][k])
            v = codes[offs[r]:offs[r + 1]]
            h = codes[offs[c]:offs[c + 1]]
            ro = int(pairs["seed"][k][2 * s])
            co = int(pairs["seed"][k][2 * s + 1])
            x = orc.xdrop_extend(h, v, co, ro, 6, gap, xdrop)
            want.append((x["beg_v"], x["end_v"], x["beg_h"], x["end_h"]))
    for ring in (16, 32):   # both per-lane variants of each cell width
        al = Aligner(cfg)
        al.set_tuning(xd_ring=ring)
        al.load_seqs(codes, offs)
        for s in range(2):
            got = al.xdrop_extents(pairs, s)
            for k in range(len(pairs)):
                assert (got["r_beg"][k], got["r_end"][k], got["c_beg"][k],
                        got["c_end"][k]) == want[s * len(pairs) + k], \
                    (xdrop, gap, ring, k, s)
        al.close()


@pytest.mark.parametrize("model", ["as_called", "as_documented"])
def test_sxa_operator_synthetic(model):
    w = small_config2()
    al = make_aligner(model)
    al.load_seqs(w.codes, w.offsets)
    got = check_align(al, w.codes, w.offsets, w.pairs, model)
    assert (got["flags"] & _lib.PA_FLAG_PASS).any()


def test_sxa_single_seed_and_edge_seeds():
    rng = np.random.default_rng(4)
    a = rand_seq(rng, 400)
    b = a.copy()
    b[::9] = rand_seq(rng, len(b[::9]))
    seqs = [a, b, rand_seq(rng, 6), rand_seq(rng, 1500), a[:200].copy()]
    codes, offs = seqset(seqs)
    # seeds at the very start, the very end and (0,0) second seeds
    rc = [(0, 1), (1, 0), (0, 4), (4, 0), (2, 2), (3, 0), (0, 3)]
    seeds = [[0, 0, 394, 394], [394, 394, 0, 0], [10, 10, 0, 0],
             [194, 194, 0, 0], [0, 0, 0, 0], [700, 100, 0, 0],
             [100, 700, 0, 0]]
    pairs = pairs_of(rc, seeds)
    for model in GAP_MODELS:
        for sc in (1, 2):
            al = make_aligner(model, seed_cnt=sc)
            al.load_seqs(codes, offs)
            check_extents(al, codes, offs, pairs, model, seed_cnt=sc)
            check_align(al, codes, offs, pairs, model, seed_cnt=sc)


def test_seed_out_of_range_is_an_error():
    rng = np.random.default_rng(6)
    codes, offs = seqset([rand_seq(rng, 20), rand_seq(rng, 30)])
    al = make_aligner("as_called")
    al.load_seqs(codes, offs)
    with pytest.raises(_lib.PaError):
        al.align(pairs_of([(0, 1)], [[15, 0, 0, 0]]))


def test_results_do_not_depend_on_tuning():
    """Launch shapes (ring slots, workgroups per CU, the kernel that takes
    the items whose band outgrows the per-lane rings) change speed only."""
    w = small_config2(n_pairs=800, seed=17)
    base = None
    for model in ("as_called", "as_documented"):
        base = None
        for ring, dp, xd, pl, wv in [(16, 0, 0, 0, 1), (32, 0, 0, 1, 1),
                                     (16, 1, 1, 1, 0), (32, 3, 2, 0, 1),
                                     (16, 0, 0, 1, 0), (32, 0, 0, 1, 0),
                                     (16, 0, 0, 1, 2)]:
            al = make_aligner(model)
            al.set_tuning(xd_ring=ring, dp_wg_per_cu=dp, xd_wg_per_cu=xd,
                          xd_coop=1 - pl, xd_wave=wv)
            got_t = al.tuning()
            assert {k: got_t[k] for k in ("dp_wg_per_cu", "xd_wg_per_cu",
                                          "xd_ring", "xd_coop")} == dict(
                dp_wg_per_cu=dp, xd_wg_per_cu=xd, xd_ring=ring,
                xd_coop=1 - pl)
            al.load_seqs(w.codes, w.offsets)
            al.reset_stats()
            got = al.align(w.pairs)
            cells = al.stats()["xdrop_cells"]
            if base is None:
                base = (got, cells)
            assert np.array_equal(got, base[0]), (model, ring, dp, xd, pl, wv)
            assert cells == base[1]
    al = make_aligner("as_called")
    with pytest.raises(_lib.PaError):
        al.set_tuning(xd_ring=24)


def _seed_of(v, h, k=6, band=40):
    """First shared k-mer near the diagonal (row offset, column offset)."""
    for i in range(0, len(v) - k):
        for j in range(max(0, i - band), min(len(h) - k, i + band)):
            if (v[i:i + k] == h[j:j + k]).all():
                return i, j
    return 0, 0


@pytest.mark.parametrize("model", ["as_called", "as_documented"])
def test_sxa_long_homologs(model):
    """Long infixes: two-pass global DP (over 1024 rows), the int16 kernels
    admitted by the residue bound (a ~970-residue infix, above the
    11 * 744 static limit), the int32 kernels (a ~2100-residue infix whose
    residue bound exceeds int16), and a 300-residue single-pass task."""
    rng = np.random.default_rng(5)
    a = rand_seq(rng, 2100)
    b = a.copy()
    b[::7] = rand_seq(rng, len(b[::7]))
    b = np.concatenate([b[:900], rand_seq(rng, 20), b[900:1500], b[1510:]])
    c = rand_seq(rng, 300)
    d = c.copy()
    d[::3] = rand_seq(rng, len(d[::3]))
    seqs = [a, b, a.copy(), c, d]
    codes, offs = seqset(seqs)
    prs = [(0, 1), (0, 2), (3, 4), (1, 0)]
    seeds = []
    for r, cc in prs:
        i, j = _seed_of(seqs[r], seqs[cc])
        seeds.append([i, j, i, j])
    pairs = pairs_of(prs, np.array(seeds, np.uint16))
    print("long homologs", model, flush=True)
    al = make_aligner(model)
    al.load_seqs(codes, offs)
    got = check_align(al, codes, offs, pairs, model)
    assert got["cols"].max() >= 2000
