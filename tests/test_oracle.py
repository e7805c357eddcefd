"""CPU oracle pinned against fixtures, brute force and closed forms.

The reference has no tests or golden vectors for this path (its numeric
core lives in the absent SeqAn library), so the oracle is pinned here by:
  * the BLOSUM62 fixture (NCBI layout) against the oracle's table, which was
    transcribed from inc/ScoreMat.hpp:104-131;
  * exhaustive path enumeration for tiny inputs (scores, local end cells,
    and that the reported statistics describe an optimal path);
  * closed forms: self-alignment = sum of the diagonal, symmetry, affine
    open > extend collapsing to linear, X-drop with X -> infinity equal to
    the linear global score of the suffixes, the reversed-direction
    identity of the left extension, and the X-drop termination bound.
"""
import os
import random

import numpy as np
import pytest

import bruteforce as bf
from oracle import oracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden")
MAT = orc.blosum62()
STD20 = "ARNDCQEGHILKMFPSTWYV"
GAPS = [(-11, -2), (-2, -11), (-3, -1), (-5, -5)]


def rnd_seq(rng, n, alphabet=STD20):
    return orc.encode("".join(rng.choice(alphabet) for _ in range(n)))


def test_blosum62_matches_ncbi_fixture():
    lines = [l for l in open(os.path.join(GOLD, "BLOSUM62.ncbi"))
             if not l.startswith("#")]
    cols = lines[0].split()
    assert len(cols) == 24
    for line in lines[1:]:
        parts = line.split()
        a = parts[0]
        for b, val in zip(cols, parts[1:]):
            ia, ib = orc.ALPHABET.index(a), orc.ALPHABET.index(b)
            assert MAT[ia * 32 + ib] == int(val), (a, b)
    m = MAT.reshape(32, 32)[:25, :25]
    assert (m == m.T).all()
    # J (last code) scores like L against the standard residues
    j, l = orc.ALPHABET.index("J"), orc.ALPHABET.index("L")
    assert (m[j, :20] == m[l, :20]).all()


@pytest.mark.parametrize("go,ge", GAPS)
def test_local_brute_force(go, ge):
    rng = random.Random(1000 + go * 7 + ge)
    for trial in range(40):
        h = rnd_seq(rng, rng.randint(1, 5), "ARNDW" if trial % 2 else STD20)
        v = rnd_seq(rng, rng.randint(1, 5), "ARNDW" if trial % 2 else STD20)
        top, i1, j1 = bf.local_first_max(h, v, MAT, go, ge)
        o = orc.sw_local(h, v, go, ge, stats=True)
        assert o["score"] == top
        assert (o["v_end"], o["h_end"]) == (i1, j1)
        if top == 0:
            assert o["cols"] == 0
            continue
        _, paths = bf.local_optimal_paths(h, v, MAT, go, ge)
        cands = [p for p in paths if (p[2], p[3]) == (i1, j1)
                 and (p[0], p[1]) == (o["v_beg"], o["h_beg"])]
        stats = [bf.path_stats(h, v, p[0], p[1], p[4]) for p in cands]
        assert {"matches": o["matches"], "cols": o["cols"],
                "gap_opens": o["gap_opens"]} in stats


@pytest.mark.parametrize("go,ge", GAPS)
def test_global_brute_force(go, ge):
    rng = random.Random(2000 - go * 5 + ge)
    for trial in range(40):
        h = rnd_seq(rng, rng.randint(1, 5), "ARNDW" if trial % 2 else STD20)
        v = rnd_seq(rng, rng.randint(1, 5), "ARNDW" if trial % 2 else STD20)
        top, paths = bf.global_optimal_paths(h, v, MAT, go, ge)
        o = orc.nw_global(h, v, go, ge, stats=True)
        assert o["score"] == top
        stats = [bf.path_stats(h, v, 0, 0, p) for p in paths]
        assert {"matches": o["matches"], "cols": o["cols"],
                "gap_opens": o["gap_opens"]} in stats


@pytest.mark.parametrize("go,ge", GAPS)
def test_local_traceback_policy(go, ge):
    """The oracle's statistics and begin cell are those of the path the
    documented tie policy selects among all optimal paths (bruteforce's
    whole-path rule), not merely of some optimal path."""
    rng = random.Random(3000 + go * 11 + ge)
    done = 0
    for trial in range(60):
        alpha = ["AW", "ARNDW", STD20][trial % 3]
        h = rnd_seq(rng, rng.randint(1, 5), alpha)
        v = rnd_seq(rng, rng.randint(1, 5), alpha)
        o = orc.sw_local(h, v, go, ge, stats=True)
        if o["score"] == 0:
            continue
        i0, j0, mv, top = bf.policy_path_local(
            h, v, MAT, go, ge, (o["v_end"], o["h_end"]))
        assert top == o["score"]
        st = bf.path_stats(h, v, i0, j0, mv)
        assert (o["v_beg"], o["h_beg"]) == (i0, j0), (h, v, mv)
        assert (o["matches"], o["cols"], o["gap_opens"]) == \
            (st["matches"], st["cols"], st["gap_opens"]), (h, v, mv)
        done += 1
    assert done >= 30


@pytest.mark.parametrize("go,ge", GAPS)
def test_global_traceback_policy(go, ge):
    rng = random.Random(4000 - go * 13 + ge)
    for trial in range(50):
        alpha = ["AW", "ARNDW", STD20][trial % 3]
        h = rnd_seq(rng, rng.randint(1, 5), alpha)
        v = rnd_seq(rng, rng.randint(1, 5), alpha)
        o = orc.nw_global(h, v, go, ge, stats=True)
        mv, top = bf.policy_path_global(h, v, MAT, go, ge)
        assert top == o["score"]
        st = bf.path_stats(h, v, 0, 0, mv)
        assert (o["matches"], o["cols"], o["gap_opens"]) == \
            (st["matches"], st["cols"], st["gap_opens"]), (h, v, mv)


# Global inputs on which the policy's statistics differ from those of the
# same rule with vertical and horizontal swapped ("hv") or with opening
# preferred to extension ("oe"); found by searching small random inputs.
DISCRIMINATING = [
    ("hv", "ASA", "SAWS", -2, -11), ("hv", "SSASW", "WWWSA", -2, -11),
    ("hv", "WAAA", "AAAWA", -3, -1), ("hv", "AAAW", "WAA", -3, -1),
    ("hv", "WAA", "AAAW", -3, -1),
    ("oe", "RA", "RRAAN", -5, -5), ("oe", "NRR", "RRRNR", -5, -5),
    ("oe", "NR", "NNNA", -5, -5), ("oe", "AARRA", "NAN", -5, -5),
    ("oe", "RNNA", "RN", -5, -5), ("oe", "NANAA", "NA", -5, -5),
]
SWAPPED = {"hv": {"stop": 6, "D": 5, "He": 4, "Ho": 3, "Ve": 2, "Vo": 1},
           "oe": {"stop": 6, "D": 5, "Vo": 4, "Ve": 3, "Ho": 2, "He": 1}}


@pytest.mark.parametrize("kind,hs,vs,go,ge", DISCRIMINATING)
def test_traceback_policy_on_discriminating_inputs(kind, hs, vs, go, ge,
                                                   monkeypatch):
    """On these inputs several optimal paths have different statistics; the
    oracle's are the documented policy's, and a swapped priority would give
    different ones (so the policy checks above can fail)."""
    h, v = orc.encode(hs), orc.encode(vs)
    o = orc.nw_global(h, v, go, ge, stats=True)
    mv, _ = bf.policy_path_global(h, v, MAT, go, ge)
    want = tuple(bf.path_stats(h, v, 0, 0, mv).values())
    assert (o["matches"], o["cols"], o["gap_opens"]) == want
    monkeypatch.setattr(bf, "RANK", SWAPPED[kind])
    mv2, _ = bf.policy_path_global(h, v, MAT, go, ge)
    assert tuple(bf.path_stats(h, v, 0, 0, mv2).values()) != want


def test_local_closed_forms():
    rng = random.Random(7)
    for _ in range(20):
        s = rnd_seq(rng, rng.randint(20, 200))
        diag = sum(int(MAT[int(c) * 33]) for c in s)
        for go, ge in GAPS:
            o = orc.sw_local(s, s, go, ge)
            assert o["score"] == diag
            assert o["matches"] == len(s) and o["cols"] == len(s)
            assert o["gap_opens"] == 0
        a, b = rnd_seq(rng, 60), rnd_seq(rng, 80)
        for go, ge in GAPS:
            assert (orc.sw_local(a, b, go, ge)["score"]
                    == orc.sw_local(b, a, go, ge)["score"])
            best_pair = max(int(MAT[int(x) * 32 + int(y)]) for x in a for y in b)
            assert orc.sw_local(a, b, go, ge)["score"] >= max(best_pair, 0)


def test_open_above_extend_is_linear():
    """With open > extend the affine DP equals linear gaps of cost `open`."""
    rng = random.Random(11)
    for _ in range(30):
        a, b = rnd_seq(rng, rng.randint(5, 60)), rnd_seq(rng, rng.randint(5, 60))
        assert (orc.sw_local(a, b, -2, -11)["score"]
                == orc.sw_local(a, b, -2, -2)["score"])
        assert (orc.nw_global(a, b, -2, -11)["score"]
                == orc.nw_global(a, b, -2, -2)["score"])


def test_single_mismatch_in_identical_flanks():
    rng = random.Random(5)
    s = rnd_seq(rng, 41)
    t = s.copy()
    t[20] = (t[20] + 1) % 20
    diag = sum(int(MAT[int(c) * 33]) for c in s)
    expect = diag - int(MAT[int(s[20]) * 33]) + int(MAT[int(s[20]) * 32 + int(t[20])])
    o = orc.nw_global(s, t, -11, -1)
    assert o["score"] == expect
    assert o["matches"] == 40 and o["cols"] == 41 and o["gap_opens"] == 0


def test_global_pure_gap_runs():
    a = orc.encode("WWWW")
    b = orc.encode("WW")
    o = orc.nw_global(a, b, -11, -1)
    assert o["score"] == 22 - 11 - 1
    assert o["cols"] == 4 and o["matches"] == 2 and o["gap_opens"] == 1


def test_predicate_thresholds():
    a = dict(score=0, v_beg=0, v_end=0, h_beg=0, h_end=0, matches=30,
             cols=100, gap_opens=0)
    aln = orc.OrAln(**a)
    import ctypes as C
    ident, sa = C.c_float(), C.c_float()
    L = orc.lib()
    L.or_predicate.argtypes = [C.POINTER(orc.OrAln), C.c_int, C.c_int,
                               C.c_double, C.c_int, C.POINTER(C.c_float),
                               C.POINTER(C.c_float)]
    assert L.or_predicate(C.byref(aln), 100, 200, 0.7, 30, C.byref(ident),
                          C.byref(sa)) == 1
    assert ident.value == pytest.approx(30.0) and sa.value == np.float32(0.3)
    aln.cols = 101  # identity 29.7 < 30
    assert L.or_predicate(C.byref(aln), 100, 200, 0.7, 30, C.byref(ident),
                          C.byref(sa)) == 0
    assert sa.value == 0.0
    aln.cols, aln.matches, aln.gap_opens = 80, 80, 10  # (80-10)/100 = 0.7
    assert L.or_predicate(C.byref(aln), 100, 300, 0.7, 30, C.byref(ident),
                          C.byref(sa)) == 1
    aln.gap_opens = 11
    assert L.or_predicate(C.byref(aln), 100, 300, 0.7, 30, C.byref(ident),
                          C.byref(sa)) == 0
    aln.cols = 0  # empty alignment: identity NaN, never passes
    aln.matches = aln.gap_opens = 0
    assert L.or_predicate(C.byref(aln), 1, 1, 0.0, 0, C.byref(ident),
                          C.byref(sa)) == 0


def test_xdrop_hand_worked():
    w = orc.encode("W")
    x = orc.xdrop_extend(np.concatenate([w, w]), np.concatenate([w, w]),
                         0, 0, 1, -11, 49)
    # seed W/W at 0; right extension W/W adds one diagonal cell of 11
    assert (x["beg_h"], x["beg_v"], x["end_h"], x["end_v"]) == (0, 0, 2, 2)
    assert x["score_right"] == 11 and x["cells"] == 1


@pytest.mark.parametrize("gap", [-11, -2, -1])
def test_xdrop_infinite_x_is_linear_global(gap):
    rng = random.Random(100 - gap)
    for _ in range(25):
        h = rnd_seq(rng, rng.randint(8, 70))
        v = rnd_seq(rng, rng.randint(8, 70))
        bh, bv = rng.randint(0, len(h) - 3), rng.randint(0, len(v) - 3)
        x = orc.xdrop_extend(h, v, bh, bv, 3, gap, 10**7)
        hs, vs = h[bh + 3:], v[bv + 3:]
        # an empty prefix or suffix on either side leaves that side alone
        exp_beg = (0, 0) if bh and bv else (bh, bv)
        exp_end = ((len(h), len(v)) if len(hs) and len(vs)
                   else (bh + 3, bv + 3))
        assert (x["beg_h"], x["beg_v"]) == exp_beg
        assert (x["end_h"], x["end_v"]) == exp_end
        if len(hs) and len(vs):
            ref = orc.nw_global(hs, vs, gap, gap, stats=False)["score"]
            assert x["score_right"] == ref
        hp, vp = h[:bh][::-1], v[:bv][::-1]
        if len(hp) and len(vp):
            ref = orc.nw_global(hp, vp, gap, gap, stats=False)["score"]
            assert x["score_left"] == ref
        full = bh * bv + (len(hs) * len(vs))
        assert x["cells"] == full


def test_xdrop_left_equals_right_on_reversed():
    rng = random.Random(3)
    for _ in range(200):
        h = rnd_seq(rng, rng.randint(10, 120))
        v = rnd_seq(rng, rng.randint(10, 120))
        k = 4
        bh, bv = rng.randint(0, len(h) - k), rng.randint(0, len(v) - k)
        for gap, X in [(-11, 49), (-2, 20), (-4, 8)]:
            x = orc.xdrop_extend(h, v, bh, bv, k, gap, X)
            hr = np.concatenate([h[:bh][::-1], h[:bh][:0]])
            vr = v[:bv][::-1]
            if len(hr) == 0 or len(vr) == 0:
                assert (x["beg_h"], x["beg_v"]) == (bh, bv)
                continue
            y = orc.xdrop_extend(hr, vr, 0, 0, 0, gap, X)
            assert y["score_right"] == x["score_left"]
            assert (bh - x["beg_h"], bv - x["beg_v"]) == (y["end_h"], y["end_v"])


def test_xdrop_termination_bound():
    """All cross scores negative: every cell is undefined by anti-diagonal
    2X+2, so at most (2X+2)^2 cells are evaluated per direction."""
    a = orc.encode("W" * 300)
    b = orc.encode("P" * 300)  # W/P = -4
    for X, gap in [(10, -11), (49, -11), (20, -2)]:
        h = np.concatenate([a[:150], orc.encode("C" * 4), a[:150]])
        v = np.concatenate([b[:150], orc.encode("C" * 4), b[:150]])
        x = orc.xdrop_extend(h, v, 150, 150, 4, gap, X)
        assert x["cells"] <= 2 * (2 * X + 2) ** 2
        assert x["end_h"] - 154 <= 2 * X + 2 and 150 - x["beg_h"] <= 2 * X + 2


def test_xdrop_identical_extends_to_shorter_end():
    rng = random.Random(9)
    s = rnd_seq(rng, 200)
    h, v = s, s[:150]
    # The longest-extension rule keeps the last defined cell in the column
    # of the query end: the identical diagonal plus floor(X/|gap|) trailing
    # gap moves into the longer sequence.
    for gap, X in [(-11, 49), (-2, 49), (-11, 10), (-7, 6)]:
        x = orc.xdrop_extend(h, v, 50, 50, 6, gap, X)
        assert (x["beg_h"], x["beg_v"]) == (0, 0)
        assert x["end_v"] == 150 and x["end_h"] == 150 + min(X // -gap, 50)


def test_sxa_seed_choice_and_lengths():
    rng = random.Random(21)
    p = orc.params()
    for _ in range(30):
        h = rnd_seq(rng, rng.randint(40, 120))
        v = h.copy()
        v[::7] = (v[::7] + 3) % 20
        seeds = [rng.randint(0, len(v) - 6), rng.randint(0, len(h) - 6), 0, 0]
        o = orc.sxa_pair(h, v, seeds, p)
        assert o["seed"] in (0, 1)
        e = o["ext"][o["seed"]]
        assert o["h_beg"] == e["beg_h"] and o["h_end"] == e["end_h"]
        assert o["v_beg"] == e["beg_v"] and o["v_end"] == e["end_v"]
        nw = orc.nw_global(h[e["beg_h"]:e["end_h"]], v[e["beg_v"]:e["end_v"]],
                           p.gap_open, p.gap_ext)
        assert nw["matches"] == o["matches"] and nw["cols"] == o["cols"]
    with pytest.raises(ValueError):
        orc.sxa_pair(orc.encode("ARND"), orc.encode("ARNDAR"), [0, 0, 0, 0], p)


def test_batch_drivers_match_single_calls():
    rng = random.Random(4)
    seqs = [rnd_seq(rng, rng.randint(10, 80)) for _ in range(20)]
    offs = np.zeros(len(seqs) + 1, np.uint64)
    offs[1:] = np.cumsum([len(s) for s in seqs])
    flat = np.concatenate(seqs)
    pr = np.array([rng.randrange(20) for _ in range(50)], np.uint32)
    pc = np.array([rng.randrange(20) for _ in range(50)], np.uint32)
    sc, ve, he = orc.sw_batch(flat, offs, pr, pc, -11, -2, nthreads=2)
    for t in range(50):
        o = orc.sw_local(seqs[pc[t]], seqs[pr[t]], -11, -2, stats=False)
        assert (sc[t], ve[t], he[t]) == (o["score"], o["v_end"], o["h_end"])
    p = orc.params()
    out = orc.sfa_batch(flat, offs, pr, pc, p, nthreads=2)
    for t in range(50):
        o = orc.sfa_pair(seqs[pc[t]], seqs[pr[t]], p)
        assert out["matches"][t] == o["matches"]
        assert out["passed"][t] == o["passed"]
