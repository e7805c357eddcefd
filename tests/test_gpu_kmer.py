"""GPU k-mer candidate generation (pa_kmer_candidates) against the oracle.

The oracle (oracle/kmer_oracle.c) restates generate_A and the
KmerIntersectLoc SpGEMM for one process and is itself pinned by brute force
(tests/test_kmer_oracle.py).  Here the GPU's candidates must equal it
exactly -- (r, c) in column-major order, count, kscore and both seeds --
on synthetic protein sets with planted families, for every alphabet, the
three symmetric rules, strides, column blocks and separate row / column
sets, with the k-mer threshold prune applied to both."""
import numpy as np
import pytest

from oracle import oracle as orc
from pastis_amd import _lib, synth
from pastis_amd.aligner import ALPHABET, Aligner, config_for_cli

pytestmark = pytest.mark.gpu


def ascii_of(codes):
    return np.frombuffer(ALPHABET.encode(), np.uint8)[codes]


def prune(ref, ckthr, mosthr):
    keep = ~((ref["count"] <= ckthr) |
             (ref["kscore"].astype(np.float32) <= np.float32(mosthr)))
    return ref[keep]


def same(got, ref):
    assert len(got) == len(ref), (len(got), len(ref))
    for f in ("r", "c", "count", "kscore", "seed"):
        bad = np.nonzero(np.any((got[f] != ref[f]).reshape(len(got), -1),
                                axis=1))[0]
        assert bad.size == 0, (f, int(bad[0]), got[bad[0]], ref[bad[0]])


@pytest.fixture(scope="module")
def fam():
    w = synth.make(400, 10, model="lognormal", a=200, b=0.5,
                   clip=(20, 1500), fam_size=8, ident=(0.4, 0.95),
                   frac_hom=0.5, seed=91, name="kmer-families")
    al = Aligner(config_for_cli(-11, -2, 6, 49, 2, "as_called"))
    al.load_seqs(w.codes, w.offsets)
    yield w, al
    al.close()


@pytest.mark.parametrize("sym", ["none", "upper", "even_odd"])
def test_candidates_match_oracle(fam, sym):
    w, al = fam
    res = ascii_of(w.codes)
    print("kmer", sym, flush=True)
    got = al.kmer_candidates(k=6, symmetric=sym, row_base=5, col_base=5)
    ref = orc.kmer_candidates(res, w.offsets, k=6, sym=sym, roff=5, coff=5)
    same(got, ref)
    assert (got["count"] > 1).any()
    # fused k-mer threshold prune (ckthr = 1 as in the C5 stream)
    got = al.kmer_candidates(k=6, symmetric=sym, ckthr=1, mosthr=20.0)
    ref = orc.kmer_candidates(res, w.offsets, k=6, sym=sym)
    same(got, prune(ref, 1, 20.0))


@pytest.mark.parametrize("alph", list(_lib.ALPHABETS))
def test_candidates_alphabets_and_strides(fam, alph):
    w, al = fam
    res = ascii_of(w.codes)
    for k, stride in [(4, 1), (5, 3)]:
        print("kmer", alph, k, stride, flush=True)
        got = al.kmer_candidates(k=k, stride=stride, alphabet=alph,
                                 symmetric="upper", ckthr=1)
        ref = orc.kmer_candidates(res, w.offsets, k=k, stride=stride,
                                  alph=alph, sym="upper")
        same(got, prune(ref, 1, -1.0))


def test_column_blocks_concatenate_to_the_whole(fam):
    w, al = fam
    whole = al.kmer_candidates(k=5, symmetric="upper")
    parts = [al.kmer_candidates(k=5, symmetric="upper", c_begin=a, c_end=b)
             for a, b in [(0, 37), (37, 200), (200, 400)]]
    same(np.concatenate(parts), whole)
    # the count-only call and the copy call agree (kept result reused)
    n = al.kmer_candidates_dev(None, 0, k=5, symmetric="upper")
    assert n == len(whole)


def test_separate_row_and_column_sets():
    w = synth.make(300, 10, model="lognormal", a=150, b=0.5, clip=(10, 900),
                   fam_size=6, ident=(0.5, 0.95), frac_hom=0.5, seed=92,
                   name="kmer-rc")
    L = w.lengths
    seqs = [w.codes[int(w.offsets[i]):int(w.offsets[i + 1])]
            for i in range(len(L))]
    rows, cols = seqs[:180], seqs[120:]

    def pack(ss):
        o = np.zeros(len(ss) + 1, np.uint64)
        o[1:] = np.cumsum([len(s) for s in ss])
        return np.concatenate(ss).astype(np.uint8), o

    rc_, ro = pack(rows)
    cc_, co = pack(cols)
    al = Aligner(config_for_cli(-11, -2, 6, 49, 2, "as_called"))
    al.load_seqs(rc_, ro, side=_lib.PA_SIDE_ROW)
    al.load_seqs(cc_, co, side=_lib.PA_SIDE_COL)
    got = al.kmer_candidates(k=6, symmetric="upper", row_base=0, col_base=120)
    ref = orc.kmer_candidates(ascii_of(rc_), ro, ascii_of(cc_), co, k=6,
                              sym="upper", roff=0, coff=120)
    same(got, ref)
    al.close()


def test_ascii_input_skips_foreign_letters():
    seqs = [b"MKVLAAGWUQMKVLA", b"mkvlaagwoqmkvla", b"MKVLA*AGWMKVLAB",
            b"XXXXXXMKVLAXXXX"]
    res = np.frombuffer(b"".join(seqs), np.uint8)
    offs = np.zeros(len(seqs) + 1, np.uint64)
    offs[1:] = np.cumsum([len(s) for s in seqs])
    al = Aligner(config_for_cli(-11, -2, 6, 49, 2, "as_called"))
    al.load_seqs(res, offs, encoding=_lib.PA_ENC_ASCII)
    got = al.kmer_candidates(k=4)
    ref = orc.kmer_candidates(res, offs, k=4)
    same(got, ref)
    al.close()
