"""The k-mer candidate oracle and the FASTA normalisation oracle, pinned.

k-mer candidates (oracle/kmer_oracle.c) against a brute force that compares
substrings position by position for every (row, column) pair and folds the
products with literal restatements of KmerIntersectLoc's multiply and add
(inc/sr.hpp:54-78), ordering products by the alphabet-mapped k-mer string
(the first residue most significant, which is the k-mer code's order), then
the column offset, then the row offset.  FASTA normalisation
(oracle/fasta_oracle.c) against hand-worked cases of the reference's
constructor (src/FastaData.cpp:13-148) and reader (src/DistFastaData.cpp:
161-263)."""
import random

import numpy as np
import pytest

from oracle import oracle as orc

PROT = "ARNDCQEGHILKMFPSTWYVBZX*J"


def _seqset(strs):
    offs = np.zeros(len(strs) + 1, np.uint64)
    offs[1:] = np.cumsum([len(s) for s in strs])
    return np.frombuffer("".join(strs).encode(), np.uint8), offs


def multiply(p):
    # OUT a (count 1, default second (0, 0)); first = (row off, col off)
    return {"count": 1, "score": p["cost"], "first": (p["i"], p["j"]),
            "second": (0, 0)}


def add(a, b):
    return {"count": (a["count"] + b["count"]) & 0xFFFF,
            "score": max(a["score"], b["score"]), "first": a["first"],
            "second": b["first"]}


def brute(rows, cols, k, stride, alph, sym="none", roff=0, coff=0):
    code, self_, _ = orc.alph_tables(alph)

    def kmers(s):
        out = []
        for p in range(0, len(s) - k + 1, stride):
            chars = [ord(ch) & 0xDF for ch in s[p:p + k]]
            if any(ch >= 128 or code[ch] == 255 for ch in chars):
                continue
            out.append((p, tuple(int(code[ch]) for ch in chars),
                        sum(int(self_[ch]) for ch in chars)))
        return out

    res = []
    for c, cs in enumerate(cols):
        for r, rs in enumerate(rows):
            R, C = r + roff, c + coff
            if sym == "upper" and not C > R:
                continue
            if sym == "even_odd":
                drop = R == C or ((R & 1) ^ (C & 1) if R > C
                                  else not ((R & 1) ^ (C & 1)))
                if drop:
                    continue
            prods = []
            for (i, ki, ci) in kmers(rs):
                for (j, kj, cj) in kmers(cs):
                    if ki == kj:
                        prods.append({"key": (ki, j, i), "i": i, "j": j,
                                      "cost": max(ci, cj)})
            if not prods:
                continue
            prods.sort(key=lambda p: p["key"])
            acc = multiply(prods[0])
            for p in prods[1:]:
                acc = add(acc, multiply(p))
            res.append((r, c, acc["count"], acc["score"]) + acc["first"] +
                       acc["second"])
    return res


def as_tuples(cands):
    return [(int(x["r"]), int(x["c"]), int(x["count"]), int(x["kscore"]),
             *(int(v) for v in x["seed"])) for x in cands]


def rnd(rng, n, alphabet):
    return "".join(rng.choice(alphabet) for _ in range(n))


@pytest.mark.parametrize("alph", list(orc.ALPHABETS))
def test_kmer_candidates_match_brute_force(alph):
    rng = random.Random(hash(alph) % 1000)
    for trial in range(6):
        # small alphabets and repeats give many shared k-mers, duplicates
        # inside one sequence, and foreign or lower-case letters
        letters = ["ACDW", "AAKLMu", "ACDEFGHIKLMNPQRSTVWYBZXJ", "AcdO*"][trial % 4]
        rows = [rnd(rng, rng.randint(0, 14), letters) for _ in range(6)]
        cols = [rnd(rng, rng.randint(0, 14), letters) for _ in range(5)]
        k, s = [(3, 1), (2, 1), (3, 2), (1, 1), (4, 3), (2, 2)][trial]
        rr, ro = _seqset(rows)
        cr, co = _seqset(cols)
        got = as_tuples(orc.kmer_candidates(rr, ro, cr, co, k=k, stride=s,
                                            alph=alph))
        assert got == brute(rows, cols, k, s, alph), (alph, trial)


@pytest.mark.parametrize("sym", ["none", "upper", "even_odd"])
def test_kmer_candidates_symmetric_pruning(sym):
    rng = random.Random(11)
    seqs = [rnd(rng, rng.randint(3, 16), "ACDW") for _ in range(9)]
    rr, ro = _seqset(seqs)
    for roff, coff in [(0, 0), (3, 0), (4, 7)]:
        got = as_tuples(orc.kmer_candidates(rr, ro, k=3, sym=sym, roff=roff,
                                            coff=coff))
        assert got == brute(seqs, seqs, 3, 1, "pdefault", sym, roff, coff)
    if sym == "even_odd":
        # exactly one of (r, c), (c, r) survives for r != c, no diagonal
        full = {(x[0], x[1]) for x in brute(seqs, seqs, 3, 1, "pdefault")}
        kept = {(x[0], x[1]) for x in got if True}
        got0 = {(x[0], x[1]) for x in as_tuples(
            orc.kmer_candidates(rr, ro, k=3, sym="even_odd"))}
        for (r, c) in full:
            if r != c:
                assert ((r, c) in got0) != ((c, r) in got0)
            else:
                assert (r, c) not in got0
        assert kept


def test_kmer_candidates_closed_forms():
    # a sequence against itself with no repeated k-mer: count = #k-mers,
    # score = best self score, first = the smallest code's offset on both
    s = "MKVLAAGWEQ"
    rr, ro = _seqset([s])
    (x,) = orc.kmer_candidates(rr, ro, k=3)
    assert x["count"] == len(s) - 2
    diag = {ch: orc.blosum62()[i * 33] for i, ch in enumerate(PROT)}
    assert x["kscore"] == max(sum(int(diag[ch]) for ch in s[p:p + 3])
                              for p in range(len(s) - 2))
    assert x["seed"][0] == x["seed"][1] and x["seed"][2] == x["seed"][3]
    # count is a ushort: 256 x 256 products of "AAA" wrap to 0
    a = "A" * 258
    rr, ro = _seqset([a])
    (x,) = orc.kmer_candidates(rr, ro, k=3)
    assert x["count"] == (256 * 256) & 0xFFFF == 0
    # one shared k-mer: second is the default (0, 0)
    rr, ro = _seqset(["WWWCA", "KKWWW"])
    got = orc.kmer_candidates(rr, ro, k=3, sym="upper")
    assert as_tuples(got) == [(0, 1, 1, 33, 0, 2, 0, 0)]


# ---- FASTA normalisation --------------------------------------------------

FASTA_CASES = [
    # the file's last byte is dropped (the reader assumes a final newline)
    (b">s1\nACDEFGH\n", 3, ["ACDEFGH"]),
    (b">s1\nACDEFGH", 3, ["ACDEFG"]),
    # newlines inside a sequence go, '*' goes, U/u become T
    (b">s1 x\nACD\nEFG*\nuUK\n", 3, ["ACDEFGTTK"]),
    # '>' inside a header does not start a record
    (b">a>b\nMKV\n>c\nWWW\n", 3, ["MKV", "WWW"]),
    # a short record in the middle goes, and takes the previous sequence's
    # last residue with it (the removal deletes one byte before its '>')
    (b">a\nACDEFG\n>b\nAC\n>c\nMKLUV*W\n", 3, ["ACDEF", "MKLTVW"]),
    # two short records in a row: two residues go
    (b">a\nACDEFG\n>b\nAC\n>c\nK\n>d\nMKLW\n", 3, ["ACDE", "MKLW"]),
    # a short last record: the previous sequence is intact
    (b">a\nACDEFG\n>b\nAC\n", 3, ["ACDEFG"]),
    # a short first record: the others are intact
    (b">a\nAC\n>b\nMKVW\n>c\nWWWW\n", 3, ["MKVW", "WWWW"]),
    # lower case and other bytes are kept as residues
    (b">a\nacdXO\n>b\nMK\r\n", 3, ["acdXO", "MK\r"]),
]


@pytest.mark.parametrize("data,k,want", FASTA_CASES)
def test_fasta_normalisation_cases(data, k, want):
    res, offs, _ = orc.fasta_normalize(data, k)
    got = [bytes(res[int(offs[i]):int(offs[i + 1])]).decode()
           for i in range(len(offs) - 1)]
    assert got == want


def test_fasta_record_indices_and_length_wrap():
    data = b">a\nAC\n>b\n" + b"A" * 65540 + b"\n>c\nMKV\n"
    res, offs, orig = orc.fasta_normalize(data, 3)
    # lengths are ushort: 65540 residues read as 4, minus the residue the
    # removal of record a does not take (a is first)
    assert list(orig) == [1, 2]
    assert int(offs[1] - offs[0]) == 65540 % 65536
    assert bytes(res[int(offs[1]):int(offs[2])]) == b"MKV"
