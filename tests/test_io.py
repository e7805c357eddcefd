"""FASTA normalisation and Matrix Market output of the library (host code,
no GPU needed) against the oracle and hand-worked files.

pa_fasta_parse must return exactly the oracle's sequences (the oracle
restates the reference's reader and FastaData constructor, and is pinned by
tests/test_kmer_oracle.py's hand-worked cases) on those cases and on random
FASTA-like files with '*', U, lower case, CR bytes, '>' inside headers,
blank lines and short records in every position."""
import random

import numpy as np
import pytest

from oracle import oracle as orc
from pastis_amd import io
from test_kmer_oracle import FASTA_CASES


def seqs_of(res, offs):
    return [bytes(res[int(offs[i]):int(offs[i + 1])])
            for i in range(len(offs) - 1)]


@pytest.mark.parametrize("data,k,want", FASTA_CASES)
def test_fasta_parse_hand_cases(data, k, want):
    res, offs, recs, wrapped = io.fasta_parse(data, k)
    assert [s.decode() for s in seqs_of(res, offs)] == want
    ores, ooffs, orecs = orc.fasta_normalize(data, k)
    assert list(recs) == list(orecs) and wrapped == 0


def random_fasta(rng):
    out = []
    if rng.random() < 0.1:
        out.append("junk line\n")
    for r in range(rng.randint(0, 12)):
        name = "".join(rng.choice("abc >*U1 ") for _ in range(rng.randint(0, 8)))
        out.append(">" + name + "\n")
        n = rng.choice([0, 1, 2, 3, 5, 8, 20, 60])
        body = "".join(rng.choice("ACDEFGHIKLMNPQRSTVWYuU*xX\r")
                       for _ in range(n))
        # wrap into lines, sometimes with blank lines
        lines, p = [], 0
        while p < len(body) or not lines:
            w = rng.randint(1, 12)
            lines.append(body[p:p + w])
            p += w
            if rng.random() < 0.1:
                lines.append("")
        out.append("\n".join(lines) + "\n")
    text = "".join(out)
    if rng.random() < 0.2 and text:
        text = text[:-1]            # no final newline
    return text.encode()


def test_fasta_parse_matches_oracle_on_random_files():
    rng = random.Random(123)
    for trial in range(1500):
        data = random_fasta(rng)
        k = rng.choice([1, 2, 3, 6])
        res, offs, recs, _ = io.fasta_parse(data, k)
        ores, ooffs, orecs = orc.fasta_normalize(data, k)
        assert seqs_of(res, offs) == seqs_of(ores, ooffs), (trial, data, k)
        assert list(recs) == list(orecs), (trial, data)


def test_fasta_lengths_wrap_like_ushort():
    data = b">a\n" + b"W" * 70000 + b"\n>b\nMKV\n"
    res, offs, recs, wrapped = io.fasta_parse(data, 3)
    assert wrapped == 1
    assert int(offs[1] - offs[0]) == 70000 % 65536
    ores, ooffs, _ = orc.fasta_normalize(data, 3)
    assert seqs_of(res, offs) == seqs_of(ores, ooffs)


def test_matrix_market_output(tmp_path):
    from pastis_amd._lib import TRIPLE_DTYPE
    t = np.zeros(4, TRIPLE_DTYPE)
    t["row"] = [3, 0, 2, 1]
    t["col"] = [1, 4, 1, 0]
    t["score_aln"] = [0.5, 0.3125, 0.987654321, 1.0]
    p = tmp_path / "sim.mtx"
    io.write_mm(str(p), 5, 6, t)
    lines = p.read_text().splitlines()
    assert lines[0] == "%%MatrixMarket matrix coordinate real general"
    assert lines[1] == "5 6 4"
    # column-major, 1-based, 6 significant digits
    assert lines[2:] == ["2 1 1", "3 2 0.987654", "4 2 0.5", "1 5 0.3125"]
    bad = t.copy()
    bad["row"][0] = 5
    from pastis_amd._lib import PaError
    with pytest.raises(PaError):
        io.write_mm(str(p), 5, 6, bad)
