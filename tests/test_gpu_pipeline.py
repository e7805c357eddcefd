"""The candidate stream end to end on the GPU against the oracle: k-mer
candidates (k = 6, strict upper triangle, ckthr = 1) -> X-drop alignment ->
similarity triples -> Matrix Market file, on a small set of the end-to-end
configuration's generator (log-normal lengths, 5 % of the sequences in
families).  The oracle path runs the same steps on the CPU: kmer_oracle's
overlap matrix, the k-mer threshold prune, or_sxa_batch, and the passing
pairs as (r, c, score_aln)."""
import numpy as np
import pytest

from helpers import oracle_params
from oracle import oracle as orc
from pastis_amd import io, synth
from pastis_amd.aligner import ALPHABET, Aligner, config_for_cli
from pastis_amd.pipeline import CandidateStream

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def e2e():
    w = synth.families(3000, median=280, sigma=0.6, fam_frac=0.2,
                       fam_size=(2, 50), ident=(0.3, 0.9), seed=95,
                       name="e2e-small")
    al = Aligner(config_for_cli(-11, -2, 6, 49, 2, "as_called"))
    al.load_seqs(w.codes, w.offsets)
    res = np.frombuffer(ALPHABET.encode(), np.uint8)[w.codes]
    cand = orc.kmer_candidates(res, w.offsets, k=6, sym="upper")
    cand = cand[cand["count"] > 1]
    ref = orc.sxa_batch(w.codes, w.offsets, cand["r"], cand["c"],
                        cand["seed"], oracle_params("as_called"))
    ok = ref["score_aln"] != 0
    want = sorted(zip(cand["r"][ok].tolist(), cand["c"][ok].tolist(),
                      ref["score_aln"][ok].tolist()))
    yield w, al, cand, want
    al.close()


def _triples(t):
    return sorted(zip(t["row"].tolist(), t["col"].tolist(),
                      t["score_aln"].tolist()))


def test_stream_matches_oracle(e2e, tmp_path):
    w, al, cand, want = e2e
    n = len(w.offsets) - 1
    print("e2e candidates", len(cand), "passing", len(want), flush=True)
    st = CandidateStream(al, k=6, ckthr=1, symmetric="upper", batch=4096,
                         col_block=700)
    tri, stats = st.run(ncols=n, keep_candidates=True)
    got_c = stats["candidate_pairs"]
    key = lambda a: a["c"].astype(np.int64) * n + a["r"]
    assert np.array_equal(np.sort(key(got_c)), np.sort(key(cand)))
    assert stats["candidates"] == len(cand)
    assert _triples(tri) == want
    assert len(want) > 100
    # Matrix Market output of the triples
    p = tmp_path / "sim.mtx"
    io.write_mm(str(p), n, n, tri)
    lines = p.read_text().splitlines()
    assert lines[1] == "%d %d %d" % (n, n, len(want))
    st.close()


def test_stream_filter_extremes(e2e):
    from pastis_amd.prefilter import init_params
    w, al, cand, want = e2e
    n = len(w.offsets) - 1
    al.filter_set_model(init_params(0), np.zeros(16, np.float32),
                        np.ones(16, np.float32))
    st = CandidateStream(al, k=6, ckthr=1, symmetric="upper", batch=5000,
                         tau=float("-inf"))
    tri, stats = st.run(ncols=n)
    assert _triples(tri) == want            # every pair kept
    st.tau = float("inf")
    tri, stats = st.run(ncols=n)
    assert len(tri) == 0 and stats["aligned"] == 0
    st.close()
