"""Shared helpers of the test suite (inputs, oracle runs, comparisons)."""
import numpy as np

from oracle import oracle as orc
from pastis_amd import synth
from pastis_amd._lib import PAIR_DTYPE

STD20 = "ARNDCQEGHILKMFPSTWYV"

# (name, DP open, DP extend, X-drop gap) for the two readings of -g -11 -e -2
GAP_MODELS = {
    "as_called": dict(gap_open=-2, gap_ext=-11, xdrop_gap=-11),
    "as_documented": dict(gap_open=-11, gap_ext=-2, xdrop_gap=-2),
}


def seqset(seqs):
    """list of code arrays -> (flat codes, offsets)."""
    offs = np.zeros(len(seqs) + 1, np.uint64)
    offs[1:] = np.cumsum([len(s) for s in seqs])
    flat = np.concatenate([np.asarray(s, np.uint8) for s in seqs]) \
        if seqs else np.zeros(0, np.uint8)
    return flat, offs


def pairs_of(rc, seeds=None):
    p = np.zeros(len(rc), PAIR_DTYPE)
    for i, (r, c) in enumerate(rc):
        p["r"][i], p["c"][i] = r, c
        p["count"][i] = 2
    if seeds is not None:
        p["seed"] = np.asarray(seeds, np.uint16).reshape(-1, 4)
    return p


def rand_seq(rng, n, alphabet=STD20):
    idx = rng.integers(0, len(alphabet), n)
    return orc.encode("".join(alphabet[i] for i in idx))


def oracle_sw(codes, offs, pairs, go, ge):
    return orc.sw_batch(codes, offs, pairs["r"], pairs["c"], go, ge)


def oracle_params(model, xdrop=49, seed_len=6, seed_cnt=2):
    g = GAP_MODELS[model]
    return orc.params(g["gap_open"], g["gap_ext"], g["xdrop_gap"], xdrop,
                      seed_len, seed_cnt, 0.7, 30)


def small_config1(n_seq=120, n_pairs=1500, seed=11):
    return synth.make(n_seq, n_pairs, model="uniform", a=50, b=500,
                      fam_size=12, ident=(0.4, 0.9), frac_hom=0.5, seed=seed,
                      name="sw-small")


def small_config2(n_seq=150, n_pairs=1500, seed=12):
    return synth.make(n_seq, n_pairs, model="lognormal", a=250, b=0.6,
                      clip=(30, 3000), fam_size=15, ident=(0.3, 0.95),
                      frac_hom=0.3, seed=seed, name="xdrop-small")
