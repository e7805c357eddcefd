"""Synthetic inputs of the benchmark configurations (libpastis_gen.so).

Deterministic (std::mt19937_64 with the seeds below).  The configurations
mirror the five workloads the aligner is measured on; `scale` shrinks
sequence and pair counts for tests.
"""
import ctypes as C
import os

import numpy as np

from ._lib import GEN_PATH, PAIR_DTYPE

_gen = None


def gen():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise ImportError("libpastis_gen.so is not built (run build())")
        G = C.CDLL(GEN_PATH)
        d, u64, i = C.c_double, C.c_uint64, C.c_int
        G.sg_new_set.argtypes = [u64, i, d, d, i, i, d, d, d, i, d, d, d, u64]
        G.sg_new_set.restype = C.c_void_p
        G.sg_make_pairs.argtypes = [C.c_void_p, u64, d, d, i, i, u64]
        G.sg_new_families.argtypes = [u64, d, d, i, i, d, i, i, d, d, d, u64]
        G.sg_new_families.restype = C.c_void_p
        for f in ("sg_num_seqs", "sg_total_residues", "sg_num_pairs"):
            getattr(G, f).argtypes = [C.c_void_p]
            getattr(G, f).restype = u64
        G.sg_mean_identity.argtypes = [C.c_void_p]
        G.sg_mean_identity.restype = d
        G.sg_copy_seqs.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        G.sg_copy_pairs.argtypes = [C.c_void_p, C.c_void_p]
        G.sg_copy_family.argtypes = [C.c_void_p, C.c_void_p]
        G.sg_free.argtypes = [C.c_void_p]
        _gen = G
    return _gen


class Workload:
    def __init__(self, name, codes, offsets, pairs, family, info):
        self.name = name
        self.codes = codes
        self.offsets = offsets
        self.pairs = pairs
        self.family = family
        self.info = info

    @property
    def lengths(self):
        return np.diff(self.offsets).astype(np.int64)

    def cells(self):
        L = self.lengths
        return int(np.sum(L[self.pairs["r"]] * L[self.pairs["c"]]))


def make(nseq, npairs, *, model="uniform", a=50, b=500, clip=(30, 3000),
         tail=(0.0, 3000, 35000), fam_size=20, ident=(0.4, 0.9),
         indel_p=0.05, frac_hom=0.5, zipf=0.0, k=6, seed=101, threads=8,
         name="custom"):
    G = gen()
    h = G.sg_new_set(nseq, 0 if model == "uniform" else 1, float(a), float(b),
                     clip[0], clip[1], float(tail[0]), float(tail[1]),
                     float(tail[2]), fam_size, ident[0], ident[1], indel_p,
                     seed)
    try:
        G.sg_make_pairs(h, npairs, frac_hom, zipf, k, threads, seed + 7)
        ns, nr = G.sg_num_seqs(h), G.sg_total_residues(h)
        codes = np.zeros(nr, np.uint8)
        offs = np.zeros(ns + 1, np.uint64)
        G.sg_copy_seqs(h, codes.ctypes.data, offs.ctypes.data)
        pairs = np.zeros(npairs, PAIR_DTYPE)
        G.sg_copy_pairs(h, pairs.ctypes.data)
        fam = np.zeros(ns, np.int32)
        G.sg_copy_family(h, fam.ctypes.data)
        info = dict(name=name, nseq=int(ns), residues=int(nr),
                    npairs=int(npairs), seed=seed,
                    mean_child_identity=G.sg_mean_identity(h))
    finally:
        G.sg_free(h)
    return Workload(name, codes, offs, pairs, fam, info)


def config1(scale=1.0, seed=101, threads=8):
    """Plain SW: 2,000 proteins U[50,500]; 1e5 pairs, half homologs 40-90%."""
    return make(max(int(2000 * scale), 40), max(int(1e5 * scale), 10),
                model="uniform", a=50, b=500, fam_size=20, ident=(0.4, 0.9),
                frac_hom=0.5, seed=seed, threads=threads, name="sw-uniform")


def config2(scale=1.0, seed=102, threads=8):
    """X-drop: 20,000 proteins log-normal(250, 0.6) in [30, 3000]; 5e6 pairs,
    30% homologs at 30-95% identity, up to 2 shared 6-mer seeds."""
    return make(max(int(20000 * scale), 100), max(int(5e6 * scale), 10),
                model="lognormal", a=250, b=0.6, clip=(30, 3000),
                fam_size=50, ident=(0.3, 0.95), frac_hom=0.3, seed=seed,
                threads=threads, name="xdrop-lognormal")


def config3(scale=1.0, seed=103, threads=8):
    """Uneven SW: 200,000 proteins log-normal(300, 0.6) plus a 1% tail
    U[3000, 35000]; 5e7 pairs with Zipf(1.1) endpoints."""
    return make(max(int(200000 * scale), 100), max(int(5e7 * scale), 10),
                model="lognormal", a=300, b=0.6, clip=(30, 3000),
                tail=(0.01, 3000, 35000), fam_size=1, frac_hom=0.0, zipf=1.1,
                seed=seed, threads=threads, name="sw-heavy-tail")


def families(nseq, *, median=280, sigma=0.6, clip=(30, 3000), fam_frac=0.05,
             fam_size=(2, 50), ident=(0.3, 0.9), indel_p=0.05, seed=105,
             name="families"):
    """Sequences only (no pairs): log-normal lengths, fam_frac of them in
    families of U[fam_size] members at identity U[ident]."""
    G = gen()
    h = G.sg_new_families(nseq, float(median), float(sigma), clip[0], clip[1],
                          float(fam_frac), fam_size[0], fam_size[1], ident[0],
                          ident[1], indel_p, seed)
    try:
        ns, nr = G.sg_num_seqs(h), G.sg_total_residues(h)
        codes = np.zeros(nr, np.uint8)
        offs = np.zeros(ns + 1, np.uint64)
        G.sg_copy_seqs(h, codes.ctypes.data, offs.ctypes.data)
        fam = np.zeros(ns, np.int32)
        G.sg_copy_family(h, fam.ctypes.data)
        info = dict(name=name, nseq=int(ns), residues=int(nr), seed=seed,
                    mean_child_identity=G.sg_mean_identity(h))
    finally:
        G.sg_free(h)
    return Workload(name, codes, offs, np.zeros(0, PAIR_DTYPE), fam, info)


def config5(scale=1.0, seed=105):
    """End-to-end candidate stream: 1M proteins, log-normal lengths (median
    280, sigma 0.6 assumed, clipped to [30, 3000]), 5 % of them in families of
    2-50 members at 30-90 % identity.  The candidates come from the k-mer
    generator (k = 6), not from here."""
    return families(max(int(1_000_000 * scale), 200), median=280, sigma=0.6,
                    fam_frac=0.05, fam_size=(2, 50), ident=(0.3, 0.9),
                    seed=seed, name="e2e-families")
